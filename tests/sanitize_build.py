"""Sanitizer builds for tests/test_sanitize.py -- TEST INFRASTRUCTURE.

    python tests/sanitize_build.py

1. tests/_asan/oracle_asan: the CPU oracle (oracle/oracle.c, oracle/ldpc.c)
   and its driver tests/cpp/oracle_asan.cpp, built with gcc / g++ and
   AddressSanitizer + UndefinedBehaviorSanitizer.
2. tests/_asan/libmegadetect_asan.so: every source of the HIP library with
   the sanitizers on its host code only (device code is compiled as usual;
   GPU sanitizers are not used), and tests/cpp/abi_asan.cpp linked against
   it, exercising the C ABI's host-side code with no GPU.

Each output is rebuilt only when a source is newer than it.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_asan")
CSRC = os.path.join(ROOT, "megastation_amd", "csrc")
INC = os.path.join(ROOT, "include")
ORACLE_BIN = os.path.join(OUT, "oracle_asan")
ABI_LIB = os.path.join(OUT, "libmegadetect_asan.so")
ABI_BIN = os.path.join(OUT, "abi_asan")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
HIPCC = "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++"


def _stale(out: str, srcs: list[str]) -> bool:
    return not os.path.exists(out) or any(os.path.getmtime(s) > os.path.getmtime(out)
                                          for s in srcs)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(" ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_oracle() -> str:
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("oracle.c", "ldpc.c")]
    drv = os.path.join(HERE, "cpp", "oracle_asan.cpp")
    deps = srcs + [drv, os.path.join(ROOT, "oracle", "oracle.h"),
                   os.path.join(CSRC, "ldpc_tables.h")]
    if not _stale(ORACLE_BIN, deps):
        return ORACLE_BIN
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for s in srcs:
        o = os.path.join(OUT, os.path.basename(s) + ".o")
        _run(["gcc", "-O1", "-g", "-std=c11", "-fopenmp", "-ffp-contract=off", *SAN, "-c", s,
              "-o", o])
        objs.append(o)
    _run(["g++", "-O1", "-g", "-std=c++17", "-fopenmp", *SAN, drv, *objs, "-lm",
          "-o", ORACLE_BIN])
    return ORACLE_BIN


def _lib_sources() -> list[str]:
    sys.path.insert(0, ROOT)
    from megastation_amd import build as mb
    return [os.path.join(CSRC, s) for s in mb.SOURCES]


def build_abi() -> str:
    srcs = _lib_sources()
    drv = os.path.join(HERE, "cpp", "abi_asan.cpp")
    deps = srcs + [drv] + glob.glob(os.path.join(CSRC, "*.h")) + \
        glob.glob(os.path.join(INC, "*.h"))
    if not _stale(ABI_BIN, deps):
        return ABI_BIN
    os.makedirs(OUT, exist_ok=True)
    # sanitizers on the host side only: each option after -Xarch_host
    host_san = []
    for f in SAN:
        host_san += ["-Xarch_host", f]
    common = ["-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              f"-I{INC}", f"-I{CSRC}", *host_san]

    def comp(s):
        o = os.path.join(OUT, os.path.basename(s) + ".o")
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        _run([HIPCC, *common, *lang, "-c", s, "-o", o])
        return o
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(comp, srcs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", ABI_LIB, *objs])
    _run([CLANGXX, "-O1", "-g", "-std=c++17", *SAN, f"-I{INC}", drv, "-L" + OUT,
          "-lmegadetect_asan", "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath," + OUT, "-Wl,-rpath,/opt/rocm/lib", "-o", ABI_BIN])
    return ABI_BIN


if __name__ == "__main__":
    print(build_oracle())
    print(build_abi())
