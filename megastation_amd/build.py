"""Build libmegadetect.so (HIP, gfx950) in-tree with hipcc.

    python -m megastation_amd.build

Every source under csrc/ is compiled for gfx950 only, with packed FP32 ops
disabled, and linked into megastation_amd/lib/libmegadetect.so. The build always
recompiles (it is the "does it build" check).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmegadetect.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["capi.cpp", "equalize.hip", "zf.hip", "detect.hip", "zf_mixed.hip",
           "equalize_mixed.hip", "batch.hip", "batch.cpp", "mlp.hip", "learned.hip", "pipeline.cpp", "precode.hip", "fft.hip", "tuning.cpp", "equalize_wide.hip", "ldpc.hip", "zf_wide.hip", "workspace.cpp", "uplink.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function",
          f"-I{os.path.join(os.path.dirname(HERE), 'include')}", f"-I{CSRC}"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    noise = "is not a recognized feature for this target"
    msgs = [l for l in (r.stderr or "").splitlines() if noise not in l and l.strip()]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return obj


CPP_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_dropin.cpp")
CPP_TEST_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "test_dropin")


def build_cpp_tests(verbose: bool = True) -> str:
    """The C++ drop-in test program (include/mega/ classes over the C ABI)."""
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           f"-I{os.path.join(os.path.dirname(HERE), 'include')}", CPP_TEST_SRC,
           f"-L{LIBDIR}", "-lmegadetect", "-Wl,-rpath,$ORIGIN/../../megastation_amd/lib",
           "-o", CPP_TEST_BIN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for test_dropin.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(CPP_TEST_BIN + ".tmp", CPP_TEST_BIN)
    if verbose:
        print("built", CPP_TEST_BIN)
    return CPP_TEST_BIN


def build(verbose: bool = True) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    shutil.rmtree(OBJDIR, ignore_errors=True)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    build_cpp_tests()
