"""CPU oracle for the uplink detection path — TEST INFRASTRUCTURE ONLY.

Python bindings (ctypes + numpy) over ``liboracle.so``, the plain-C
restatement in ``oracle.c`` of MegaStation's ZF beamforming
(``src/operators/beamform/beamform.cu:74-122``), fused equalize + max-log demod
(``src/operators/sgemv/equalize_kernel.cuh:415-480``,
``src/operators/modulation/demodulation.h:19-67``) and precoder normalisation
(``src/operators/beamform/beam_utils.cuh:40-80``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package: it is the checker, never the product path.
Every function's parity pin is listed in the header of ``oracle.c``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with gcc (FP contraction off, OpenMP on)."""
    cmd = [
        "gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp",
        "-ffp-contract=off", "-fno-fast-math", "-Wall",
        "-o", LIB_PATH + ".tmp", os.path.join(_HERE, "oracle.c"),
        os.path.join(_HERE, "ldpc.c"), "-lm",
    ]
    subprocess.run(cmd, check=True)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        u8, u16, u32, u64, i32 = (ctypes.c_uint8, ctypes.c_uint16,
                                  ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int)
        P = ctypes.c_void_p
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f32_to_f16.restype = u16
        L.orc_f16_to_f32.argtypes = [u16]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f16_sub.argtypes = [u16, u16]
        L.orc_f16_sub.restype = u16
        L.orc_demod.argtypes = [ctypes.c_float, ctypes.c_float, i32, P]
        L.orc_demod.restype = i32
        eq = [u64, u64, u64, P, P, P, u64, u64, i32, u64, u64, u64, u64, P]
        L.orc_equalize_f32.argtypes = eq
        L.orc_equalize_f32.restype = i32
        L.orc_equalize_f64.argtypes = eq
        L.orc_equalize_f64.restype = i32
        L.orc_zf_f64.argtypes = [u32, u32, u32, P, P, P, P]
        L.orc_zf_f64.restype = i32
        L.orc_zf_f32.argtypes = [u32, u32, u32, P, P, P]
        L.orc_zf_f32.restype = i32
        L.orc_scale_absmax.argtypes = [u32, u32, u32, P, P]
        L.orc_scale_absmax.restype = None
        L.orc_precode.argtypes = [u32, u32, u32, u32, P, P, P, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_precode.restype = None
        L.orc_detect_cpu.argtypes = [u32, u32, u32, u32, u32, u32, i32, P, P, P,
                                     u64, u64, u64, u64, u64, i32]
        L.orc_detect_cpu.restype = i32
        del u8
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c64(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.complex64), requirements="C")


def f32_to_f16_bits(x: float) -> int:
    return int(lib().orc_f32_to_f16(float(x)))


def f16_bits_to_f32(h: int) -> float:
    return float(lib().orc_f16_to_f32(int(h)))


def f16_sub_bits(a: int, b: int) -> int:
    return int(lib().orc_f16_sub(int(a), int(b)))


def demod(xhat: np.ndarray, Q: int) -> np.ndarray:
    """Max-log LLRs (binary16 bit patterns, uint16) of each complex value:
    returns shape xhat.shape + (Q,)."""
    x = _c64(xhat)
    flat = x.reshape(-1)
    out = np.zeros((flat.size, Q), dtype=np.uint16)
    L = lib()
    buf = np.zeros(8, dtype=np.uint16)
    for i, c in enumerate(flat):
        if L.orc_demod(float(c.real), float(c.imag), Q, _p(buf)) != 0:
            raise ValueError(f"unsupported modulation order {Q}")
        out[i] = buf[:Q]
    return out.reshape(x.shape + (Q,))


def equalize(csi: np.ndarray, y: np.ndarray, U: int, B: int, N: int, sg: int,
             Q: int, ue_stride: int, cb_size: int, cb_stride: int, zeros: int,
             valid_bits: int, precision: str = "f32", llr_init=None):
    """One symbol. csi: complex64 [G][B][U] flat; y: complex64 [B][N].
    Returns (llr uint16 [U][ue_stride], xhat [U][N] complex64 / complex128)."""
    csi = _c64(csi)
    y = _c64(y)
    llr = (np.zeros((U, ue_stride), dtype=np.uint16) if llr_init is None
           else np.array(llr_init, dtype=np.uint16).reshape(U, ue_stride).copy())
    if precision == "f32":
        x = np.zeros((U, N), dtype=np.complex64)
        rc = lib().orc_equalize_f32(U, B, N, _p(csi), _p(y), _p(llr), ue_stride,
                                    sg, Q, cb_size, cb_stride, zeros, valid_bits,
                                    _p(x))
    else:
        x = np.zeros((U, N), dtype=np.complex128)
        rc = lib().orc_equalize_f64(U, B, N, _p(csi), _p(y), _p(llr), ue_stride,
                                    sg, Q, cb_size, cb_stride, zeros, valid_bits,
                                    _p(x))
    if rc != 0:
        raise ValueError("bad equalize arguments")
    return llr, x


def zf(csi: np.ndarray, U: int, B: int, batch: int, precision: str = "f64"):
    """W_g = (A_g A_g^H)^-1 A_g per group. Returns (W complex64 flat,
    W64 complex128 flat or None, info int32[batch])."""
    a = _c64(csi).reshape(-1)
    out = np.zeros_like(a)
    info = np.zeros(batch, dtype=np.int32)
    if precision == "f64":
        out64 = np.zeros(a.size, dtype=np.complex128)
        lib().orc_zf_f64(U, B, batch, _p(a), _p(out), _p(out64), _p(info))
        return out, out64, info
    lib().orc_zf_f32(U, B, batch, _p(a), _p(out), _p(info))
    return out, None, info


def scale_absmax(w: np.ndarray, U: int, B: int, batch: int) -> np.ndarray:
    a = _c64(w).reshape(-1)
    out = np.zeros_like(a)
    lib().orc_scale_absmax(U, B, batch, _p(a), _p(out))
    return out


def precode(prec: np.ndarray, ue: np.ndarray, B: int, U: int, N: int, sg: int,
            bs_stride: int, bs_start: int, out: np.ndarray | None = None) -> np.ndarray:
    """Downlink precode of one symbol (orc_precode): prec [G][U][B], ue [U][N],
    returns out [B][bs_stride] complex64 (a copy of `out` if given, else
    zeros, with the written entries replaced)."""
    p = _c64(prec).reshape(-1)
    x = _c64(ue).reshape(-1)
    o = (np.zeros(B * bs_stride, np.complex64) if out is None
         else _c64(out).reshape(-1).copy())
    lib().orc_precode(B, U, N, sg, _p(p), _p(x), _p(o), bs_stride, bs_start)
    return o.reshape(B, bs_stride)


def detect_cpu(csi: np.ndarray, y: np.ndarray, slots: int, S: int, U: int,
               B: int, N: int, sg: int, Q: int, ue_stride: int, cb_size: int,
               cb_stride: int, zeros: int, valid_bits: int, n_threads: int = 0,
               llr: np.ndarray | None = None):
    """CPU baseline over `slots` slots. Returns (llr uint16
    [slots][S][U][ue_stride], threads used)."""
    csi = _c64(csi)
    y = _c64(y)
    if llr is None:
        llr = np.zeros((slots, S, U, ue_stride), dtype=np.uint16)
    nt = lib().orc_detect_cpu(slots, S, U, B, N, sg, Q, _p(csi), _p(y), _p(llr),
                              ue_stride, cb_size, cb_stride, zeros, valid_bits,
                              n_threads)
    if nt < 0:
        raise ValueError("bad detect arguments")
    return llr, nt
