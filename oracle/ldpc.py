"""LDPC oracle -- TEST INFRASTRUCTURE ONLY.

Bindings of oracle/ldpc.c (layered normalized min-sum decoding, the
operation behind src/operators/ldpc/decoder.h:73-86), plus numpy helpers the
tests use: the base-graph tables read from the library's generated header
(3GPP TS 38.212 Tables 5.3.2-2/-3), the lifting-size set of Zc (Table
5.3.2-1), parity checks of a full codeword, and a systematic encoder
(38.212 Section 5.3.2: information bits, then the four core parity blocks
solved through the double-diagonal structure, then the extension parity
blocks one row at a time). Only tests/, smoke() and bench.py's CPU-baseline
leg use this module.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from . import _p, lib

_HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(_HERE), "megastation_amd", "csrc", "ldpc_tables.h")
KB = {1: 22, 2: 10}
# Table 5.3.2-1: lifting sizes Zc = a * 2^j of set iLS
SET_A = [2, 3, 5, 7, 9, 11, 13, 15]
_tables = {}


def set_index(zc: int) -> int:
    for i, a in enumerate(SET_A):
        z = a
        while z <= 384:
            if z == zc:
                return i
            z *= 2
    raise ValueError(f"{zc} is not a 5G NR lifting size")


def tables(bg: int):
    """(row_start int16[rows+1], col int8[edges], shift int16[8][edges])."""
    if bg not in _tables:
        txt = open(HEADER).read()

        def arr(name, dt):
            m = re.search(name + r"\[[^\]]*\](?:\[[^\]]*\])?\s*=\s*\{(.*?)\};", txt, re.S)
            return np.array([int(v) for v in re.findall(r"-?\d+", m.group(1))], dtype=dt)
        rs = arr(f"kBg{bg}RowStart", np.int16)
        col = arr(f"kBg{bg}Col", np.int8)
        sh = arr(f"kBg{bg}Shift", np.int16).reshape(8, -1)
        _tables[bg] = (rs, col, sh)
    return _tables[bg]


def _circ(x: np.ndarray, p: int) -> np.ndarray:
    """I(p) x: row z of the shifted identity picks x[(z + p) mod Zc]."""
    return np.roll(x, -p)


def parity_syndrome(cw: np.ndarray, bg: int, zc: int, mb: int) -> int:
    """Number of unsatisfied checks of the first mb block rows for a full
    codeword cw ((Kb + mb) * Zc bits, punctured bits included)."""
    rs, col, sh = tables(bg)
    s = sh[set_index(zc)]
    bad = 0
    for r in range(mb):
        acc = np.zeros(zc, np.uint8)
        for e in range(rs[r], rs[r + 1]):
            c = int(col[e])
            acc ^= _circ(cw[c * zc:(c + 1) * zc], int(s[e]) % zc)
        bad += int(acc.sum())
    return bad


def encode(info: np.ndarray, bg: int, zc: int, mb: int) -> np.ndarray:
    """Full codeword ((Kb + mb) * Zc bits) of Kb * Zc information bits."""
    kb = KB[bg]
    rs, col, sh = tables(bg)
    s = sh[set_index(zc)]
    x = np.zeros((kb + mb, zc), np.uint8)
    x[:kb] = np.asarray(info, np.uint8).reshape(kb, zc)

    def row_sum(r, upto):  # sum of the row's blocks in columns < upto
        acc = np.zeros(zc, np.uint8)
        for e in range(rs[r], rs[r + 1]):
            if int(col[e]) < upto:
                acc ^= _circ(x[int(col[e])], int(s[e]) % zc)
        return acc
    lam = [row_sum(r, kb) for r in range(4)]
    # column kb has three blocks in rows 0-3; two shifts are equal and cancel
    # in the sum of the four rows (the double diagonal of columns kb+1..kb+3
    # cancels too), leaving I(k) p0 = sum of lambda
    pk = [int(s[e]) % zc for r in range(4) for e in range(rs[r], rs[r + 1])
          if int(col[e]) == kb]
    odd = [p for p in pk if pk.count(p) % 2 == 1]
    assert len(pk) == 3 and len(odd) == 1, pk
    tot = lam[0] ^ lam[1] ^ lam[2] ^ lam[3]
    x[kb] = np.roll(tot, odd[0])            # I(k)^-1
    for r, newc in ((0, kb + 1), (1, kb + 2), (2, kb + 3)):
        # row r: every block left of column newc is known, and newc enters
        # with the identity (the double diagonal), so it is their sum
        assert all(int(s[e]) % zc == 0 for e in range(rs[r], rs[r + 1])
                   if int(col[e]) == newc)
        x[newc] = row_sum(r, newc)
    for r in range(4, mb):                   # extension: identity in column kb + r
        x[kb + r] = row_sum(r, kb + 4)
    cw = x.reshape(-1)
    assert parity_syndrome(cw, bg, zc, mb) == 0
    return cw


def decode(llr: np.ndarray, bg: int, zc: int, mb: int, max_iter: int, alpha: float,
           want_app: bool = False):
    """Hard decisions of the Kb * Zc information bits, bit-packed (uint32,
    bit k at word k//32, bit k%32) for llr float32 [n][(Kb + mb) * Zc]; with
    want_app also the final a-posteriori LLRs."""
    kb = KB[bg]
    rs, col, sh = tables(bg)
    s = np.ascontiguousarray(sh[set_index(zc)])
    llr = np.ascontiguousarray(llr, np.float32)
    if llr.ndim == 1:
        llr = llr[None]
    n, nv = llr.shape
    assert nv >= (kb + mb) * zc
    words = (kb * zc + 31) // 32
    out = np.zeros((n, words), np.uint32)
    app = np.zeros((n, (kb + mb) * zc), np.float32) if want_app else None
    L = lib()
    f = L.orc_ldpc_decode
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_float,
                                                                 ctypes.c_void_p, ctypes.c_uint64,
                                                                 ctypes.c_int, ctypes.c_void_p,
                                                                 ctypes.c_uint64, ctypes.c_void_p]
    f.restype = ctypes.c_int
    rc = f(kb, mb, zc, _p(rs), _p(col), _p(s), max_iter, alpha, _p(llr), nv, n, _p(out), words,
           _p(app) if want_app else None)
    assert rc == 0, rc
    return (out, app) if want_app else out


def unpack_bits(words: np.ndarray, nbits: int) -> np.ndarray:
    """uint32 words (bit k at word k//32, bit k%32) -> uint8 bits."""
    b = np.unpackbits(np.ascontiguousarray(words).view(np.uint8).reshape(words.shape[0], -1),
                      axis=1, bitorder="little")
    return b[:, :nbits]
