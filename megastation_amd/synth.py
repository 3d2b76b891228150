"""Seeded synthetic uplink slots for the detection path.

Constellations, bit order and the CSI convention follow the reference:

* Gray QAM tables restated from ``src/operators/modulation/modulation.cu:22-72``
  (same values, same index bit split); the table index is built from the bit
  stream MSB-first (``modulation_kernel.cuh:54-58``: the first bit of a symbol
  becomes the index MSB), so LLR ``j`` of the demod belongs to stream bit ``j``.
* The reference 256-QAM table is the bitwise complement of what
  ``demod256QAM`` decodes (``demodulation.h:54-67``); the generator therefore
  indexes that table with the complemented index, which is the mapping the
  demod (and 3GPP TS 38.211 with LLR > 0 <=> bit 0) expects.
* CSI buffer ``A_g = H_g^H`` stored ``[G][B][U]`` (element (u,b,g) at
  ``u + U*b + U*B*g``), as the pilot storer writes it
  (``src/operators/fft/fft_callback.cuh:87-105``).
* ``y = H x + n`` per subcarrier, stored ``[S][B][N]`` (subcarrier fastest,
  ``src/baseband/buffers.cc:87-88``); unit-energy symbols, noise
  ``CN(0, 10^(-SNR/10))``.

Random numbers come from numpy's counter-based Philox generator, seeded.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def qam_table(Q: int) -> np.ndarray:
    """complex64[2^Q] constellation indexed by the Q-bit symbol index."""
    M = 1 << Q
    t = np.zeros(M, dtype=np.complex64)
    if Q == 2:
        s = 1 / math.sqrt(2)
        lv = [-s, s]
        for i in range(4):
            t[i] = complex(lv[i // 2], lv[i % 2])
    elif Q == 4:
        s = 1 / math.sqrt(10)
        lv = [1 * s, 3 * s, -1 * s, -3 * s]
        for i in range(16):
            im_i = (((i >> 2) & 1) << 1) + (i & 1)
            re_i = (((i >> 3) & 1) << 1) + ((i >> 1) & 1)
            t[i] = complex(lv[re_i], lv[im_i])
    elif Q == 6:
        s = 1 / math.sqrt(42)
        lv = [3 * s, 1 * s, 5 * s, 7 * s, -3 * s, -1 * s, -5 * s, -7 * s]
        for i in range(64):
            im_i = (((i >> 4) & 1) << 2) + (((i >> 2) & 1) << 1) + (i & 1)
            re_i = (((i >> 5) & 1) << 2) + (((i >> 3) & 1) << 1) + ((i >> 1) & 1)
            t[i] = complex(lv[re_i], lv[im_i])
    elif Q == 8:
        s = 1 / math.sqrt(170)
        lv = [-15 * s, -13 * s, -9 * s, -11 * s, -1 * s, -3 * s, -7 * s, -5 * s,
              15 * s, 13 * s, 9 * s, 11 * s, 1 * s, 3 * s, 7 * s, 5 * s]
        for i in range(256):
            j = (~i) & 0xFF  # complement: see module docstring
            im_i = ((j & 0x40) >> 3) + ((j & 0x10) >> 2) + ((j & 0x4) >> 1) + (j & 1)
            re_i = ((j & 0x80) >> 4) + ((j & 0x20) >> 3) + ((j & 0x8) >> 2) + ((j & 0x2) >> 1)
            t[i] = complex(lv[re_i], lv[im_i])
    else:
        raise ValueError(f"unsupported modulation order {Q}")
    return t


def reference_table(Q: int) -> np.ndarray:
    """The reference's own table for Q (for Q=8 *without* the complement),
    used only to document defect D3 in the tests."""
    if Q != 8:
        return qam_table(Q)
    s = 1 / math.sqrt(170)
    lv = [-15 * s, -13 * s, -9 * s, -11 * s, -1 * s, -3 * s, -7 * s, -5 * s,
          15 * s, 13 * s, 9 * s, 11 * s, 1 * s, 3 * s, 7 * s, 5 * s]
    t = np.zeros(256, dtype=np.complex64)
    for i in range(256):
        im_i = ((i & 0x40) >> 3) + ((i & 0x10) >> 2) + ((i & 0x4) >> 1) + (i & 1)
        re_i = ((i & 0x80) >> 4) + ((i & 0x20) >> 3) + ((i & 0x8) >> 2) + ((i & 0x2) >> 1)
        t[i] = complex(lv[re_i], lv[im_i])
    return t


def bits_to_index(bits: np.ndarray) -> np.ndarray:
    """bits[..., Q] (first stream bit first) -> symbol index, MSB-first."""
    Q = bits.shape[-1]
    w = (1 << np.arange(Q - 1, -1, -1)).astype(np.int64)
    return (bits.astype(np.int64) * w).sum(-1)


def modulate(bits: np.ndarray) -> np.ndarray:
    return qam_table(bits.shape[-1])[bits_to_index(bits)]


# --------------------------------------------------------------------------
# channels
# --------------------------------------------------------------------------

def _cn(rng, shape, var=1.0):
    s = math.sqrt(var / 2)
    return (rng.standard_normal(shape, dtype=np.float32) * s
            + 1j * rng.standard_normal(shape, dtype=np.float32) * s).astype(np.complex64)


def channel_iid(rng, G: int, B: int, U: int) -> np.ndarray:
    """H[G][B][U] i.i.d. CN(0,1)."""
    return _cn(rng, (G, B, U))


def exp_corr(n: int, rho: float) -> np.ndarray:
    i = np.arange(n)
    return rho ** np.abs(i[:, None] - i[None, :])


def _sqrtm_psd(R: np.ndarray) -> np.ndarray:
    w, V = np.linalg.eigh(R)
    return (V * np.sqrt(np.maximum(w, 0))) @ V.conj().T


def channel_kronecker(rng, G: int, B: int, U: int, rho_rx: float,
                      rho_tx: float) -> np.ndarray:
    """H = R_rx^{1/2} H_w R_tx^{1/2}, [R]_ij = rho^|i-j| (per group)."""
    Rr = _sqrtm_psd(exp_corr(B, rho_rx)).astype(np.complex64)
    Rt = _sqrtm_psd(exp_corr(U, rho_tx)).astype(np.complex64)
    Hw = _cn(rng, (G, B, U))
    return (Rr[None] @ Hw @ Rt[None]).astype(np.complex64)


class TdlJakes:
    """Time-varying tapped-delay-line channel for sequences of slots.

    n_taps taps at spacing tap_ns with an exponential power-delay profile
    whose discrete RMS delay spread is rms_ns (the decay constant is solved
    for that), powers summing to 1; every (antenna, user, tap) path fades
    with Jakes' Doppler spectrum at fd_hz, built as a sum of n_sin sinusoids
    with equally spaced arrival angles (random offset) and random phases
    (unit mean power). Slot t samples the paths at t * slot_s; the group
    response is the tap sum at the group's centre subcarrier,
    H_g[b][u] = sum_l h_l[b][u](t) exp(-2 pi i f_g tau_l),
    f_g = (g*sg + sg/2) * scs. slot(t) returns A = H^H as [G][B][U]."""

    def __init__(self, B: int, U: int, N: int, sg: int, seed: int, n_taps: int = 24,
                 rms_ns: float = 300.0, tap_ns: float = 65.0, fd_hz: float = 100.0,
                 slot_s: float = 0.5e-3, scs_hz: float = 30e3, n_sin: int = 16):
        self.B, self.U, self.N, self.sg = B, U, N, sg
        self.G = (N + sg - 1) // sg
        self.slot_s, self.fd = slot_s, fd_hz
        tau = np.arange(n_taps) * tap_ns * 1e-9
        self.tau = tau
        self.power = self._pdp(tau, rms_ns * 1e-9)
        rng = np.random.Generator(np.random.Philox(seed))
        P = B * U * n_taps
        m = np.arange(n_sin)
        theta = rng.uniform(0, 2 * np.pi, size=(P, 1))
        alpha = (2 * np.pi * m[None, :] + theta) / n_sin           # arrival angles
        self.omega = 2 * np.pi * fd_hz * np.cos(alpha)               # [P][n_sin]
        self.phi = rng.uniform(0, 2 * np.pi, size=(P, n_sin))
        fg = (np.arange(self.G) * sg + sg / 2.0) * scs_hz
        # [n_taps][G] tap phase at each group centre
        self.steer = np.exp(-2j * np.pi * tau[:, None] * fg[None, :])
        self.amp = np.sqrt(np.repeat(self.power[None, :], B * U, axis=0).reshape(P) / n_sin)
        self.n_taps = n_taps

    @staticmethod
    def _pdp(tau: np.ndarray, rms: float) -> np.ndarray:
        def spread(beta):
            p = np.exp(-tau / beta)
            p /= p.sum()
            m1 = (p * tau).sum()
            return math.sqrt(max((p * tau ** 2).sum() - m1 * m1, 0.0))
        lo, hi = 1e-12, 1e-3
        for _ in range(200):            # bisection: spread() grows with beta
            mid = math.sqrt(lo * hi)
            if spread(mid) < rms:
                lo = mid
            else:
                hi = mid
        p = np.exp(-tau / hi)
        return p / p.sum()

    def rms_delay_spread(self) -> float:
        """Discrete RMS delay spread of the power-delay profile (s)."""
        m1 = (self.power * self.tau).sum()
        return math.sqrt((self.power * self.tau ** 2).sum() - m1 * m1)

    def paths(self, t: int) -> np.ndarray:
        """h[b][u][l] at slot t (complex128)."""
        ts = t * self.slot_s
        h = self.amp * np.exp(1j * (self.omega * ts + self.phi)).sum(axis=1)
        return h.reshape(self.B, self.U, self.n_taps)

    def slot(self, t: int) -> np.ndarray:
        """A = H^H per group, complex64 [G][B][U]."""
        H = np.einsum("bul,lg->gbu", self.paths(t), self.steer, optimize=True)
        return np.conj(H).astype(np.complex64)


# --------------------------------------------------------------------------
# LLR layout (src/config/config.cc:84-135, src/operators/ldpc/config.h:61-72,
# src/baseband/buffers.cc:97-103 with the 128-byte kCoalesce padding)
# --------------------------------------------------------------------------

LDPC_ZC = sorted([2, 4, 8, 16, 32, 64, 128, 256, 3, 6, 12, 24, 48, 96, 192,
                  384, 5, 10, 20, 40, 80, 160, 320, 7, 14, 28, 56, 112, 224, 9,
                  18, 36, 72, 144, 288, 11, 22, 44, 88, 176, 352, 13, 26, 52,
                  104, 208, 15, 30, 60, 120, 240])


@dataclass(frozen=True)
class LlrLayout:
    cb_size: int      # bits of one code block (encoded bits)
    zeros: int        # punctured zero prefix per code block
    cb_stride: int    # halves per code-block row (padded to 64 halves)
    cb_blocks: int    # code blocks per user per symbol

    @property
    def ue_stride(self) -> int:
        return self.cb_stride * self.cb_blocks

    @property
    def valid_bits(self) -> int:
        return self.cb_size * self.cb_blocks


def pad64(n: int) -> int:
    """Leading dimension of a kCoalesce binary16 tensor: 128-byte multiple."""
    return (n + 63) // 64 * 64


def standalone_layout(N: int, Q: int) -> LlrLayout:
    """One code block holding every bit of the symbol, no punctured prefix."""
    return LlrLayout(cb_size=N * Q, zeros=0, cb_stride=pad64(N * Q), cb_blocks=1)


def ldpc_uplink_layout(N: int, Q: int, code_rate: float = 1 / 3,
                       base_graph: int = 1, cb_per_symbol: int = 1) -> LlrLayout:
    """The layout the reference pipeline hands to equalize: cb_size =
    encoded_bits, zeros = punctured_bits (config.cc:97-117, ldpc/config.h)."""
    info = 22 if base_graph == 1 else 10
    uncoded = int(N * Q * code_rate)
    zc = LDPC_ZC[-1]
    for i in range(len(LDPC_ZC) - 1):
        if info * LDPC_ZC[i] * cb_per_symbol <= uncoded < info * LDPC_ZC[i + 1] * cb_per_symbol:
            zc = LDPC_ZC[i]
            break
    rows = int(round(info / code_rate)) - (info - 2)
    encoded = (info - 2 + rows) * zc
    punct = 2 * zc
    return LlrLayout(cb_size=encoded, zeros=punct, cb_stride=pad64(encoded + punct),
                     cb_blocks=cb_per_symbol)


# --------------------------------------------------------------------------
# slots
# --------------------------------------------------------------------------

# Uneven batch of cells: antenna x user, subcarriers, group size and QAM
# order drawn per cell from these sets (seeded), 13 data symbols, i.i.d.
# CN(0,1) channels at 20 dB SNR.
BATCH_SHAPES = [(64, 8), (64, 16), (128, 16), (128, 32), (256, 64)]   # (B, U)
BATCH_N = [624, 1200, 3276]
BATCH_SG = [12, 16]
BATCH_Q = [4, 6, 8]


def uneven_cells(n_cells: int = 40, seed: int = 7) -> list[dict]:
    """Cell configurations of the uneven batch workload (dicts with U, B, N,
    sg, S, slots, Q, snr, channel)."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    for _ in range(n_cells):
        B, U = BATCH_SHAPES[rng.integers(len(BATCH_SHAPES))]
        out.append(dict(U=U, B=B, N=int(BATCH_N[rng.integers(3)]),
                        sg=int(BATCH_SG[rng.integers(2)]), S=13, slots=1,
                        Q=int(BATCH_Q[rng.integers(3)]), snr=20.0, channel="iid"))
    return out


@dataclass
class Slot:
    csi: np.ndarray    # complex64 [G][B][U]  (A = H^H)
    y: np.ndarray      # complex64 [S][B][N]
    x: np.ndarray      # complex64 [S][U][N]  transmitted symbols
    bits: np.ndarray   # uint8 [S][U][N][Q]
    H: np.ndarray      # complex64 [G][B][U]


def make_slot(rng, U: int, B: int, N: int, sg: int, S: int, Q: int,
              snr_db: float, channel: str = "iid", rho_rx: float = 0.5,
              rho_tx: float = 0.3, H: np.ndarray | None = None) -> Slot:
    G = (N + sg - 1) // sg
    if H is None:
        if channel == "iid":
            H = channel_iid(rng, G, B, U)
        elif channel == "kronecker":
            H = channel_kronecker(rng, G, B, U, rho_rx, rho_tx)
        else:
            raise ValueError(channel)
    bits = rng.integers(0, 2, size=(S, U, N, Q), dtype=np.uint8)
    x = modulate(bits).astype(np.complex64)                     # [S][U][N]
    sigma2 = 10.0 ** (-snr_db / 10.0)
    gidx = np.arange(N) // sg
    Hk = H[gidx]                                                # [N][B][U]
    # y[s, b, k] = sum_u H[g(k), b, u] x[s, u, k]
    y = np.einsum("kbu,suk->sbk", Hk, x, optimize=True)
    y = (y + _cn(rng, y.shape, sigma2)).astype(np.complex64)
    return Slot(csi=np.conj(H).astype(np.complex64), y=y, x=x, bits=bits, H=H)


def index_bits(idx: np.ndarray, Q: int) -> np.ndarray:
    """Symbol indices -> bits[..., Q], first stream bit (the index MSB)
    first: the order of a symbol's Q LLRs."""
    sh = np.arange(Q - 1, -1, -1)
    return ((np.asarray(idx)[..., None] >> sh) & 1).astype(np.uint8)


# Per-level thresholds of the max-log demod in units of the LLR scale 100
# (src/operators/modulation/demodulation.h:35-67): 200/sqrt(10);
# 400/sqrt(42), 200/sqrt(42); 800/sqrt(170), 400/sqrt(170), 200/sqrt(170).
_DEMOD_T = {4: (2 / math.sqrt(10),), 6: (4 / math.sqrt(42), 2 / math.sqrt(42)),
            8: (8 / math.sqrt(170), 4 / math.sqrt(170), 2 / math.sqrt(170))}


def hard_bits_xhat(xhat: np.ndarray, Q: int) -> np.ndarray:
    """Hard decisions of the max-log demod taken in float64 straight from
    x_hat (no binary16 rounding): bit j = 1 iff LLR j < 0, with
    l0 = re, l1 = im, l(2i+2) = t_i - |l(2i)|, l(2i+3) = t_i - |l(2i+1)|
    (QPSK: l = -re, -im). Returns uint8 [..., Q]."""
    x = np.asarray(xhat).astype(np.complex128)
    re, im = x.real, x.imag
    if Q == 2:
        ls = [-re, -im]
    else:
        ls = [re, im]
        for t in _DEMOD_T[Q]:
            ls += [t - np.abs(ls[-2]), t - np.abs(ls[-1])]
    return np.stack([(l < 0) for l in ls], axis=-1).astype(np.uint8)


def hard_bits_llr(llr_u16: np.ndarray) -> np.ndarray:
    """Hard decisions from binary16 LLR bit patterns: the sign bit
    (LLR > 0 means bit 0)."""
    return (np.asarray(llr_u16).astype(np.uint16) >> 15).astype(np.uint8)


def evm_db(xhat: np.ndarray, x: np.ndarray) -> float:
    """EVM in dB of equalized symbols against the transmitted ones."""
    err = np.mean(np.abs(xhat.astype(np.complex128) - x.astype(np.complex128)) ** 2)
    ref = np.mean(np.abs(x.astype(np.complex128)) ** 2)
    return 10 * math.log10(err / ref)
