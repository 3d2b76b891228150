"""Shared helpers of the GPU parity tests: device buffers, the equalize call
through the C ABI, and the comparison policy against the oracle."""
from __future__ import annotations

import numpy as np

import megastation_amd as m
import oracle
from megastation_amd import synth

SENTINEL = 0x7E01  # a binary16 NaN pattern the demod never produces


def torch():
    import torch as _t
    return _t


def dev_c64(a):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to("cuda:0")


def llr_buffer(n_halves: int, fill: int = SENTINEL):
    t = torch()
    return t.full((n_halves,), np.int16(np.uint16(fill).view(np.int16)).item(),
                  dtype=t.int16, device="cuda:0")


def host_u16(t) -> np.ndarray:
    return t.cpu().numpy().view(np.uint16)


def host_c64(t) -> np.ndarray:
    return t.cpu().numpy()


def rel_fro(a, b) -> float:
    """||a - b||_F / ||b||_F. Complex inputs are compared as complex128 (real
    and imaginary parts both count); real inputs as float64."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def assert_imag_counts(ref, floor: float = 0.1) -> None:
    """Guard for a complex check: for the test's own reference data the
    conjugate must be far from it under rel_fro, so a comparison that
    dropped the imaginary parts could not pass."""
    assert np.iscomplexobj(ref), "complex check on real data"
    assert rel_fro(np.conj(ref), ref) > floor, "reference has no imaginary content"


def gpu_equalize_slots(csi, y, U, B, N, sg, Q, lay, *, slots, S, want_xhat=True,
                       fill=SENTINEL):
    """csi complex64 [slots][G][B][U], y [slots][S][B][N]. Returns (llr uint16
    [slots][S][U][ue_stride], xhat complex64 [slots][S][U][N] or None)."""
    t = torch()
    G = (N + sg - 1) // sg
    assert csi.shape == (slots, G, B, U) and y.shape == (slots, S, B, N)
    d_csi, d_y = dev_c64(csi), dev_c64(y)
    llr = llr_buffer(slots * S * U * lay.ue_stride, fill)
    xh = t.zeros(slots * S * U * N, dtype=t.complex64, device="cuda:0") if want_xhat else None
    m.equalize_slots(d_csi, d_y, llr, num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                     num_syms=S, num_slots=slots, mod_order=Q,
                     ue_stride=lay.ue_stride, cb_size=lay.cb_size,
                     cb_stride=lay.cb_stride, zeros=lay.zeros,
                     valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                     y_sym_stride=B * N, y_slot_stride=S * B * N,
                     llr_sym_stride=U * lay.ue_stride,
                     llr_slot_stride=S * U * lay.ue_stride, xhat=xh)
    t.cuda.synchronize()
    llr_h = host_u16(llr).reshape(slots, S, U, lay.ue_stride)
    xh_h = host_c64(xh).reshape(slots, S, U, N) if want_xhat else None
    return llr_h, xh_h


def oracle_llr_from_xhat(xhat_un: np.ndarray, N: int, sg: int, Q: int, lay,
                         fill=SENTINEL) -> np.ndarray:
    """LLRs the oracle places for given equalized symbols x[U][N]: the oracle
    equalizes y = x with W_g = I (exact: products with 1 and 0, sums with 0)."""
    U = xhat_un.shape[0]
    G = (N + sg - 1) // sg
    eye = np.broadcast_to(np.eye(U, dtype=np.complex64), (G, U, U)).copy()
    init = np.full((U, lay.ue_stride), fill, dtype=np.uint16)
    llr, _ = oracle.equalize(eye, xhat_un, U, U, N, sg, Q, lay.ue_stride,
                             lay.cb_size, lay.cb_stride, lay.zeros, lay.valid_bits,
                             "f32", llr_init=init)
    return llr


def f16(bits: np.ndarray) -> np.ndarray:
    return bits.view(np.float16).astype(np.float64)


# Comparison policy (stated once, used by every parity test):
#   * x_hat (FP32) within 1e-4 relative of the FP64 oracle, relative to the
#     largest |x_hat| of the symbol (the GEMV's forward-error scale);
#   * LLRs bit-exact against the oracle's demod + placement of the kernel's
#     own x_hat (pins demod and index math exactly);
#   * hard decisions equal to the FP64 oracle's except where the FP64 LLR is
#     within HARD_BIT_MARGIN of the decision boundary (0.05 LLR units = 5e-4
#     in x_hat units at scale 100, far above FP32 GEMV error).
XHAT_RTOL = 1e-4
HARD_BIT_MARGIN = 0.05


def check_symbol(llr_gpu, xhat_gpu, llr64, xhat64, N, sg, Q, lay, rtol=XHAT_RTOL):
    """One symbol: llr [U][ue_stride] uint16, xhat [U][N]."""
    scale = np.abs(xhat64).max()
    err = np.abs(xhat_gpu.astype(np.complex128) - xhat64).max()
    assert err <= rtol * scale, (err, scale)
    own = oracle_llr_from_xhat(xhat_gpu, N, sg, Q, lay)
    mism = np.flatnonzero(own.reshape(-1) != llr_gpu.reshape(-1))
    assert mism.size == 0, f"{mism.size} LLRs differ from demod of own x_hat, first at {mism[:5]}"
    written = llr64 != SENTINEL
    assert np.array_equal(written, llr_gpu != SENTINEL)
    a, b = f16(llr_gpu[written]), f16(llr64[written])
    bad = (np.signbit(a) != np.signbit(b)) & (np.abs(b) >= HARD_BIT_MARGIN)
    assert not bad.any(), f"{bad.sum()} hard bits differ away from the boundary"


def oracle_symbol_f64(w_csi, y_bn, U, B, N, sg, Q, lay):
    init = np.full((U, lay.ue_stride), SENTINEL, dtype=np.uint16)
    return oracle.equalize(w_csi, y_bn, U, B, N, sg, Q, lay.ue_stride, lay.cb_size,
                           lay.cb_stride, lay.zeros, lay.valid_bits, "f64",
                           llr_init=init)


def make_slots(seed, slots, U, B, N, sg, S, Q, snr_db=20.0, channel="iid"):
    rng = np.random.Generator(np.random.Philox(seed))
    return [synth.make_slot(rng, U, B, N, sg, S, Q, snr_db, channel) for _ in range(slots)]
