"""Parity of the mixed form (MEGA_FORM_MIXED: split-BF16 matrix-core
products, refined FP32 inverse, FP64 route for ill-conditioned groups)
against the FP64 oracle, through the C ABI on cuda:0.

Bars: W within 1e-4 (relative Frobenius per group) of the FP64 oracle on
well-conditioned groups; FP64-route groups within 1e-6 (FP64 arithmetic
rounded to FP32); symbols by the same policy as the FP32 form (x_hat within
1e-4 of the FP64 oracle relative to the symbol's largest |x_hat|, LLRs
bit-exact against the demod of the kernel's own x_hat, hard bits equal away
from the boundary) and EVM within 0.1 dB of the FP64 oracle's."""
import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth

from gpu_helpers import (SENTINEL, check_symbol, dev_c64, host_c64, host_u16,
                         llr_buffer, make_slots, oracle_symbol_f64, torch)

pytestmark = pytest.mark.gpu

KAPPA_MAX = 1e3


def _sync():
    torch().cuda.synchronize()


def kappa1(a: np.ndarray, U: int, B: int, G: int) -> np.ndarray:
    """1-norm condition numbers of the Gram matrices A_g A_g^H (FP64)."""
    A = a.reshape(G, B, U).astype(np.complex128)          # A[g][b][u] = A_g[u][b]
    gram = np.matmul(np.swapaxes(A, 1, 2), A.conj())      # [g][u][v]
    n1 = lambda M: np.abs(M).sum(axis=-2).max(axis=-1)
    return n1(gram) * n1(np.linalg.inv(gram))


def _zf_mixed(a, U, B, G):
    t = torch()
    d_in = dev_c64(a)
    d_out = t.zeros_like(d_in)
    d_info = t.full((G,), -7, dtype=t.int32, device="cuda:0")
    m.zf(d_in, d_out, B, U, G, info=d_info, form=m.FORM_MIXED)
    _sync()
    assert np.array_equal(host_c64(d_in), a)   # out of place: input intact
    return host_c64(d_out).reshape(G, B * U), d_info.cpu().numpy()


@pytest.mark.parametrize("U,B,G,channel", [(16, 64, 75, "iid"), (32, 256, 40, "kronecker"),
                                           (8, 24, 17, "iid"), (64, 256, 6, "iid"),
                                           (5, 300, 9, "iid"), (3, 7, 4, "iid")])
def test_zf_mixed_matches_fp64_oracle(U, B, G, channel):
    print(f"case: zf mixed U={U} B={B} G={G} {channel}", flush=True)
    rng = np.random.Generator(np.random.Philox(U * 1000 + B + 1))
    H = (synth.channel_iid(rng, G, B, U) if channel == "iid"
         else synth.channel_kronecker(rng, G, B, U, 0.5, 0.3))
    a = np.conj(H).astype(np.complex64)
    assert kappa1(a, U, B, G).max() < KAPPA_MAX / 2   # all on the split-BF16 route
    _, w64, info64 = oracle.zf(a, U, B, G, "f64")
    w, info = _zf_mixed(a, U, B, G)
    ref = w64.reshape(G, B * U)
    rel = np.linalg.norm(w - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-4, rel.max()
    assert np.array_equal(info, info64)


def test_zf_mixed_routes_ill_conditioned_groups_to_fp64():
    """Groups with kappa_1(G) > 1e3 (a user almost parallel to another) take the
    FP64 route and match the FP64 oracle to FP32 rounding; an exactly
    repeated user is reported singular with W = 0; the rest stay on the
    split-BF16 route. route flags come back through mega_detect_slots."""
    print("case: zf mixed routes", flush=True)
    U, B, G = 32, 128, 8
    rng = np.random.Generator(np.random.Philox(77))
    H = synth.channel_iid(rng, G, B, U)
    a = np.conj(H).astype(np.complex64).reshape(G, B, U)
    for g, eps in ((1, 1e-2), (3, 3e-3), (6, 1e-1)):
        noise = (rng.standard_normal(B) + 1j * rng.standard_normal(B)) * eps
        a[g, :, 1] = (a[g, :, 0] + noise).astype(np.complex64)
    a[5, :, 7] = a[5, :, 2]                               # exactly singular
    a = a.reshape(-1)
    k = np.full(G, np.inf)
    ok = np.ones(G, bool)
    ok[5] = False
    k[ok] = kappa1(a.reshape(G, B, U)[ok].reshape(-1), U, B, int(ok.sum()))
    far = ok & ((k > 2 * KAPPA_MAX) | (k < KAPPA_MAX / 2))  # away from the threshold
    assert far.sum() >= 6 and (k[far] > KAPPA_MAX).sum() >= 2
    _, w64, info64 = oracle.zf(a, U, B, G, "f64")
    w, info = _zf_mixed(a, U, B, G)
    assert info[5] > 0 and np.all(w[5] == 0)
    assert np.all(info[ok] == 0) and np.array_equal(info[ok], info64[ok])
    ref = w64.reshape(G, B * U)
    rel = np.linalg.norm(w - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    fp64 = far & (k > KAPPA_MAX)
    assert rel[fp64].max() < 1e-6, rel[fp64]
    assert rel[far & ~fp64].max() < 1e-4, rel[far & ~fp64]

    # the same groups through detect_slots, which reports the route
    t = torch()
    N, sg, S, Q = G * 12, 12, 1, 4
    lay = synth.standalone_layout(N, Q)
    y = (rng.standard_normal((S, B, N)) + 1j * rng.standard_normal((S, B, N))).astype(np.complex64)
    work = t.empty(m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                                       num_slots=1, form=m.FORM_MIXED),
                   dtype=t.uint8, device="cuda:0")
    route = t.full((G,), -1, dtype=t.int32, device="cuda:0")
    info_d = t.full((G,), -1, dtype=t.int32, device="cuda:0")
    w_d = t.zeros(G * B * U, dtype=t.complex64, device="cuda:0")
    llr = llr_buffer(S * U * lay.ue_stride)
    m.detect_slots(dev_c64(a), dev_c64(y), llr, num_ue=U, num_bs=B, num_sc=N,
                   sc_group=sg, num_syms=S, num_slots=1, mod_order=Q,
                   ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
                   zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                   y_sym_stride=B * N, y_slot_stride=S * B * N,
                   llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
                   w_out=w_d, info=info_d, form=m.FORM_MIXED, work=work, route=route)
    _sync()
    r = route.cpu().numpy()
    assert r[5] == 1
    assert np.array_equal(r[far], (k[far] > KAPPA_MAX).astype(np.int32))
    assert np.array_equal(info_d.cpu().numpy(), info)
    assert np.array_equal(host_c64(w_d).reshape(G, B * U), w)   # same kernel as mega_zf


def _detect_mixed(csi, y, U, B, N, sg, Q, lay, slots, S):
    t = torch()
    G = (N + sg - 1) // sg
    llr = llr_buffer(slots * S * U * lay.ue_stride)
    xh = t.zeros(slots * S * U * N, dtype=t.complex64, device="cuda:0")
    info = t.full((slots * G,), -5, dtype=t.int32, device="cuda:0")
    route = t.full((slots * G,), -5, dtype=t.int32, device="cuda:0")
    work = t.empty(m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                                       num_slots=slots, form=m.FORM_MIXED),
                   dtype=t.uint8, device="cuda:0")
    m.detect_slots(dev_c64(csi), dev_c64(y), llr, num_ue=U, num_bs=B, num_sc=N,
                   sc_group=sg, num_syms=S, num_slots=slots, mod_order=Q,
                   ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
                   zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                   y_sym_stride=B * N, y_slot_stride=S * B * N,
                   llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
                   xhat=xh, info=info, form=m.FORM_MIXED, work=work, route=route)
    _sync()
    return (host_u16(llr).reshape(slots, S, U, lay.ue_stride),
            host_c64(xh).reshape(slots, S, U, N), info.cpu().numpy(),
            route.cpu().numpy())


@pytest.mark.parametrize("U,B,N,sg,Q,channel,snr,layout,slots", [
    (16, 64, 1200, 16, 4, "iid", 20.0, "standalone", 2),
    (16, 64, 1200, 16, 6, "iid", 30.0, "ldpc", 1),
    (32, 256, 3276, 12, 6, "kronecker", 25.0, "ldpc", 1),
    (64, 128, 600, 12, 8, "iid", 30.0, "standalone", 1),
    (5, 24, 100, 7, 2, "iid", 15.0, "standalone", 1),
    (8, 40, 204, 12, 4, "iid", 20.0, "ldpc", 1),
    # W-in-registers equalize: 8 antenna steps x 16 users (partial last group),
    # 4 steps x 32 users over two slots
    (16, 128, 610, 12, 6, "iid", 25.0, "ldpc", 1),
    (32, 64, 480, 16, 4, "iid", 20.0, "standalone", 2),
])
def test_detect_mixed_matches_fp64_oracle(U, B, N, sg, Q, channel, snr, layout, slots):
    print(f"case: detect mixed U={U} B={B} N={N} sg={sg} Q={Q} {channel}", flush=True)
    S = 2
    sl = make_slots(U * 13 + N + Q, slots, U, B, N, sg, S, Q, snr, channel)
    lay = synth.ldpc_uplink_layout(N, Q) if layout == "ldpc" else synth.standalone_layout(N, Q)
    G = (N + sg - 1) // sg
    csi = np.stack([s.csi for s in sl])
    y = np.stack([s.y for s in sl])
    llr, xh, info, route = _detect_mixed(csi, y, U, B, N, sg, Q, lay, slots, S)
    assert not info.any() and not route.any()
    for s_ in range(slots):
        w64, _, _ = oracle.zf(sl[s_].csi, U, B, G, "f64")
        for t_ in range(S):
            llr64, x64 = oracle_symbol_f64(w64, sl[s_].y[t_], U, B, N, sg, Q, lay)
            check_symbol(llr[s_, t_], xh[s_, t_], llr64, x64, N, sg, Q, lay)
            e_mix = synth.evm_db(xh[s_, t_], sl[s_].x[t_])
            e_64 = synth.evm_db(x64, sl[s_].x[t_])
            assert abs(e_mix - e_64) < 0.1, (e_mix, e_64)


@pytest.mark.parametrize("U,B,N,ks", [(32, 256, 600, 1), (32, 256, 600, 2), (32, 256, 600, 4),
                                      (64, 128, 300, 2), (16, 256, 250, 4)])
def test_mixed_equalize_antenna_split(tuning, U, B, N, ks):
    """The streamed mixed equalize with the antennas of a unit split over
    ks waves of a workgroup (partial sums met in LDS), every split the
    launcher can take, forced by MEGA_EQ_XKS; a split changes the FP32
    summation order only, so each is held to the FP64 oracle's bar
    (partial last group included: N % sg != 0)."""
    tuning(MEGA_EQ_XKS=ks)
    print(f"case: mixed equalize antenna split ks={ks} U={U} B={B} N={N}", flush=True)
    sg, Q, S, slots = 12, 6, 2, 1
    sl = make_slots(U + B + N + ks, slots, U, B, N, sg, S, Q, 25.0, "iid")
    lay = synth.ldpc_uplink_layout(N, Q)
    G = (N + sg - 1) // sg
    csi = np.stack([s.csi for s in sl])
    y = np.stack([s.y for s in sl])
    llr, xh, info, route = _detect_mixed(csi, y, U, B, N, sg, Q, lay, slots, S)
    assert not info.any() and not route.any()
    w64, _, _ = oracle.zf(sl[0].csi, U, B, G, "f64")
    for t_ in range(S):
        llr64, x64 = oracle_symbol_f64(w64, sl[0].y[t_], U, B, N, sg, Q, lay)
        check_symbol(llr[0, t_], xh[0, t_], llr64, x64, N, sg, Q, lay)


@pytest.mark.parametrize("rho_tx,kmin,kmax", [(0.3, 100.0, KAPPA_MAX / 2),
                                              (0.64, KAPPA_MAX / 2, KAPPA_MAX)])
def test_mixed_split_route_holds_fp32_bar_up_to_route_threshold(rho_tx, kmin, kmax):
    """Split-route accuracy on ill-conditioned groups that stay on that route
    (32 users on 64 correlated antennas; kappa_1 from about 180 up to just
    under the FP64 route's 1e3): the Gram matrix the inverse is refined
    against and W = X A are both taken on three-piece splits, so the
    condition number multiplies only FP32-sized errors. Bar: the FP32 form's
    (x_hat within 1e-4 of the FP64 oracle relative to the symbol's largest
    |x_hat|, LLRs bit-exact against the demod of the kernel's own x_hat),
    EVM within 0.1 dB of the FP64 oracle's, no group on the FP64 route."""
    U, B, N, sg, Q, S, slots = 32, 64, 480, 16, 4, 2, 2
    print(f"case: detect mixed kappa U={U} B={B} kronecker rho_tx={rho_tx}", flush=True)
    rng = np.random.Generator(np.random.Philox(U * 13 + N + Q + int(100 * rho_tx)))
    sl = [synth.make_slot(rng, U, B, N, sg, S, Q, 20.0, "kronecker", rho_tx=rho_tx)
          for _ in range(slots)]
    lay = synth.standalone_layout(N, Q)
    G = (N + sg - 1) // sg
    kap = max(kappa1(s_.csi, U, B, G).max() for s_ in sl)
    assert kmin < kap < kmax, kap
    csi = np.stack([s_.csi for s_ in sl])
    y = np.stack([s_.y for s_ in sl])
    llr, xh, info, route = _detect_mixed(csi, y, U, B, N, sg, Q, lay, slots, S)
    assert not info.any() and not route.any()
    for s_ in range(slots):
        w64, _, _ = oracle.zf(sl[s_].csi, U, B, G, "f64")
        for t_ in range(S):
            llr64, x64 = oracle_symbol_f64(w64, sl[s_].y[t_], U, B, N, sg, Q, lay)
            check_symbol(llr[s_, t_], xh[s_, t_], llr64, x64, N, sg, Q, lay, rtol=1e-4)
            e_mix = synth.evm_db(xh[s_, t_], sl[s_].x[t_])
            e_64 = synth.evm_db(x64, sl[s_].x[t_])
            assert abs(e_mix - e_64) < 0.1, (e_mix, e_64)


def test_mixed_halves_equal_detect_slots():
    """mega_zf_split followed by mega_equalize_slots(form=MIXED) is the mixed
    mega_detect_slots, bit for bit (same kernels), also with a padded slot
    stride of the channel estimates."""
    print("case: zf_split + equalize_slots == detect_slots (mixed)", flush=True)
    t = torch()
    U, B, N, sg, Q, S, slots = 16, 64, 240, 16, 4, 3, 2
    G = (N + sg - 1) // sg
    sl = make_slots(41, slots, U, B, N, sg, S, Q)
    lay = synth.standalone_layout(N, Q)
    pad = 3 * B * U                                  # padded csi slot stride
    csi = np.zeros((slots, G * B * U + pad), np.complex64)
    for s_ in range(slots):
        csi[s_, :G * B * U] = sl[s_].csi.reshape(-1)
    y = np.stack([s_.y for s_ in sl])
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=slots,
              mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
              cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
              y_sym_stride=B * N, y_slot_stride=S * B * N,
              llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride)
    d_csi, d_y = dev_c64(csi), dev_c64(y)
    work = t.empty(m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                                       num_slots=slots, form=m.FORM_MIXED),
                   dtype=t.uint8, device="cuda:0")
    llr_a = llr_buffer(slots * S * U * lay.ue_stride)
    m.detect_slots(d_csi, d_y, llr_a, csi_slot_stride=G * B * U + pad,
                   form=m.FORM_MIXED, work=work, **kw)
    Bp = (B + 15) // 16 * 16
    assert work.numel() == slots * G * U * Bp * 8
    split = t.empty(slots * G * U * Bp, dtype=t.int64, device="cuda:0")
    for s_ in range(slots):
        m.zf_split(d_csi[s_], split[s_ * G * U * Bp:], B, U, G)
    llr_b = llr_buffer(slots * S * U * lay.ue_stride)
    m.equalize_slots(split, d_y, llr_b, csi_slot_stride=G * U * Bp, form=m.FORM_MIXED, **kw)
    _sync()
    a, b = host_u16(llr_a), host_u16(llr_b)
    assert np.array_equal(a, b)
    assert (a != SENTINEL).sum() == slots * S * U * min(N * Q, lay.valid_bits)


def test_mixed_rejects_missing_work():
    print("case: mixed argument checks", flush=True)
    U, B, N, sg, Q = 8, 16, 32, 16, 4
    lay = synth.standalone_layout(N, Q)
    t = torch()
    d = t.zeros(2 * B * U, dtype=t.complex64, device="cuda:0")
    yv = t.zeros(B * N, dtype=t.complex64, device="cuda:0")
    llr = llr_buffer(U * lay.ue_stride, 0)
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=1, num_slots=1,
              mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
              cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
              csi_slot_stride=2 * B * U, y_sym_stride=B * N, y_slot_stride=B * N,
              llr_sym_stride=U * lay.ue_stride, llr_slot_stride=U * lay.ue_stride)
    with pytest.raises(m.MegaError) as e:
        m.detect_slots(d, yv, llr, form=m.FORM_MIXED, **kw)
    assert e.value.code == m.MEGA_ERR_INVALID_ARG
    with pytest.raises(m.MegaError) as e:
        m.equalize_slots(d, yv, llr, form=7, **kw)
    assert e.value.code == m.MEGA_ERR_UNSUPPORTED
    _sync()
    assert np.all(host_u16(llr) == 0)
