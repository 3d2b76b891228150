"""Parity of the HIP detection path (through the C ABI) with the CPU oracle.

All tests run on cuda:0 of an MI355X. Inputs are seeded; sizes are chosen so
the oracle finishes in seconds. Full-size workloads are checked through
size-independent properties (EVM against FP64, demod bit-exactness of the
kernel's own x_hat, run-to-run determinism)."""
import ctypes
import math

import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth

from gpu_helpers import (SENTINEL, check_symbol, dev_c64, gpu_equalize_slots,
                         host_c64, host_u16, llr_buffer, make_slots,
                         oracle_symbol_f64, torch)

pytestmark = pytest.mark.gpu


def _sync():
    torch().cuda.synchronize()


# ----------------------------------------------------------------------------
# reference fixtures through the drop-in entry points
# ----------------------------------------------------------------------------


def test_equalize_reference_fixture_bit_exact(golden):
    """test/sgemv/test_sgemv.cc:126-159 via mega_equalize (raw overload)."""
    print("case: sgemv fixture U=4 B=8 N=16 QPSK", flush=True)
    U, B, N = 4, 8, 16
    csi = dev_c64(golden["sgemv_csi"])
    y = dev_c64(golden["sgemv_bsvec_kb"].reshape(N, B).T)
    out = llr_buffer(U * 2 * N, 0)
    m.Equalize.equalize_raw(U, B, N, csi, y, out, 2 * N, 16, 2, 2 * N, 2 * N, 0, 2 * N)
    _sync()
    assert np.array_equal(host_u16(out), golden["sgemv_equalize_f16bits"])


def test_beamformer_reference_fixture(golden):
    """test/beamform/test_beamform.cc:67-95: Beamform(8,4,2).beamformer."""
    print("case: beamform fixture U=4 B=8 batch=2", flush=True)
    t = torch()
    a = np.concatenate([golden["beam_csi_in"]] * 2)
    csi = dev_c64(a)
    prec = t.zeros_like(csi)
    bf = m.Beamform(8, 4, 2)
    bf.beamformer(csi, prec, batch=2)
    info = bf.info(2)
    w, p = host_c64(csi), host_c64(prec)
    assert not info.any()
    for g in range(2):
        assert np.abs(w[32 * g:32 * g + 32] - golden["beam_csi_out"]).max() < 1e-5
        assert np.abs(p[32 * g:32 * g + 32] - golden["beam_precoder"]).max() < 1e-5
    bf.close()


# ----------------------------------------------------------------------------
# ZF
# ----------------------------------------------------------------------------


@pytest.mark.parametrize("U,B,G,channel", [(4, 8, 3, "iid"), (16, 64, 75, "iid"),
                                           (32, 256, 40, "kronecker"),
                                           (8, 24, 17, "iid"), (64, 256, 6, "iid"),
                                           (5, 300, 9, "iid"),
                                           # 65 to 128 users: 16-wave workgroups
                                           (128, 256, 5, "iid"), (96, 384, 4, "kronecker"),
                                           (65, 130, 3, "iid")])
def test_zf_matches_fp64_oracle(U, B, G, channel):
    print(f"case: zf U={U} B={B} G={G} {channel}", flush=True)
    rng = np.random.Generator(np.random.Philox(U * 1000 + B))
    H = (synth.channel_iid(rng, G, B, U) if channel == "iid"
         else synth.channel_kronecker(rng, G, B, U, 0.5, 0.3))
    a = np.conj(H).astype(np.complex64)
    _, w64, info64 = oracle.zf(a, U, B, G, "f64")
    t = torch()
    d_in = dev_c64(a)
    d_out = t.zeros_like(d_in)
    d_info = t.full((G,), -7, dtype=t.int32, device="cuda:0")
    m.zf(d_in, d_out, B, U, G, info=d_info)
    _sync()
    w = host_c64(d_out).reshape(G, B * U).astype(np.complex128)
    ref = w64.reshape(G, B * U)
    rel = np.linalg.norm(w - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-4, rel.max()
    assert np.array_equal(d_info.cpu().numpy(), info64)
    assert np.array_equal(host_c64(d_in), a)       # out-of-place leaves input intact


@pytest.mark.parametrize("U,B,G", [(64, 256, 40), (40, 200, 17)])
def test_zf_thread_counts_are_bit_identical(U, B, G, tuning):
    """33 to 64 users on 4 or 16 waves per group (MEGA_ZF_NTH): the Gram's
    chunks are added in the same order, every sweep entry and apply product
    is the same arithmetic, so W and the status are the same bits; both
    against the FP64 oracle."""
    print(f"case: zf thread counts U={U} B={B} G={G}", flush=True)
    rng = np.random.Generator(np.random.Philox(U * 7 + B))
    H = synth.channel_kronecker(rng, G, B, U, 0.5, 0.3)
    H[3] = 0                                           # one failed group
    a = np.conj(H).astype(np.complex64)
    _, w64, info64 = oracle.zf(a, U, B, G, "f64")
    t = torch()
    outs = []
    for nth in (256, 0):                              # 0: the default, 16 waves
        tuning(MEGA_ZF_NTH=nth)
        d_out = t.zeros(G * B * U, dtype=t.complex64, device="cuda:0")
        d_info = t.full((G,), -7, dtype=t.int32, device="cuda:0")
        m.zf(dev_c64(a), d_out, B, U, G, info=d_info)
        _sync()
        outs.append((host_c64(d_out), d_info.cpu().numpy()))
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][1], info64)
    ok = info64 == 0
    w = outs[0][0].reshape(G, B * U).astype(np.complex128)[ok]
    ref = w64.reshape(G, B * U)[ok]
    assert (np.linalg.norm(w - ref, axis=1) / np.linalg.norm(ref, axis=1)).max() < 1e-4


@pytest.mark.parametrize("U,B", [(8, 32), (100, 200)])
def test_zf_in_place_and_failed_groups(U, B):
    print(f"case: zf in place with rank-deficient groups U={U} B={B}", flush=True)
    rng = np.random.Generator(np.random.Philox(9))
    G = 10
    H = synth.channel_iid(rng, G, B, U)
    H[3, :, 5] = H[3, :, 1]
    H[7] = 0
    a = np.conj(H).astype(np.complex64)
    w32, _, info64 = oracle.zf(a, U, B, G, "f64")
    bf = m.Beamform(B, U, G)
    d = dev_c64(a)
    bf.beamform_uplink(d)
    info = bf.info(G)
    assert info[7] == 1 and info[3] != 0
    ok = info == 0
    assert np.array_equal(ok, info64 == 0)
    w = host_c64(d).reshape(G, -1)
    assert not w[~ok].any()
    ref = w32.reshape(G, -1)
    assert np.abs(w[ok] - ref[ok]).max() < 1e-4 * np.abs(ref[ok]).max()
    with pytest.raises(m.MegaError) as e:
        bf.beamform_uplink(d, batch=G + 1)
    assert e.value.code == m.MEGA_ERR_BATCH
    bf.close()


def test_beamform_downlink_normalisation(orc):
    print("case: scale_absmax", flush=True)
    rng = np.random.Generator(np.random.Philox(4))
    U, B, G = 16, 64, 33
    w = synth._cn(rng, (G, B, U))
    t = torch()
    d = dev_c64(w)
    p = t.zeros_like(d)
    m.Beamform.beamform_downlink(d, p, U, B, G)
    _sync()
    ref = orc.scale_absmax(w, U, B, G)
    assert np.abs(host_c64(p).reshape(-1) - ref).max() < 1e-6


# ----------------------------------------------------------------------------
# equalize + demod
# ----------------------------------------------------------------------------

EQ_CASES = [
    # U, B, N, sg, Q, layout
    (4, 8, 16, 16, 2, "standalone"),
    (16, 64, 1200, 16, 4, "standalone"),
    (16, 64, 1200, 16, 6, "ldpc"),
    (32, 256, 3276, 12, 6, "ldpc"),
    (8, 64, 624, 12, 8, "standalone"),
    (32, 128, 3276, 16, 8, "standalone"),   # partial last group (204.75 groups)
    (64, 128, 600, 12, 4, "ldpc"),
    (3, 5, 37, 7, 6, "standalone"),
    (20, 200, 500, 12, 6, "ldpc"),          # wide form: 20 users, 200 antennas
    (32, 96, 300, 7, 4, "standalone"),      # wide form: a part-filled antenna slice
    (128, 256, 144, 12, 4, "standalone"),   # 128 users: four user groups of 32
    (48, 256, 240, 12, 6, "ldpc"),          # wide form, four user tiles, the last empty
    (80, 128, 96, 16, 6, "ldpc"),           # a part-filled user group
]


# MEGA_EQ_MODE: 0 = the launcher's choice, 1 = W streamed from memory,
# 2 = W held in registers (antennas split over waves when B is wide),
# 3 = the wide form (antennas split over a workgroup's waves, three-product
# MFMAs) where it applies (up to 32 users, 33 to 256 antennas); 3w8 = the
# wide form on eight waves per workgroup (MEGA_EQ_WNW=8); 3p = its persistent
# form (MEGA_EQ_WP=2: 512 workgroups over balanced tile ranges)
@pytest.mark.parametrize("mode", ["0", "1", "2", "3", "3w8", "3p"])
@pytest.mark.parametrize("U,B,N,sg,Q,layout", EQ_CASES)
def test_equalize_matches_oracle(U, B, N, sg, Q, layout, mode, tuning):
    print(f"case: equalize U={U} B={B} N={N} sg={sg} Q={Q} {layout} mode={mode}", flush=True)
    if mode == "3w8":
        tuning(MEGA_EQ_MODE=3, MEGA_EQ_WNW=8)
    elif mode == "3p":
        tuning(MEGA_EQ_MODE=3, MEGA_EQ_WP=2)
    else:
        tuning(MEGA_EQ_MODE=mode)
    S, slots = 2, 1
    sl = make_slots(U * 7 + N, slots, U, B, N, sg, S, Q)
    lay = synth.ldpc_uplink_layout(N, Q) if layout == "ldpc" else synth.standalone_layout(N, Q)
    G = (N + sg - 1) // sg
    w, _, _ = oracle.zf(sl[0].csi, U, B, G, "f64")   # equalize with exact weights
    w = w.reshape(1, G, B, U)
    y = np.stack([s.y for s in sl])
    llr, xh = gpu_equalize_slots(w, y, U, B, N, sg, Q, lay, slots=slots, S=S)
    for s in range(S):
        llr64, x64 = oracle_symbol_f64(w[0], y[0, s], U, B, N, sg, Q, lay)
        check_symbol(llr[0, s], xh[0, s], llr64, x64, N, sg, Q, lay)


@pytest.mark.parametrize("U,B,N,sg,Q", [(32, 256, 3276, 12, 6), (24, 128, 610, 16, 4)])
def test_equalize_wide_chunking_is_bit_identical(U, B, N, sg, Q, tuning):
    """The wide form over two slots and three symbols: against the FP64
    oracle, and bit-identical LLRs and x_hat whatever the number of column
    tiles per workgroup (a tile's arithmetic does not depend on the chunk
    it is in)."""
    print(f"case: equalize wide chunking U={U} B={B} N={N} sg={sg} Q={Q}", flush=True)
    S, slots = 3, 2
    sl = make_slots(U + B + N, slots, U, B, N, sg, S, Q, 25.0)
    lay = synth.ldpc_uplink_layout(N, Q)
    G = (N + sg - 1) // sg
    w = np.stack([oracle.zf(s_.csi, U, B, G, "f64")[0].reshape(G, B, U) for s_ in sl])
    y = np.stack([s_.y for s_ in sl])
    for nw in (4, 8):                                 # waves per workgroup
        outs = []
        for tc in (0, 1, 4, 100):
            tuning(MEGA_EQ_MODE=3, MEGA_EQ_WTC=tc, MEGA_EQ_WNW=nw)
            outs.append(gpu_equalize_slots(w, y, U, B, N, sg, Q, lay, slots=slots, S=S))
        if nw == 8:                                   # the persistent form: same bits
            for k in (1, 2, 8):
                tuning(MEGA_EQ_MODE=3, MEGA_EQ_WP=k)
                outs.append(gpu_equalize_slots(w, y, U, B, N, sg, Q, lay, slots=slots, S=S))
        for llr, xh in outs[1:]:
            assert np.array_equal(llr, outs[0][0]) and np.array_equal(xh, outs[0][1])
        llr, xh = outs[0]
        for s_ in range(slots):
            for t_ in range(S):
                llr64, x64 = oracle_symbol_f64(w[s_], y[s_, t_], U, B, N, sg, Q, lay)
                check_symbol(llr[s_, t_], xh[s_, t_], llr64, x64, N, sg, Q, lay)


def test_equalize_slots_equals_per_symbol_calls():
    """Batched launch == one mega_equalize per symbol, bit for bit."""
    print("case: slots vs per-symbol", flush=True)
    U, B, N, sg, Q, S, slots = 16, 64, 240, 16, 4, 3, 2
    sl = make_slots(77, slots, U, B, N, sg, S, Q)
    lay = synth.standalone_layout(N, Q)
    csi = np.stack([s.csi for s in sl])
    y = np.stack([s.y for s in sl])
    llr, _ = gpu_equalize_slots(csi, y, U, B, N, sg, Q, lay, slots=slots, S=S,
                                want_xhat=False)
    for s in range(slots):
        for t_ in range(S):
            out = llr_buffer(U * lay.ue_stride)
            m.Equalize.equalize_raw(U, B, N, dev_c64(csi[s]), dev_c64(y[s, t_]), out,
                                    lay.ue_stride, sg, Q, lay.cb_size, lay.cb_stride,
                                    lay.zeros, lay.valid_bits)
            _sync()
            assert np.array_equal(host_u16(out).reshape(U, -1), llr[s, t_])


def test_equalize_ownership_and_determinism():
    """Two runs with different sentinels: the same positions are written, with
    the same values (each (user, subcarrier, bit) stored exactly once, no
    stores into a neighbouring group's outputs at sg = 12)."""
    print("case: ownership sg=12", flush=True)
    U, B, N, sg, Q, S = 32, 64, 1200, 12, 6, 3
    sl = make_slots(5, 1, U, B, N, sg, S, Q)
    lay = synth.ldpc_uplink_layout(N, Q)
    csi = sl[0].csi[None]
    y = sl[0].y[None]
    a, _ = gpu_equalize_slots(csi, y, U, B, N, sg, Q, lay, slots=1, S=S,
                              want_xhat=False, fill=SENTINEL)
    b, _ = gpu_equalize_slots(csi, y, U, B, N, sg, Q, lay, slots=1, S=S,
                              want_xhat=False, fill=0x7E02)
    wa, wb = a != SENTINEL, b != 0x7E02
    assert np.array_equal(wa, wb)
    assert np.array_equal(a[wa], b[wb])
    per_user = wa.sum(-1)
    assert np.all(per_user == min(N * Q, lay.valid_bits))


def test_equalize_rejects_bad_modulation_without_writing():
    print("case: bad Q", flush=True)
    out = llr_buffer(64, 0)
    with pytest.raises(m.MegaError) as e:
        m.Equalize.equalize_raw(4, 8, 16, dev_c64(np.ones(32)), dev_c64(np.ones(128)),
                                out, 32, 16, 5, 32, 32, 0, 32)
    assert e.value.code == m.MEGA_ERR_UNSUPPORTED
    _sync()
    assert not host_u16(out).any()


# ----------------------------------------------------------------------------
# end to end at the cell shapes of the benchmark configurations
# ----------------------------------------------------------------------------


@pytest.mark.parametrize("U,B,N,sg,Q,channel,snr", [
    (16, 64, 1200, 16, 4, "iid", 20.0),
    (16, 64, 1200, 16, 6, "iid", 30.0),
    (32, 256, 3276, 12, 6, "kronecker", 25.0),
])
def test_detect_end_to_end_evm(U, B, N, sg, Q, channel, snr):
    """ZF on the GPU, then equalize+demod: EVM within 0.1 dB of the FP64
    oracle's, symbol-level checks as in test_equalize_matches_oracle."""
    print(f"case: e2e U={U} B={B} N={N} sg={sg} Q={Q} {channel}", flush=True)
    S = 2
    sl = make_slots(U + N + Q, 1, U, B, N, sg, S, Q, snr, channel)[0]
    lay = synth.ldpc_uplink_layout(N, Q)
    G = (N + sg - 1) // sg
    t = torch()
    d_a = dev_c64(sl.csi)
    d_w = t.zeros_like(d_a)
    m.zf(d_a, d_w, B, U, G)
    _sync()
    w_gpu = host_c64(d_w).reshape(1, G, B, U)
    llr, xh = gpu_equalize_slots(w_gpu, sl.y[None], U, B, N, sg, Q, lay, slots=1, S=S)
    w64, _, _ = oracle.zf(sl.csi, U, B, G, "f64")
    for s in range(S):
        llr64, x64 = oracle_symbol_f64(w64, sl.y[s], U, B, N, sg, Q, lay)
        check_symbol(llr[0, s], xh[0, s], llr64, x64, N, sg, Q, lay)
        e_gpu = synth.evm_db(xh[0, s], sl.x[s])
        e_64 = synth.evm_db(x64, sl.x[s])
        assert abs(e_gpu - e_64) < 0.1, (e_gpu, e_64)


# ----------------------------------------------------------------------------
# fused detection (mega_detect_slots)
# ----------------------------------------------------------------------------


def _detect(csi, y, U, B, N, sg, Q, lay, slots, S, want_w=True, want_x=True):
    t = torch()
    G = (N + sg - 1) // sg
    d_a, d_y = dev_c64(csi), dev_c64(y)
    llr = llr_buffer(slots * S * U * lay.ue_stride)
    w = t.zeros_like(d_a) if want_w else None
    xh = t.zeros(slots * S * U * N, dtype=t.complex64, device="cuda:0") if want_x else None
    info = t.full((slots * G,), -5, dtype=t.int32, device="cuda:0")
    m.detect_slots(d_a, d_y, llr, num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                   num_syms=S, num_slots=slots, mod_order=Q, ue_stride=lay.ue_stride,
                   cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                   valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                   y_sym_stride=B * N, y_slot_stride=S * B * N,
                   llr_sym_stride=U * lay.ue_stride,
                   llr_slot_stride=S * U * lay.ue_stride, xhat=xh, w_out=w, info=info)
    _sync()
    return (host_u16(llr).reshape(slots, S, U, lay.ue_stride),
            host_c64(xh).reshape(slots, S, U, N) if want_x else None,
            host_c64(w).reshape(slots, G, B, U) if want_w else None,
            info.cpu().numpy().reshape(slots, G))


@pytest.mark.parametrize("U,B,N,sg,Q", [(16, 64, 1200, 16, 4), (8, 64, 624, 12, 8),
                                        (16, 128, 600, 16, 6), (12, 192, 300, 12, 2),
                                        (16, 256, 240, 16, 4), (5, 24, 100, 7, 6),
                                        (32, 128, 480, 16, 6), (128, 256, 96, 12, 4)])
def test_detect_slots_equals_zf_then_equalize(U, B, N, sg, Q, tuning):
    """Fused kernel == ZF kernel followed by the equalize kernel with W in
    registers, bit for bit (same arithmetic and summation order), and x_hat /
    EVM against the FP64 oracle."""
    print(f"case: detect U={U} B={B} N={N} sg={sg} Q={Q}", flush=True)
    tuning(MEGA_EQ_MODE=2)
    S, slots = 3, 2
    sl = make_slots(U * 31 + B, slots, U, B, N, sg, S, Q, 25.0)
    lay = synth.standalone_layout(N, Q)
    G = (N + sg - 1) // sg
    csi = np.stack([s.csi for s in sl])
    y = np.stack([s.y for s in sl])
    llr_f, x_f, w_f, info = _detect(csi, y, U, B, N, sg, Q, lay, slots, S)
    assert not info.any()
    t = torch()
    d_w = t.zeros(slots * G * B * U, dtype=t.complex64, device="cuda:0")
    m.zf(dev_c64(csi), d_w, B, U, slots * G)
    _sync()
    w_sep = host_c64(d_w).reshape(slots, G, B, U)
    assert np.array_equal(w_f, w_sep)
    llr_s, x_s = gpu_equalize_slots(w_sep, y, U, B, N, sg, Q, lay, slots=slots, S=S)
    assert np.array_equal(llr_f, llr_s)
    assert np.array_equal(x_f, x_s)
    w64, _, _ = oracle.zf(csi[0], U, B, G, "f64")
    for s in range(S):
        llr64, x64 = oracle_symbol_f64(w64, y[0, s], U, B, N, sg, Q, lay)
        check_symbol(llr_f[0, s], x_f[0, s], llr64, x64, N, sg, Q, lay)
        assert abs(synth.evm_db(x_f[0, s], sl[0].x[s]) - synth.evm_db(x64, sl[0].x[s])) < 0.1


def test_detect_slots_failed_group_and_no_w_out():
    print("case: detect failed group, no w_out", flush=True)
    U, B, N, sg, Q, S = 8, 32, 64, 16, 4, 2
    sl = make_slots(3, 1, U, B, N, sg, S, Q)
    csi = sl[0].csi[None].copy()
    csi[0, 2] = 0                              # group 2: no channel -> singular
    lay = synth.standalone_layout(N, Q)
    llr, xh, _, info = _detect(csi, sl[0].y[None], U, B, N, sg, Q, lay, 1, S,
                               want_w=False)
    assert info[0, 2] == 1 and not np.delete(info[0], 2).any()
    assert not xh[0, :, :, 32:48].any()        # W = 0 -> x_hat = 0 for its subcarriers
    z = oracle.demod(np.zeros(1, np.complex64), Q).reshape(-1)
    for u in range(U):
        for k in range(32, 48):
            assert np.array_equal(llr[0, 0, u, k * Q:(k + 1) * Q], z)


def test_detect_slots_requires_w_out_outside_fused_shapes():
    print("case: detect U=32 without w_out", flush=True)
    U, B, N, sg, Q = 32, 64, 48, 16, 4
    sl = make_slots(1, 1, U, B, N, sg, 1, Q)
    lay = synth.standalone_layout(N, Q)
    with pytest.raises(m.MegaError) as e:
        _detect(sl[0].csi[None], sl[0].y[None], U, B, N, sg, Q, lay, 1, 1, want_w=False)
    assert e.value.code == m.MEGA_ERR_INVALID_ARG
