"""The fused uplink FFT stage on cuda:0 (mega_uplink_fft, mega_pilot_fft)
against the reference's FFT test fixtures and the oracle (oracle/fft.py:
numpy's DFT in complex128 around the reference's FP32 load and store). FP32
radix-2 transforms: the bar is the reference tests' 1e-5 (1e-4 for the pilot
product), relative to the spectrum's scale at the larger sizes. Outputs are
prefilled with a sentinel, so entries the reference leaves untouched (users
past U in the CSI, stride padding) are checked too."""
import numpy as np
import pytest

import megastation_amd as m
from oracle import fft as offt

from gpu_helpers import dev_c64, host_c64, torch

pytestmark = pytest.mark.gpu

SENT = np.complex64(7 - 3j)


def _iq(t, a):
    return t.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_uplink_fft_reference_fixture(golden):
    print("case: uplink FFT fixture 8 antennas, 32 bins, 16 subcarriers", flush=True)
    t = torch()
    y = t.zeros(8 * 16, dtype=t.complex64, device="cuda:0")
    m.uplink_fft(_iq(t, golden["fft_uplink_in"]), y, ofdm_start=8, ofdm_ca=32, ofdm_num=16,
                 bs_num=8)
    t.cuda.synchronize()
    ref = golden["fft_uplink_out_kb"].reshape(16, 8).T
    assert np.abs(host_c64(y).reshape(8, 16) - ref).max() < 1e-5


def test_pilot_fft_reference_fixture(golden):
    print("case: pilot FFT fixture 8 antennas, 4 users, groups of 4", flush=True)
    t = torch()
    p = dev_c64(m.pilot_sign(16))
    csi = t.zeros(4 * 8 * 4, dtype=t.complex64, device="cuda:0")
    m.pilot_fft(_iq(t, golden["fft_pilot_in"]), csi, p, ofdm_start=8, ofdm_ca=32, ofdm_num=16,
                bs_num=8, ue_num=4, sc_group=4)
    t.cuda.synchronize()
    assert np.abs(host_c64(csi) - golden["fft_pilot_csi"]).max() < 1e-4


@pytest.mark.parametrize("B,ca,num,start,S,dtype", [
    (64, 2048, 1200, 424, 2, np.int16),   # 64-antenna cell, 2048-point
    (256, 4096, 3276, 410, 3, np.int16),  # 256-antenna cell, 4096-point
    (12, 512, 300, 100, 1, np.int8),      # char samples
    (5, 16, 16, 15, 2, np.int16),         # smallest size, every bin
    (7, 64, 64, 3, 1, np.int16),          # last pass of two stages (6 = 4 + 2)
    (4, 128, 100, 60, 2, np.int16),       # last pass of three stages (7 = 4 + 3)
    (9, 1024, 600, 200, 2, np.int8),      # 10 = 4 + 4 + 2
    (6, 2048, 1200, 424, 2, np.int8),     # compile-time size 2048, char samples
    (3, 4096, 3276, 410, 1, np.int8),     # compile-time size 4096, char samples
])
def test_uplink_fft_matches_oracle(B, ca, num, start, S, dtype):
    print(f"case: uplink FFT B={B} ca={ca} num={num} start={start} S={S} {dtype.__name__}",
          flush=True)
    t = torch()
    rng = np.random.default_rng(ca + B)
    hi = np.iinfo(dtype).max
    iq = rng.integers(-hi // 4, hi // 4, size=(S, B, ca, 2)).astype(dtype)
    pad = 7
    y = t.from_numpy(np.full(S * (B * num + pad), SENT, np.complex64)).to("cuda:0")
    m.uplink_fft(_iq(t, iq), y, ofdm_start=start, ofdm_ca=ca, ofdm_num=num, bs_num=B,
                 num_syms=S, out_sym_stride=B * num + pad)
    t.cuda.synchronize()
    got = host_c64(y).reshape(S, B * num + pad)
    for s in range(S):
        exp = offt.uplink_fft(iq[s], start, ca, num, B)
        scale = float(np.abs(exp).max())
        np.testing.assert_allclose(got[s, :B * num].reshape(B, num), exp, rtol=0,
                                   atol=1e-5 * max(1.0, scale))
        assert (got[s, B * num:] == SENT).all()


@pytest.mark.parametrize("B,ca,num,start,U,sg,ue_start", [
    (64, 2048, 1200, 424, 16, 16, 0),
    (256, 4096, 3276, 410, 32, 12, 0),
    (16, 256, 200, 30, 6, 8, 3),  # users past U skipped, groups of 8, offset
])
def test_pilot_fft_matches_oracle(B, ca, num, start, U, sg, ue_start):
    print(f"case: pilot FFT B={B} ca={ca} num={num} U={U} sg={sg} ue_start={ue_start}",
          flush=True)
    t = torch()
    rng = np.random.default_rng(num + U)
    iq = rng.integers(-8000, 8000, size=(B, ca, 2)).astype(np.int16)
    p = m.pilot_sign(num)
    G = (num + sg - 1) // sg
    csi = t.from_numpy(np.full(G * B * U, SENT, np.complex64)).to("cuda:0")
    m.pilot_fft(_iq(t, iq), csi, dev_c64(p), ofdm_start=start, ofdm_ca=ca, ofdm_num=num,
                bs_num=B, ue_num=U, sc_group=sg, ue_start=ue_start)
    t.cuda.synchronize()
    fill = np.full((G, B, U), SENT, np.complex64)
    exp = offt.pilot_fft(iq, start, ca, num, B, U, sg, p, ue_start=ue_start, out=fill)
    got = host_c64(csi).reshape(G, B, U)
    untouched = exp == SENT
    assert np.array_equal(got[untouched], exp[untouched])
    scale = float(np.abs(exp[~untouched]).max())
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-5 * max(1.0, scale))


def test_downlink_ifft_reference_fixture(golden):
    print("case: downlink IFFT fixture 8 antennas, 32 points", flush=True)
    t = torch()
    iq = t.zeros(8 * 32 * 2, dtype=t.int16, device="cuda:0")
    m.downlink_ifft(dev_c64(golden["ifft_in"]), iq, ofdm_ca=32, bs_num=8)
    t.cuda.synchronize()
    assert np.array_equal(iq.cpu().numpy(), golden["ifft_out"])


@pytest.mark.parametrize("B,ca,S,dtype,amp", [
    (64, 2048, 2, np.int16, 3e-3),   # typical level
    (16, 4096, 1, np.int16, 0.5),    # overdriven: saturation
    (8, 512, 2, np.int8, 5e-3),
])
def test_downlink_ifft_matches_oracle(B, ca, S, dtype, amp):
    """FP32 radix-2 against numpy's complex128 inverse: integers equal except
    where the scaled value lies within 1e-3 of a rounding boundary, there at
    most one apart; saturation exact."""
    print(f"case: downlink IFFT B={B} ca={ca} S={S} {dtype.__name__} amp={amp}", flush=True)
    t = torch()
    rng = np.random.default_rng(ca + S)
    x = ((rng.standard_normal((S, B, ca)) + 1j * rng.standard_normal((S, B, ca))) * amp)
    x = x.astype(np.complex64)
    iq = t.zeros(S * B * ca * 2, dtype=t.int16 if dtype == np.int16 else t.int8, device="cuda:0")
    m.downlink_ifft(dev_c64(x.reshape(-1)), iq, ofdm_ca=ca, bs_num=B, num_syms=S)
    t.cuda.synchronize()
    got = iq.cpu().numpy().reshape(S, B, ca, 2).astype(np.int32)
    info = np.iinfo(dtype)
    for s in range(S):
        exp = offt.downlink_ifft(x[s], ca, B, dtype).astype(np.int32)
        Xs = np.fft.ifft(x[s].astype(np.complex128), axis=1) * ca * ((info.max + 1.0) / ca)
        v = np.stack([Xs.real, Xs.imag], -1)
        near = np.abs(np.abs(v - np.floor(v)) - 0.5) < 1e-3
        d = np.abs(got[s] - exp)
        assert (d <= 1).all() and (d[~near] == 0).all()
        sat = (v > info.max) | (v < info.min)
        assert np.array_equal(got[s][sat], exp[sat])


def test_pilot_fft_symbols_of_a_slot_in_one_launch():
    """num_syms = 3: pilot symbol s carries users s*sg .. s*sg+sg-1 into one
    CSI; equal bit for bit to three one-symbol calls with ue_start = s*sg
    (the reference's calls), users past U skipped."""
    print("case: three pilot symbols in one launch", flush=True)
    t = torch()
    rng = np.random.default_rng(31)
    B, ca, num, start, U, sg, P = 32, 1024, 600, 212, 30, 12, 3
    iq = t.from_numpy(rng.integers(-9000, 9000, size=(P, B, ca, 2)).astype(np.int16)).to("cuda:0")
    p = dev_c64(m.pilot_sign(num))
    G = (num + sg - 1) // sg
    one = t.from_numpy(np.full(G * B * U, SENT, np.complex64)).to("cuda:0")
    sep = one.clone()
    kw = dict(ofdm_start=start, ofdm_ca=ca, ofdm_num=num, bs_num=B, ue_num=U, sc_group=sg)
    m.pilot_fft(iq, one, p, num_syms=P, **kw)
    for s in range(P):
        m.pilot_fft(iq[s], sep, p, ue_start=s * sg, **kw)
    t.cuda.synchronize()
    a, b = host_c64(one), host_c64(sep)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    # users 30..35 of the third symbol are past U: exactly users 0..29 written
    written = (a.reshape(G, B, U) != SENT).any(axis=(0, 1))
    assert written.all()


def test_fft_stage_feeds_detection():
    """The stages compose: a slot built in the frequency domain (i.i.d.
    channel, 16-QAM at 40 dB), taken to int16 time-domain samples, through
    mega_pilot_fft (CSI, A = H^H) -> mega_zf -> mega_uplink_fft (y) ->
    mega_equalize_slots; the receiver's gain cancels in ZF, so x_hat must
    give back the transmitted symbols (every nearest-point decision right,
    EVM of the chain within 0.5 dB of equalizing the frequency-domain slot
    directly)."""
    print("case: pilot FFT -> ZF -> uplink FFT -> equalize", flush=True)
    t = torch()
    from megastation_amd import synth
    B, U, ca, N, start, sg, S, Q = 32, 8, 256, 192, 32, 8, 2, 4
    G = N // sg
    rng = np.random.Generator(np.random.Philox(123))
    sl = synth.make_slot(rng, U, B, N, sg, S, Q, 40.0, "iid")
    p = m.pilot_sign(N)
    cols = (np.arange(N) + start + ca // 2) % ca
    Xp = np.zeros((B, ca), np.complex128)
    for i in range(N):  # subcarrier i carries user i % sg; transmitted 1/conj(p)
        Xp[:, cols[i]] = sl.H[i // sg, :, i % sg] / np.conj(p[i])
    Xd = np.zeros((S, B, ca), np.complex128)
    Xd[:, :, cols] = sl.y
    xt = np.fft.ifft(np.concatenate([Xp[None], Xd]), axis=2)
    gain = 0.4 / np.abs(np.stack([xt.real, xt.imag])).max()  # one receiver gain
    iq = np.round(np.stack([xt.real, xt.imag], -1) * 32768 * gain).astype(np.int16)
    dev = "cuda:0"
    kw = dict(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B)
    csi = t.zeros(G * B * U, dtype=t.complex64, device=dev)
    m.pilot_fft(t.from_numpy(iq[0].copy()).to(dev), csi, dev_c64(p), ue_num=U, sc_group=sg, **kw)
    w = t.empty_like(csi)
    m.zf(csi, w, B, U, G)
    y = t.empty(S * B * N, dtype=t.complex64, device=dev)
    m.uplink_fft(t.from_numpy(iq[1:].copy()).to(dev), y, num_syms=S, **kw)
    lay = synth.standalone_layout(N, Q)
    eq = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=1, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
              y_sym_stride=B * N, y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
              llr_slot_stride=S * U * lay.ue_stride)
    llr = t.empty(S * U * lay.ue_stride, dtype=t.int16, device=dev)
    xh = t.zeros(S * U * N, dtype=t.complex64, device=dev)
    m.equalize_slots(w, y, llr, xhat=xh, **eq)
    # the same slot equalized straight from the frequency domain
    w2 = t.empty_like(csi)
    m.zf(dev_c64(sl.csi.reshape(-1)), w2, B, U, G)
    xh2 = t.zeros_like(xh)
    m.equalize_slots(w2, dev_c64(sl.y.reshape(-1)), t.empty_like(llr), xhat=xh2, **eq)
    t.cuda.synchronize()
    x_chain = host_c64(xh).reshape(S, U, N)
    x_direct = host_c64(xh2).reshape(S, U, N)
    table = synth.qam_table(Q)
    dec = np.abs(x_chain[..., None] - table).argmin(-1)
    ref = np.abs(sl.x[..., None] - table).argmin(-1)
    assert np.array_equal(dec, ref)
    e_chain, e_direct = synth.evm_db(x_chain, sl.x), synth.evm_db(x_direct, sl.x)
    assert e_chain < -25 and abs(e_chain - e_direct) < 0.5, (e_chain, e_direct)
