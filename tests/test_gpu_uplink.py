"""The uplink data path (mega_uplink_*: FFT of symbol chunks into a
two-chunk ring, equalize + demod of each chunk beside the next chunk's FFT)
against the standalone stages it chains, through the C ABI on cuda:0.

Bar: LLRs and x_hat equal mega_uplink_fft of the whole slot followed by
mega_equalize_slots, bit for bit (FP32 form), at 2048 and 4096 points and
for chunk sizes from one symbol to the whole slot; the stages themselves are
held to the FFT and FP64 detection oracles by test_gpu_fft.py and
test_gpu_parity.py. An end-to-end slot built in the frequency domain must
come back as the transmitted symbols (reference data flow:
src/baseband/process.cc:175-186)."""
import numpy as np
import pytest

import megastation_amd as m
from megastation_amd import synth

from gpu_helpers import SENTINEL, dev_c64, host_c64, host_u16, llr_buffer, torch

pytestmark = pytest.mark.gpu


def _samples(rng, S, B, ca):
    t = torch()
    iq = rng.integers(-20000, 20000, size=(S, B, ca, 2), dtype=np.int16)
    return t.from_numpy(iq).to("cuda:0")


def _weights(rng, U, B, G):
    t = torch()
    H = synth.channel_iid(rng, G, B, U)
    a = dev_c64(np.conj(H).reshape(-1))
    w = t.empty_like(a)
    m.zf(a, w, B, U, G)
    return a, w


def _layout(kind, N, Q):
    return synth.ldpc_uplink_layout(N, Q) if kind == "ldpc" else synth.standalone_layout(N, Q)


@pytest.mark.parametrize("U,B,ca,N,start,sg,S,Q,lay_kind,chunks", [
    (32, 256, 4096, 3276, 410, 12, 13, 6, "standalone", (0, 1, 4, 13)),
    (16, 64, 2048, 1200, 424, 16, 13, 4, "ldpc", (0, 3)),
    (8, 24, 512, 300, 100, 12, 5, 2, "standalone", (2,)),
])
def test_uplink_equals_fft_then_equalize(U, B, ca, N, start, sg, S, Q, lay_kind, chunks):
    t = torch()
    G = (N + sg - 1) // sg
    rng = np.random.default_rng(U * 100 + B + ca)
    iq = _samples(rng, S, B, ca)
    _, w = _weights(rng, U, B, G)
    lay = _layout(lay_kind, N, Q)
    y = t.empty(S * B * N, dtype=t.complex64, device="cuda:0")
    m.uplink_fft(iq, y, ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S)
    eq = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=1, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
              y_sym_stride=B * N, y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
              llr_slot_stride=S * U * lay.ue_stride)
    llr_ref = llr_buffer(S * U * lay.ue_stride)
    x_ref = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
    m.equalize_slots(w, y, llr_ref, xhat=x_ref, **eq)
    t.cuda.synchronize()
    ref_l, ref_x = host_u16(llr_ref), host_c64(x_ref)
    for K in chunks:
        print(f"case: uplink U={U} B={B} ca={ca} N={N} S={S} chunk={K}", flush=True)
        up = m.UplinkDetect(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S,
                            num_ue=U, sc_group=sg, mod_order=Q, ue_stride=lay.ue_stride,
                            cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                            valid_bits=lay.valid_bits, chunk_syms=K)
        llr = llr_buffer(S * U * lay.ue_stride)
        xh = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
        for _ in range(2):                      # a second run reuses the ring
            up.run(iq, w, llr, xh)
        t.cuda.synchronize()
        assert np.array_equal(host_u16(llr), ref_l)
        assert np.array_equal(host_c64(xh).view(np.uint32), ref_x.view(np.uint32))
        up.close()
    assert not (ref_l == SENTINEL).all()


def test_uplink_mixed_form_matches_stages():
    """The mixed form's equalize from mega_zf_split's planes, chunked: equal to
    the standalone mixed equalize of the FFT'd slot bit for bit."""
    print("case: uplink mixed 256x32", flush=True)
    t = torch()
    U, B, ca, N, start, sg, S, Q = 32, 256, 4096, 3276, 410, 12, 13, 6
    G = (N + sg - 1) // sg
    Bp = (B + 15) // 16 * 16
    rng = np.random.default_rng(5)
    iq = _samples(rng, S, B, ca)
    a, _ = _weights(rng, U, B, G)
    split = t.empty(G * U * Bp, dtype=t.int64, device="cuda:0")
    m.zf_split(a, split, B, U, G)
    lay = synth.standalone_layout(N, Q)
    y = t.empty(S * B * N, dtype=t.complex64, device="cuda:0")
    m.uplink_fft(iq, y, ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S)
    llr_ref = llr_buffer(S * U * lay.ue_stride)
    m.equalize_slots(split, y, llr_ref, num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S,
                     num_slots=1, mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
                     cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
                     csi_slot_stride=G * U * Bp, y_sym_stride=B * N, y_slot_stride=S * B * N,
                     llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
                     form=m.FORM_MIXED)
    up = m.UplinkDetect(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S,
                        num_ue=U, sc_group=sg, mod_order=Q, ue_stride=lay.ue_stride,
                        cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                        valid_bits=lay.valid_bits, form=m.FORM_MIXED)
    llr = llr_buffer(S * U * lay.ue_stride)
    up.run(iq, split, llr)
    t.cuda.synchronize()
    assert np.array_equal(host_u16(llr), host_u16(llr_ref))
    up.close()


def test_uplink_recovers_transmitted_symbols():
    """A slot built in the frequency domain (i.i.d. channel, 16-QAM at 40 dB),
    taken to int16 samples; CSI from the channel, ZF, then the uplink data
    path: every nearest-point decision equals the transmitted symbol."""
    print("case: uplink end to end", flush=True)
    t = torch()
    B, U, ca, N, start, sg, S, Q = 32, 8, 256, 192, 32, 8, 4, 4
    G = N // sg
    rng = np.random.Generator(np.random.Philox(321))
    sl = synth.make_slot(rng, U, B, N, sg, S, Q, 40.0, "iid")
    cols = (np.arange(N) + start + ca // 2) % ca
    Xd = np.zeros((S, B, ca), np.complex128)
    Xd[:, :, cols] = sl.y
    xt = np.fft.ifft(Xd, axis=2)
    gain = 0.4 / np.abs(np.stack([xt.real, xt.imag])).max()
    iq = np.round(np.stack([xt.real, xt.imag], -1) * 32768 * gain).astype(np.int16)
    # the FFT gives y * gain (samples scaled by 1/32768; the unnormalised DFT
    # undoes numpy's normalised inverse): the channel estimates carry the same
    # scale, which ZF then cancels
    a = dev_c64(sl.csi.reshape(-1) * np.float32(gain))
    w = t.empty_like(a)
    m.zf(a, w, B, U, G)
    lay = synth.standalone_layout(N, Q)
    up = m.UplinkDetect(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S,
                        num_ue=U, sc_group=sg, mod_order=Q, ue_stride=lay.ue_stride,
                        cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                        valid_bits=lay.valid_bits, chunk_syms=2)
    llr = llr_buffer(S * U * lay.ue_stride)
    xh = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
    up.run(t.from_numpy(iq).to("cuda:0"), w, llr, xh)
    t.cuda.synchronize()
    x = host_c64(xh).reshape(S, U, N)
    table = synth.qam_table(Q)
    assert np.array_equal(np.abs(x[..., None] - table).argmin(-1),
                          np.abs(sl.x[..., None] - table).argmin(-1))
    up.close()


def test_uplink_rejects_bad_shapes():
    print("case: uplink argument checks", flush=True)
    lay = synth.standalone_layout(300, 4)
    kw = dict(ofdm_start=0, ofdm_ca=512, ofdm_num=300, bs_num=16, num_syms=2, num_ue=4,
              sc_group=12, mod_order=4, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
              cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits)
    m.UplinkDetect(**kw).close()
    for bad, code in ((dict(ofdm_ca=500), m.MEGA_ERR_UNSUPPORTED),
                      (dict(ofdm_num=600), m.MEGA_ERR_UNSUPPORTED),
                      (dict(quant_bits=12), m.MEGA_ERR_UNSUPPORTED),
                      (dict(mod_order=5), m.MEGA_ERR_UNSUPPORTED)):
        with pytest.raises(m.MegaError) as e:
            m.UplinkDetect(**dict(kw, **bad))
        assert e.value.code == code, bad
