"""The LDPC consumer (mega_ldpc_decode) on cuda:0: the reference's decoder
test (test/ldpc/test_ldpc.cc:155-190: noiseless LLRs of encode.data decode to
input.data, one code block and two), bit-identical to the oracle's layered
min-sum on noisy code blocks of both base graphs (binary16 and float LLRs,
padded strides), and the chain the decoder exists for: LLRs of the
equalizer's LDPC layout (zeros = 2 * Zc, cb_size = (22 - 2 + mb) * Zc,
128-byte padded rows) decoded back into the transmitted information bits."""
import numpy as np
import pytest

import megastation_amd as m
from megastation_amd import synth
from oracle import ldpc

from gpu_helpers import dev_c64, make_slots, torch

pytestmark = pytest.mark.gpu


def _decode_gpu(llr_np, bg, zc, mb, iters, half, n_cb, llr_stride, out_stride, alpha=0.0):
    t = torch()
    dt = t.float16 if half else t.float32
    llr = t.from_numpy(llr_np.astype(np.float16 if half else np.float32).reshape(-1)).to("cuda:0")
    out = t.full((n_cb * out_stride,), -1, dtype=t.int32, device="cuda:0")
    m.ldpc_decode(llr.to(dt), out, base_graph=bg, lifting_size=zc, parity_nodes=mb,
                  max_iterations=iters, num_cb=n_cb, llr_stride=llr_stride,
                  out_stride=out_stride, normalization=alpha)
    t.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32).reshape(n_cb, out_stride)


@pytest.mark.parametrize("half", [False, True])
def test_reference_codewords_decode_to_input_data(golden, half):
    print(f"case: ldpc reference fixture half={half}", flush=True)
    zc, mb = 64, 46
    enc = np.unpackbits(golden["ldpc_encoded_bits"], bitorder="little").reshape(2, 66 * zc)
    llr = np.zeros((2, 68 * zc), np.float32)
    llr[:, 2 * zc:] = np.where(enc == 1, -127.0, 127.0)
    words = 22 * zc // 32
    for n in (1, 2):                                # LDPCDecoderVec and LDPCDecoderArr
        out = _decode_gpu(llr[:n], 1, zc, mb, 8, half, n, 68 * zc, words)
        got = out.view(np.uint8).reshape(-1)
        assert np.array_equal(got, golden["ldpc_input_bits"][:n * 176])


@pytest.mark.parametrize("bg,zc,mb,half", [(1, 64, 46, False), (1, 288, 46, True),
                                           (2, 52, 42, False), (2, 384, 12, True),
                                           (1, 15, 4, False)])
def test_decoder_matches_oracle_on_noisy_code_blocks(bg, zc, mb, half):
    """Noise at a level the decoder only partly corrects in 4 iterations, so
    the comparison covers messages of both signs and unfinished decodes."""
    print(f"case: ldpc vs oracle bg={bg} zc={zc} mb={mb} half={half}", flush=True)
    kb = ldpc.KB[bg]
    n_cb = 5
    rng = np.random.default_rng(zc * 7 + mb)
    info = rng.integers(0, 2, (n_cb, kb * zc)).astype(np.uint8)
    cws = np.stack([ldpc.encode(b, bg, zc, mb) for b in info])
    sigma = 1.0
    llr = ((2 * ((1 - 2.0 * cws) + sigma * rng.standard_normal(cws.shape))) / sigma ** 2)
    llr[:, :2 * zc] = 0
    if half:
        llr = llr.astype(np.float16)
    llr = llr.astype(np.float32)                    # exactly what the kernel reads
    nv = (kb + mb) * zc
    stride = synth.pad64(nv)                        # 128-byte rows, like the equalizer's
    buf = np.zeros((n_cb, stride), np.float32)
    buf[:, :nv] = llr
    words = (kb * zc + 31) // 32
    out = _decode_gpu(buf, bg, zc, mb, 4, half, n_cb, stride, words + 3)
    ref = ldpc.decode(llr, bg, zc, mb, 4, 0.8125)
    assert np.array_equal(out[:, :words], ref)
    assert np.all(out[:, words:] == 0xFFFFFFFF)     # nothing written past the bits
    errs = (ldpc.unpack_bits(ref, kb * zc) != info).sum()
    print(f"  residual bit errors after 4 iterations: {errs}", flush=True)


def test_equalizer_llrs_decode_to_transmitted_bits():
    """64 x 16 cell, 1200 subcarriers, 16-QAM at -4 dB per antenna (about
    13 dB after ZF's array gain; the hard decisions of about 2 % of the coded
    bits are wrong): each user's symbol carries one rate-1/3 BG1 code block
    (Zc = 72: 22 * 72 <= 1200 * 4 / 3); ZF + equalize write its LLRs into the
    decoder's layout; the decoder gives back every information bit."""
    print("case: ldpc chain equalize -> decode", flush=True)
    t = torch()
    U, B, N, sg, Q, S, snr = 16, 64, 1200, 16, 4, 2, -4.0
    G = N // sg
    lay = synth.ldpc_uplink_layout(N, Q)
    zc = lay.zeros // 2
    mb = lay.cb_size // zc - 20
    assert (zc, mb) == (72, 46)
    rng = np.random.default_rng(11)
    info = rng.integers(0, 2, (S, U, 22 * zc)).astype(np.uint8)
    bits = rng.integers(0, 2, (S, U, N * Q)).astype(np.uint8)   # filler past cb_size
    for s in range(S):
        for u in range(U):
            bits[s, u, :lay.cb_size] = ldpc.encode(info[s, u], 1, zc, mb)[2 * zc:]
    sl = make_slots(5, 1, U, B, N, sg, S, Q, snr)[0]
    # the same channel and noise draw as make_slots, with these bits
    x = synth.modulate(bits.reshape(S, U, N, Q)).astype(np.complex64)
    Hk = sl.H[np.arange(N) // sg]
    y = np.einsum("kbu,suk->sbk", Hk, x).astype(np.complex64)
    y += (sl.y - np.einsum("kbu,suk->sbk", Hk, sl.x)).astype(np.complex64)   # its noise
    w = t.zeros(G * B * U, dtype=t.complex64, device="cuda:0")
    m.zf(dev_c64(sl.csi), w, B, U, G)
    llr = t.zeros(S * U * lay.ue_stride, dtype=t.int16, device="cuda:0")   # zero prefix
    m.equalize_slots(w, dev_c64(y), llr, num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                     num_syms=S, num_slots=1, mod_order=Q, ue_stride=lay.ue_stride,
                     cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                     valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                     y_sym_stride=B * N, y_slot_stride=S * B * N,
                     llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride)
    words = 22 * zc // 32 + (1 if (22 * zc) % 32 else 0)
    out = t.zeros(S * U * words, dtype=t.int32, device="cuda:0")
    m.ldpc_decode(llr.view(t.float16), out, base_graph=1, lifting_size=zc, parity_nodes=mb,
                  max_iterations=8, num_cb=S * U, llr_stride=lay.cb_stride, out_stride=words)
    t.cuda.synchronize()
    dec = ldpc.unpack_bits(out.cpu().numpy().view(np.uint32).reshape(S * U, words), 22 * zc)
    raw = np.array(llr.cpu().numpy().view(np.float16), dtype=np.float32).reshape(
        S * U, lay.cb_stride)[:, lay.zeros:lay.zeros + lay.cb_size]
    raw_err = int(((raw < 0) != bits.reshape(S * U, -1)[:, :lay.cb_size]).sum())
    print(f"  channel bit errors before decoding: {raw_err}", flush=True)
    assert raw_err > 0                               # the decoder has work to do
    assert np.array_equal(dec, info.reshape(S * U, -1))


def test_ldpc_rejects_bad_arguments():
    print("case: ldpc argument checks", flush=True)
    t = torch()
    llr = t.zeros(68 * 64, dtype=t.float32, device="cuda:0")
    out = t.zeros(64, dtype=t.int32, device="cuda:0")
    kw = dict(base_graph=1, lifting_size=64, parity_nodes=46, max_iterations=8, num_cb=1,
              llr_stride=68 * 64, out_stride=44)
    m.ldpc_decode(llr, out, **kw)
    for bad, code in ((dict(base_graph=3), m.MEGA_ERR_UNSUPPORTED),
                      (dict(lifting_size=100), m.MEGA_ERR_UNSUPPORTED),
                      (dict(parity_nodes=47), m.MEGA_ERR_UNSUPPORTED),
                      (dict(max_iterations=0), m.MEGA_ERR_UNSUPPORTED),
                      (dict(llr_stride=68 * 64 - 1), m.MEGA_ERR_INVALID_ARG),
                      (dict(out_stride=43), m.MEGA_ERR_INVALID_ARG)):
        with pytest.raises(m.MegaError) as e:
            m.ldpc_decode(llr, out, **dict(kw, **bad))
        assert e.value.code == code, bad
