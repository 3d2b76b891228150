"""The LDPC oracle (oracle/ldpc.c, oracle/ldpc.py) pinned before it judges the
HIP decoder: the base-graph tables against the reference's BG1 codewords and
the standard's structure, the encoder against those codewords, the decoder
against the reference's decoder test (test/ldpc/test_ldpc.cc:155-190), a
brute-force tiny code and the min-sum update's closed form."""
import ctypes
import itertools

import numpy as np
import pytest

from oracle import _p, ldpc

ZC, MB = 64, 46   # the reference test's code: BG1, Zc = 64, 46 parity rows


def _fixture_bits(golden):
    inp = np.unpackbits(golden["ldpc_input_bits"], bitorder="little").reshape(2, 22 * ZC)
    enc = np.unpackbits(golden["ldpc_encoded_bits"], bitorder="little").reshape(2, 66 * ZC)
    return inp, enc


def test_tables_have_the_standard_structure():
    for bg, rows, cols, kb, nnz in ((1, 46, 68, 22, 316), (2, 42, 52, 10, 197)):
        rs, col, sh = ldpc.tables(bg)
        assert len(rs) == rows + 1 and rs[-1] == nnz and sh.shape == (8, nnz)
        assert col.min() >= 0 and col.max() < cols
        for r in range(rows):
            c = col[rs[r]:rs[r + 1]].astype(int)
            assert np.all(np.diff(c) > 0)                      # sorted, no repeats
            if r >= 4:                                          # extension: identity
                assert (kb + r) in c and c[-1] == kb + r
                e = rs[r] + list(c).index(kb + r)
                assert np.all(sh[:, e] == 0)
    assert ldpc.set_index(64) == 0 and ldpc.set_index(288) == 4 and ldpc.set_index(52) == 6
    with pytest.raises(ValueError):
        ldpc.set_index(100)


def test_reference_codewords_satisfy_every_check(golden):
    inp, enc = _fixture_bits(golden)
    for i in range(2):
        cw = np.concatenate([inp[i][:2 * ZC], enc[i]])        # punctured bits restored
        assert np.array_equal(cw[:22 * ZC], inp[i])           # systematic
        assert ldpc.parity_syndrome(cw, 1, ZC, MB) == 0
        bad = cw.copy()
        bad[1000] ^= 1
        assert ldpc.parity_syndrome(bad, 1, ZC, MB) > 0


def test_encoder_reproduces_reference_codewords(golden):
    inp, enc = _fixture_bits(golden)
    for i in range(2):
        assert np.array_equal(ldpc.encode(inp[i], 1, ZC, MB)[2 * ZC:], enc[i])


def test_noiseless_decode_gives_reference_information_bits(golden):
    """test_ldpc.cc:155-190: LLR +-127 of the encoded bits after a zero
    prefix of 2 * Zc, 8 iterations: the decoded bytes equal input.data."""
    inp, enc = _fixture_bits(golden)
    llr = np.zeros((2, 68 * ZC), np.float32)
    llr[:, 2 * ZC:] = np.where(enc == 1, -127.0, 127.0)
    out = ldpc.decode(llr, 1, ZC, MB, 8, 0.8125)
    got = out.view(np.uint8).reshape(2, -1)[:, :22 * ZC // 8]
    assert np.array_equal(got.reshape(-1), golden["ldpc_input_bits"])


def _raw_decode(kb, mb, zc, rs, col, sh, iters, alpha, llr):
    L = ldpc.lib()
    f = L.orc_ldpc_decode
    n, nv = llr.shape
    words = (kb * zc + 31) // 32
    out = np.zeros((n, words), np.uint32)
    app = np.zeros((n, nv), np.float32)
    rs, col, sh = (np.ascontiguousarray(a) for a in (rs, col, sh))
    llr = np.ascontiguousarray(llr, np.float32)
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 3 + [
        ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
        ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    f.restype = ctypes.c_int
    assert f(kb, mb, zc, _p(rs), _p(col), _p(sh), iters, alpha, _p(llr), nv, n, _p(out), words,
             _p(app)) == 0
    return ldpc.unpack_bits(out, kb * zc), app


def test_min_sum_update_closed_form():
    """One check of degree 4 (one block row, Zc = 1), one iteration: each
    variable's LLR becomes L + alpha * (product of the other signs) * (smallest
    other magnitude)."""
    rs = np.array([0, 4], np.int16)
    col = np.array([0, 1, 2, 3], np.int8)
    sh = np.zeros(4, np.int16)
    rng = np.random.default_rng(5)
    for _ in range(50):
        l0 = rng.standard_normal(4).astype(np.float32) * 3
        _, app = _raw_decode(3, 1, 1, rs, col, sh, 1, 0.75, l0[None])
        for e in range(4):
            others = np.delete(l0.astype(np.float64), e)
            ext = 0.75 * np.prod(np.sign(others)) * np.abs(others).min()
            assert abs(app[0, e] - (l0[e] + ext)) <= 1e-6 * (abs(l0[e]) + abs(ext))


def test_brute_force_tiny_code():
    """A 2 x 4 base graph lifted by Zc = 3 (12 bits): every codeword found by
    enumerating all 2^12 words decodes to itself from noiseless LLRs, and a
    weak single wrong LLR is corrected."""
    rs = np.array([0, 3, 6], np.int16)
    col = np.array([0, 1, 2, 0, 1, 3], np.int8)
    sh = np.array([0, 1, 0, 2, 0, 0], np.int16)
    zc, kb, mb = 3, 2, 2

    def ok(cw):
        for r in range(mb):
            acc = np.zeros(zc, np.uint8)
            for e in range(rs[r], rs[r + 1]):
                acc ^= np.roll(cw[col[e] * zc:(col[e] + 1) * zc], -sh[e])
            if acc.any():
                return False
        return True
    words = [np.array(b, np.uint8) for b in itertools.product((0, 1), repeat=12)]
    code = [w for w in words if ok(w)]
    assert len(code) == 2 ** (kb * zc)               # full-rank parity part
    llr = np.stack([np.where(w == 1, -4.0, 4.0) for w in code]).astype(np.float32)
    dec, _ = _raw_decode(kb, mb, zc, rs, col, sh, 5, 0.8125, llr)
    assert np.array_equal(dec, np.stack([w[:kb * zc] for w in code]))
    flipped = llr.copy()
    flipped[:, 1] *= -0.25                             # one weak, wrong information LLR
    dec, _ = _raw_decode(kb, mb, zc, rs, col, sh, 5, 0.8125, flipped)
    assert np.array_equal(dec, np.stack([w[:kb * zc] for w in code]))


@pytest.mark.parametrize("bg,zc,mb", [(1, 64, 46), (2, 52, 42), (1, 288, 46), (2, 384, 20)])
def test_decoding_corrects_channel_errors(bg, zc, mb):
    """Codewords of the encoder through BPSK + AWGN: after 8 iterations
    fewer bit errors than hard decisions of the channel LLRs, none at a
    moderate SNR."""
    rng = np.random.default_rng(bg * 1000 + zc)
    kb = ldpc.KB[bg]
    info = rng.integers(0, 2, (4, kb * zc)).astype(np.uint8)
    cws = np.stack([ldpc.encode(b, bg, zc, mb) for b in info])
    for snr_db, want_clean in ((-1.0, False), (2.0, True)):
        sigma = np.sqrt(0.5 / 10 ** (snr_db / 10))
        rx = (1 - 2.0 * cws) + sigma * rng.standard_normal(cws.shape)
        llr = (2 * rx / sigma ** 2).astype(np.float32)
        llr[:, :2 * zc] = 0
        dec = ldpc.unpack_bits(ldpc.decode(llr, bg, zc, mb, 8, 0.8125), kb * zc)
        err_dec = int((dec != info).sum())
        err_ch = int(((llr[:, 2 * zc:kb * zc] < 0) != info[:, 2 * zc:]).sum())
        assert err_dec < err_ch
        if want_clean:
            assert err_dec == 0
