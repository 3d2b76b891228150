"""Downlink precode on cuda:0 (mega_precode / mega_precode_syms) against the
oracle (orc_precode, pinned to the reference's test/sgemv/precoded.data) and
the reference fixture itself. FP32 products: the bar is the reference test's
1e-5 (relative to the output scale for the large shapes). Output buffers are
prefilled with a sentinel, so the entries the reference leaves untouched
(guard columns, a partial last group) are checked too."""
import numpy as np
import pytest

import megastation_amd as m
import oracle

from gpu_helpers import dev_c64, host_c64, torch

pytestmark = pytest.mark.gpu

SENT = np.complex64(7 - 3j)


def test_precode_reference_fixture(golden):
    print("case: precode reference fixture 8x4, 16 subcarriers, row 32", flush=True)
    t = torch()
    prec = dev_c64(golden["precode_prec"])
    ue = dev_c64(golden["precode_uevec_ku"].reshape(16, 4).T.copy().reshape(-1))
    out = t.zeros(8 * 32, dtype=t.complex64, device="cuda:0")
    rc = m.lib().mega_precode(8, 4, 16, m._ptr(prec), m._ptr(ue), m._ptr(out), 32, 16, 8,
                              m._stream(None))
    assert rc == 0, m.last_error()
    t.cuda.synchronize()
    got = host_c64(out).reshape(8, 32)
    assert np.abs(got - golden["precode_out"].reshape(8, 32)).max() < 1e-5


@pytest.mark.parametrize("B,U,N,sg,row,start,S", [
    (64, 16, 1200, 16, 2048, 424, 1),    # the reference's PrecodeLarge shape
    (256, 32, 3276, 12, 4096, 410, 3),   # 256x32 cell, 273 groups, three symbols
    (24, 5, 100, 12, 128, 17, 2),        # partial last group left untouched
    (130, 64, 200, 40, 256, 0, 1),       # sg > 32 (column chunks), 64 users
    (8, 3, 7, 8, 16, 3, 1),              # no full group: nothing written
])
def test_precode_matches_oracle(B, U, N, sg, row, start, S):
    print(f"case: precode B={B} U={U} N={N} sg={sg} row={row} start={start} S={S}", flush=True)
    t = torch()
    rng = np.random.default_rng(B * 7 + U + N)
    G = max(1, N // sg)
    P = (rng.standard_normal((G, U, B)) + 1j * rng.standard_normal((G, U, B))).astype(np.complex64)
    ue_pad, out_pad = U * N + 5, B * row + 3  # padded symbol strides
    ue = np.zeros((S, ue_pad), np.complex64)
    ue[:, :U * N] = (rng.standard_normal((S, U * N)) + 1j * rng.standard_normal((S, U * N)))
    out = t.from_numpy(np.full(S * out_pad, SENT, np.complex64)).to("cuda:0")
    m.precode(dev_c64(P.reshape(-1)), dev_c64(ue.reshape(-1)), out, num_bs=B, num_ue=U,
              num_sc=N, sc_group=sg, bs_stride=row, bs_start=start, num_syms=S,
              ue_sym_stride=ue_pad, bs_sym_stride=out_pad)
    t.cuda.synchronize()
    got = host_c64(out).reshape(S, out_pad)
    for s in range(S):
        fill = np.full((B, row), SENT, np.complex64)
        exp = oracle.precode(P, ue[s, :U * N].reshape(U, N), B, U, N, sg, row, start, out=fill)
        g = got[s, :B * row].reshape(B, row)
        untouched = exp == SENT
        assert np.array_equal(g[untouched], exp[untouched])
        scale = max(1.0, float(np.abs(exp[~untouched]).max(initial=0.0)))
        np.testing.assert_allclose(g, exp, rtol=1e-5, atol=1e-5 * scale)
        assert (got[s, B * row:] == SENT).all()  # stride padding untouched
    if N < sg:
        assert (got == SENT).all()

