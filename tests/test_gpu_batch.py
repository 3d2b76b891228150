"""The batching form (mega_batch_*): uneven cells planned, autotuned and
captured into one hipGraph, checked cell by cell against the single-cell
entry points and the FP64 oracle on cuda:0.

Bars: W bit-identical to mega_zf of the same cell (same group kernel code);
symbols by the parity policy of the FP32 form (x_hat within 1e-4 of the FP64
oracle relative to the symbol's largest |x_hat|, LLRs bit-exact against the
demod of the kernel's own x_hat, hard bits equal away from the boundary);
graph replays bit-identical to the first run."""
import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth

from gpu_helpers import (SENTINEL, check_symbol, dev_c64, host_c64, host_u16,
                         oracle_llr_from_xhat,
                         llr_buffer, make_slots, oracle_symbol_f64, torch)

pytestmark = pytest.mark.gpu

# U, B, N, sg, Q, layout: every antenna x user size of the uneven-batch
# workload, group sizes 12 and 16, partial last groups, both LLR layouts
CELLS = [
    (8, 64, 624, 12, 4, "standalone"),
    (16, 64, 300, 16, 6, "ldpc"),
    (16, 128, 240, 16, 8, "standalone"),
    (32, 128, 204, 12, 6, "ldpc"),
    (64, 256, 96, 12, 4, "standalone"),
    (8, 64, 100, 16, 8, "standalone"),      # 6.25 groups
    (32, 256, 150, 16, 2, "standalone"),    # 9.4 groups
]


def _make_cells(seed=3, S=2):
    t = torch()
    out = []
    for k, (U, B, N, sg, Q, layout) in enumerate(CELLS):
        sl = make_slots(seed * 100 + k, 1, U, B, N, sg, S, Q, 20.0)[0]
        lay = synth.ldpc_uplink_layout(N, Q) if layout == "ldpc" else synth.standalone_layout(N, Q)
        G = (N + sg - 1) // sg
        out.append(dict(U=U, B=B, N=N, sg=sg, S=S, Q=Q, layout=lay, slot=sl,
                        csi=dev_c64(sl.csi), w=t.zeros(G * B * U, dtype=t.complex64, device="cuda:0"),
                        y=dev_c64(sl.y), llr=llr_buffer(S * U * lay.ue_stride),
                        xhat=t.zeros(S * U * N, dtype=t.complex64, device="cuda:0"),
                        info=t.full((G,), -3, dtype=t.int32, device="cuda:0")))
    return out


def test_batch_matches_single_cell_paths_and_oracle(tmp_path):
    print("case: batch of 7 uneven cells", flush=True)
    t = torch()
    cells = _make_cells()
    cache = str(tmp_path / "tune.txt")
    b = m.Batch(cells, tune_path=cache)
    plan = b.plan()
    print(plan, flush=True)
    b.run()
    t.cuda.synchronize()
    first = [host_u16(c["llr"]).copy() for c in cells]
    for c in cells:
        U, B, N, sg, Q, S, lay = c["U"], c["B"], c["N"], c["sg"], c["Q"], c["S"], c["layout"]
        G = (N + sg - 1) // sg
        assert not c["info"].cpu().numpy().any()
        w1 = t.zeros_like(c["w"])
        m.zf(c["csi"], w1, B, U, G)
        t.cuda.synchronize()
        assert np.array_equal(host_c64(c["w"]), host_c64(w1)), (U, B)
        llr = host_u16(c["llr"]).reshape(S, U, lay.ue_stride)
        xh = host_c64(c["xhat"]).reshape(S, U, N)
        w64, _, _ = oracle.zf(c["slot"].csi, U, B, G, "f64")
        for s in range(S):
            llr64, x64 = oracle_symbol_f64(w64, c["slot"].y[s], U, B, N, sg, Q, lay)
            check_symbol(llr[s], xh[s], llr64, x64, N, sg, Q, lay)
    # replays (graph nodes re-run, counters reset by a kernel) give the same bits
    for _ in range(3):
        for c in cells:
            c["llr"].fill_(SENTINEL)
        b.run()
        t.cuda.synchronize()
        for c, f in zip(cells, first):
            assert np.array_equal(host_u16(c["llr"]), f)
    b.close()
    # a second plan of the same batch takes its tile plans from the cache
    b2 = m.Batch(cells, tune_path=cache)
    assert "(cached)" in b2.plan()
    for c in cells:
        c["llr"].fill_(SENTINEL)
    b2.run()
    t.cuda.synchronize()
    for c, f in zip(cells, first):   # same tile plans -> same arithmetic
        assert np.array_equal(host_u16(c["llr"]), f)
    b2.close()


def test_batch_rejects_bad_cells():
    print("case: batch argument checks", flush=True)
    cells = _make_cells()[:2]
    bad = dict(cells[0], Q=5)
    with pytest.raises(m.MegaError) as e:
        m.Batch([cells[1], bad])
    assert e.value.code == m.MEGA_ERR_UNSUPPORTED
    bad = dict(cells[0], U=65)
    with pytest.raises(m.MegaError) as e:
        m.Batch([bad])
    assert e.value.code == m.MEGA_ERR_UNSUPPORTED


def _gpu_cell(cfg, seed):
    """Inputs of one cell generated on the GPU: A = H^H [G][B][U] (i.i.d.
    CN(0,1)), y = H x + n [S][B][N] at the cell's SNR, x from the QAM table."""
    t = torch()
    U, B, N, sg, S, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "Q"))
    G = (N + sg - 1) // sg
    gen = t.Generator(device="cuda:0")
    gen.manual_seed(seed)
    H = t.randn((G, B, U), dtype=t.complex64, device="cuda:0", generator=gen)
    table = t.from_numpy(synth.qam_table(Q)).to("cuda:0")
    x = table[t.randint(0, 1 << Q, (S, U, G * sg), device="cuda:0", generator=gen)]
    X = x.reshape(S, U, G, sg).permute(2, 1, 0, 3).reshape(G, U, S * sg).contiguous()
    Hr, Hi, Xr, Xi = (v.contiguous() for v in (H.real, H.imag, X.real, X.imag))
    Y = t.complex(t.bmm(Hr, Xr) - t.bmm(Hi, Xi), t.bmm(Hr, Xi) + t.bmm(Hi, Xr))
    Y = Y + (10 ** (-cfg["snr"] / 20)) * t.randn(Y.shape, dtype=t.complex64, device="cuda:0",
                                                  generator=gen)
    y = Y.reshape(G, B, S, sg).permute(2, 1, 0, 3).reshape(S, B, G * sg)[..., :N].contiguous()
    csi = t.conj(H).resolve_conj().contiguous()
    return csi, y, x[..., :N].contiguous()


def test_batch_40_uneven_cells_full_size():
    """The uneven 40-cell workload at its full sizes (13 symbols, N up to
    3276, 256 x 64 cells, QAM 16/64/256, partial last groups): every cell's W
    bit-identical to mega_zf; every symbol's LLRs bit-identical to the
    oracle's demod and placement of the kernel's own x_hat; symbols 0 and 12
    of every cell within the FP64 bar (x_hat within 1e-4 of FP64 zero-forcing
    relative to the symbol's largest |x_hat|, EVM within 0.1 dB)."""
    t = torch()
    cfgs = synth.uneven_cells(40, 7)
    print(f"case: batch of 40 uneven cells, full size", flush=True)
    cells = []
    for k, cfg in enumerate(cfgs):
        U, B, N, sg, S, Q = (cfg[x] for x in ("U", "B", "N", "sg", "S", "Q"))
        G = (N + sg - 1) // sg
        csi, y, x = _gpu_cell(cfg, 1000 + k)
        lay = synth.standalone_layout(N, Q)
        cells.append(dict(U=U, B=B, N=N, sg=sg, S=S, Q=Q, layout=lay, csi=csi, y=y, x=x,
                          w=t.zeros(G * B * U, dtype=t.complex64, device="cuda:0"),
                          llr=llr_buffer(S * U * lay.ue_stride),
                          xhat=t.zeros(S * U * N, dtype=t.complex64, device="cuda:0"),
                          info=t.full((G,), -3, dtype=t.int32, device="cuda:0")))
    b = m.Batch(cells)
    print(b.plan(), flush=True)
    b.run()
    t.cuda.synchronize()
    for k, c in enumerate(cells):
        U, B, N, sg, Q, S, lay = c["U"], c["B"], c["N"], c["sg"], c["Q"], c["S"], c["layout"]
        G = (N + sg - 1) // sg
        print(f"  cell {k}: {B}x{U} N={N} sg={sg} Q={Q}", flush=True)
        assert not c["info"].cpu().numpy().any()
        w1 = t.zeros_like(c["w"])
        m.zf(c["csi"], w1, B, U, G)
        t.cuda.synchronize()
        assert np.array_equal(host_c64(c["w"]), host_c64(w1)), k
        llr = host_u16(c["llr"]).reshape(S, U, lay.ue_stride)
        xh = host_c64(c["xhat"]).reshape(S, U, N)
        for s in range(S):
            assert np.array_equal(llr[s], oracle_llr_from_xhat(xh[s], N, sg, Q, lay)), (k, s)
        # FP64 zero-forcing (numpy complex128) of symbols 0 and S-1
        A = host_c64(c["csi"]).astype(np.complex128).reshape(G, B, U).transpose(0, 2, 1)
        W64 = np.linalg.solve(A @ A.conj().transpose(0, 2, 1), A)       # [G][U][B]
        yh = host_c64(c["y"]).reshape(S, B, N)
        xt = host_c64(c["x"]).reshape(S, U, N)
        for s in (0, S - 1):
            yp = np.zeros((B, G * sg), np.complex128)
            yp[:, :N] = yh[s]
            x64 = (W64 @ yp.reshape(B, G, sg).transpose(1, 0, 2)).transpose(1, 0, 2)
            x64 = x64.reshape(U, G * sg)[:, :N]
            err = np.abs(xh[s] - x64).max() / np.abs(x64).max()
            assert err < 1e-4, (k, s, err)
            e_gpu, e_64 = synth.evm_db(xh[s], xt[s]), synth.evm_db(x64, xt[s])
            assert abs(e_gpu - e_64) < 0.1, (k, s, e_gpu, e_64)
    b.close()
