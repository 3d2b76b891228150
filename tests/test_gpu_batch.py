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
