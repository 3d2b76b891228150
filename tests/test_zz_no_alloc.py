"""mega_detect_slots_noalloc against mega_detect_slots, through the C ABI on
cuda:0: the same kernels with the split weight planes in the library's
static scratch instead of a caller's work buffer, so every output (LLRs,
x_hat, W, group status, route) must be bit-identical. Two calls in a row,
the second smaller, show that nothing of the first call's scratch contents
leaks into the second; a desc needing more than the scratch's maximum is
refused before anything is launched."""
import numpy as np
import pytest

import megastation_amd as m
from megastation_amd import synth

from gpu_helpers import SENTINEL, dev_c64, host_c64, host_u16, llr_buffer, make_slots, torch

pytestmark = pytest.mark.gpu


def _run(csi, y, U, B, N, sg, Q, S, slots, lay, form, noalloc):
    t = torch()
    G = (N + sg - 1) // sg
    llr = llr_buffer(slots * S * U * lay.ue_stride)
    xh = t.zeros(slots * S * U * N, dtype=t.complex64, device="cuda:0")
    w = t.zeros(slots * G * B * U, dtype=t.complex64, device="cuda:0")
    info = t.full((slots * G,), -5, dtype=t.int32, device="cuda:0")
    route = t.full((slots * G,), -5, dtype=t.int32, device="cuda:0")
    work = None
    if not noalloc and form == m.FORM_MIXED:
        work = t.empty(m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                                           num_slots=slots, form=form),
                       dtype=t.uint8, device="cuda:0")
    m.detect_slots(dev_c64(csi), dev_c64(y), llr, num_ue=U, num_bs=B, num_sc=N,
                   sc_group=sg, num_syms=S, num_slots=slots, mod_order=Q,
                   ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
                   zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                   y_sym_stride=B * N, y_slot_stride=S * B * N,
                   llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
                   xhat=xh, w_out=w, info=info, form=form, work=work,
                   route=route if form == m.FORM_MIXED else None, noalloc=noalloc)
    t.cuda.synchronize()
    return (host_u16(llr), host_c64(xh), host_c64(w), info.cpu().numpy(),
            route.cpu().numpy())


# (U, B, N, sg, Q, S, slots, channel): a 256 x 32 slot (17.9 MB of split
# planes), then a smaller 64 x 16 case of two slots
CASES = [(32, 256, 3276, 12, 6, 2, 1, "kronecker"), (16, 64, 600, 16, 4, 2, 2, "iid")]


@pytest.mark.parametrize("form", ["mixed", "fp32"])
def test_noalloc_equals_detect_slots_on_two_calls(form):
    f = m.FORM_MIXED if form == "mixed" else m.FORM_FP32
    for U, B, N, sg, Q, S, slots, channel in CASES:
        print(f"case: noalloc {form} U={U} B={B} N={N} sg={sg} slots={slots}", flush=True)
        sl = make_slots(U + B + N, slots, U, B, N, sg, S, Q, 25.0, channel)
        lay = synth.ldpc_uplink_layout(N, Q)
        csi = np.stack([s.csi for s in sl])
        y = np.stack([s.y for s in sl])
        if f == m.FORM_MIXED:
            need = m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                                       num_slots=slots, form=f)
            assert 0 < need <= m.DETECT_SCRATCH_BYTES
        ref = _run(csi, y, U, B, N, sg, Q, S, slots, lay, f, noalloc=False)
        got = _run(csi, y, U, B, N, sg, Q, S, slots, lay, f, noalloc=True)
        names = ("llr", "xhat", "w", "info", "route")
        for name, a, b in zip(names, got, ref):
            if name == "route" and f != m.FORM_MIXED:
                continue
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
        assert not (ref[0] == SENTINEL).all()


def test_noalloc_rejects_oversize_and_work():
    print("case: noalloc argument checks", flush=True)
    t = torch()
    U, B, N, sg, Q, S, slots = 32, 256, 3276, 12, 6, 1, 2    # 35.8 MB > 32 MiB
    G = (N + sg - 1) // sg
    assert m.detect_work_bytes(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_slots=slots,
                               form=m.FORM_MIXED) > m.DETECT_SCRATCH_BYTES
    lay = synth.standalone_layout(N, Q)
    csi = t.zeros(slots * G * B * U, dtype=t.complex64, device="cuda:0")
    y = t.zeros(slots * S * B * N, dtype=t.complex64, device="cuda:0")
    llr = llr_buffer(slots * S * U * lay.ue_stride)
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=slots,
              mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
              cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
              csi_slot_stride=G * B * U, y_sym_stride=B * N, y_slot_stride=S * B * N,
              llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
              form=m.FORM_MIXED, noalloc=True)
    with pytest.raises(m.MegaError) as e:
        m.detect_slots(csi, y, llr, **kw)
    assert e.value.code == m.MEGA_ERR_UNSUPPORTED
    assert "scratch" in str(e.value)
    work = t.empty(16, dtype=t.uint8, device="cuda:0")
    with pytest.raises(m.MegaError) as e:
        m.detect_slots(csi, y, llr, work=work, **dict(kw, num_slots=1))
    assert e.value.code == m.MEGA_ERR_INVALID_ARG
    t.cuda.synchronize()
    assert (host_u16(llr) == SENTINEL).all()                 # nothing was launched
