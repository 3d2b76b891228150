"""The slot pipeline (mega_pipeline_* through megastation_amd.pipeline) against
the sequential path on cuda:0: ZF of slot i+1 overlapped with the equalize of
slot i on a second stream, the weights triple-buffered, over more slots than
buffers. Same kernels on the same inputs, so LLRs and x_hat must be equal bit
for bit to those of ZF then equalize slot by slot, in both forms; one slot is
also checked against the FP64 oracle's EVM. The staged case writes each
slot's csi into a two-deep ring on the pipeline's ZF stream just before its
push, as a channel estimator feeding the pipeline would (a read of a stale or
overwritten ring entry would change the LLRs)."""
import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth
from megastation_amd.pipeline import PipelineGraph, SlotPipeline

from gpu_helpers import SENTINEL, dev_c64, host_c64, host_u16, llr_buffer, torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form,staged", [(m.FORM_FP32, False), (m.FORM_MIXED, False),
                                         (m.FORM_MIXED, True)])
def test_pipeline_equals_sequential(form, staged):
    print(f"case: slot pipeline form={form} staged={staged}", flush=True)
    t = torch()
    U, B, N, sg, S, Q, n_slots = 32, 128, 600, 12, 2, 6, 7
    G = (N + sg - 1) // sg
    lay = synth.ldpc_uplink_layout(N, Q)
    rng = np.random.Generator(np.random.Philox(77 + form))
    sl = [synth.make_slot(rng, U, B, N, sg, S, Q, 25.0, "kronecker") for _ in range(n_slots)]
    csi = [dev_c64(s_.csi.reshape(-1)) for s_ in sl]
    y = [dev_c64(s_.y.reshape(-1)) for s_ in sl]
    n_llr = S * U * lay.ue_stride
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits)
    # sequential: ZF then equalize, slot by slot
    Bp = (B + 15) // 16 * 16
    seq = []
    for i in range(n_slots):
        llr = llr_buffer(n_llr)
        xh = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
        if form == m.FORM_MIXED:
            w = t.empty(G * U * Bp, dtype=t.int64, device="cuda:0")
            m.zf_split(csi[i], w, B, U, G)
            slot_stride = G * U * Bp
        else:
            w = t.empty(G * B * U, dtype=t.complex64, device="cuda:0")
            m.zf(csi[i], w, B, U, G)
            slot_stride = G * B * U
        m.equalize_slots(w, y[i], llr, xhat=xh, form=form, num_slots=1,
                         csi_slot_stride=slot_stride, y_sym_stride=B * N,
                         y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
                         llr_slot_stride=n_llr, **kw)
        seq.append((llr, xh))
    t.cuda.synchronize()
    # pipelined
    p = SlotPipeline(form=form, **kw)
    ring = [t.zeros_like(csi[0]) for _ in range(2)]
    p.zs.wait_stream(t.cuda.current_stream())  # ring and csi are ready
    out = []
    for i in range(n_slots):
        llr = llr_buffer(n_llr)
        xh = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
        if staged:
            with t.cuda.stream(p.zs):
                ring[i % 2].copy_(csi[i])
            p.push(ring[i % 2], y[i], llr, xh)
        else:
            p.push(csi[i], y[i], llr, xh)
        out.append((llr, xh))
    p.flush()
    t.cuda.synchronize()
    p.close()
    for i in range(n_slots):
        assert np.array_equal(host_u16(out[i][0]), host_u16(seq[i][0])), i
        assert np.array_equal(host_c64(out[i][1]), host_c64(seq[i][1])), i
    # and the symbols are right: EVM of the last slot within 0.1 dB of FP64's
    w64, _, _ = oracle.zf(sl[-1].csi, U, B, G, "f64")
    xg = host_c64(out[-1][1]).reshape(S, U, N)
    for s_ in range(S):
        init = np.full((U, lay.ue_stride), SENTINEL, dtype=np.uint16)
        _, x64 = oracle.equalize(w64, sl[-1].y[s_], U, B, N, sg, Q, lay.ue_stride, lay.cb_size,
                                 lay.cb_stride, lay.zeros, lay.valid_bits, "f64",
                                 llr_init=init)
        e = synth.evm_db(xg[s_], sl[-1].x[s_]) - synth.evm_db(x64, sl[-1].x[s_])
        assert abs(e) < 0.1, e


@pytest.mark.parametrize("form", [m.FORM_FP32, m.FORM_MIXED])
def test_pipeline_graph_equals_sequential(form):
    """mega_pipeline_graph_*: five slots with distinct buffers captured into
    one graph, launched twice (the second launch rewrites the same outputs);
    LLRs and x_hat of every slot bit-equal to ZF then equalize slot by slot."""
    print(f"case: pipeline graph form={form}", flush=True)
    t = torch()
    U, B, N, sg, S, Q, n_slots = 32, 256, 600, 12, 2, 6, 5
    G = (N + sg - 1) // sg
    lay = synth.ldpc_uplink_layout(N, Q)
    rng = np.random.Generator(np.random.Philox(91 + form))
    sl = [synth.make_slot(rng, U, B, N, sg, S, Q, 25.0, "iid") for _ in range(n_slots)]
    csi = [dev_c64(s_.csi.reshape(-1)) for s_ in sl]
    y = [dev_c64(s_.y.reshape(-1)) for s_ in sl]
    n_llr = S * U * lay.ue_stride
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits)
    Bp = (B + 15) // 16 * 16
    seq = []
    for i in range(n_slots):
        llr = llr_buffer(n_llr)
        xh = t.zeros(S * U * N, dtype=t.complex64, device="cuda:0")
        if form == m.FORM_MIXED:
            w = t.empty(G * U * Bp, dtype=t.int64, device="cuda:0")
            m.zf_split(csi[i], w, B, U, G)
        else:
            w = t.empty(G * B * U, dtype=t.complex64, device="cuda:0")
            m.zf(csi[i], w, B, U, G)
        m.equalize_slots(w, y[i], llr, xhat=xh, form=form, num_slots=1,
                         csi_slot_stride=w.numel(), y_sym_stride=B * N,
                         y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
                         llr_slot_stride=n_llr, **kw)
        seq.append((llr, xh))
    t.cuda.synchronize()
    outs = [(llr_buffer(n_llr), t.zeros(S * U * N, dtype=t.complex64, device="cuda:0"))
            for _ in range(n_slots)]
    gr = PipelineGraph([(csi[i], y[i], outs[i][0], outs[i][1]) for i in range(n_slots)],
                       form=form, **kw)
    for rep in range(2):
        if rep:
            for o in outs:
                o[1].zero_()
        gr.launch()
        t.cuda.synchronize()
        for i in range(n_slots):
            assert np.array_equal(host_u16(outs[i][0]), host_u16(seq[i][0])), (rep, i)
            assert np.array_equal(host_c64(outs[i][1]), host_c64(seq[i][1])), (rep, i)
    gr.close()


@pytest.mark.parametrize("form", [m.FORM_FP32, m.FORM_MIXED])
def test_pipeline_two_slots_per_push(form):
    """slots=2: a push takes two consecutive slots (one ZF launch over both
    slots' groups, one equalize launch); three pushes eager and as a graph,
    bit-equal to zf + equalize_slots called on the same two slots."""
    print(f"case: pipeline two slots per push form={form}", flush=True)
    t = torch()
    U, B, N, sg, S, Q, n_push = 32, 256, 600, 12, 2, 6, 3
    G = (N + sg - 1) // sg
    lay = synth.standalone_layout(N, Q)
    rng = np.random.Generator(np.random.Philox(17 + form))
    pairs = [[synth.make_slot(rng, U, B, N, sg, S, Q, 25.0, "iid") for _ in range(2)]
             for _ in range(n_push)]
    csi = [dev_c64(np.concatenate([s_.csi.reshape(-1) for s_ in p_])) for p_ in pairs]
    y = [dev_c64(np.concatenate([s_.y.reshape(-1) for s_ in p_])) for p_ in pairs]
    n_llr = 2 * S * U * lay.ue_stride
    kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits)
    Bp = (B + 15) // 16 * 16
    seq = []
    for i in range(n_push):
        llr = llr_buffer(n_llr)
        if form == m.FORM_MIXED:
            w = t.empty(2 * G * U * Bp, dtype=t.int64, device="cuda:0")
            m.zf_split(csi[i], w, B, U, 2 * G)
        else:
            w = t.empty(2 * G * B * U, dtype=t.complex64, device="cuda:0")
            m.zf(csi[i], w, B, U, 2 * G)
        m.equalize_slots(w, y[i], llr, form=form, num_slots=2, csi_slot_stride=w.numel() // 2,
                         y_sym_stride=B * N, y_slot_stride=S * B * N,
                         llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride,
                         **kw)
        seq.append(host_u16(llr))
    t.cuda.synchronize()
    p = SlotPipeline(form=form, slots=2, **kw)
    p.zs.wait_stream(t.cuda.current_stream())
    outs = [llr_buffer(n_llr) for _ in range(n_push)]
    for i in range(n_push):
        p.push(csi[i], y[i], outs[i])
    p.flush()
    t.cuda.synchronize()
    p.close()
    for i in range(n_push):
        assert np.array_equal(host_u16(outs[i]), seq[i]), i
    outs_g = [llr_buffer(n_llr) for _ in range(n_push)]
    gr = PipelineGraph([(csi[i], y[i], outs_g[i]) for i in range(n_push)], form=form, slots=2,
                       **kw)
    gr.launch()
    t.cuda.synchronize()
    gr.close()
    for i in range(n_push):
        assert np.array_equal(host_u16(outs_g[i]), seq[i]), i


def test_pipeline_rejects_bad_arguments():
    with pytest.raises(ValueError):
        SlotPipeline(num_ue=4, num_bs=8, num_sc=16, sc_group=4, num_syms=1, mod_order=4,
                     ue_stride=64, cb_size=64, cb_stride=64, zeros=0, valid_bits=64, form=7)
    with pytest.raises(ValueError):
        SlotPipeline(num_ue=4, num_bs=8, num_sc=16, sc_group=4, num_syms=1, mod_order=4,
                     ue_stride=64, cb_size=64, cb_stride=64, zeros=0, valid_bits=64, depth=1)
    with pytest.raises(ValueError):
        SlotPipeline(num_ue=4, num_bs=8, num_sc=16, sc_group=4, num_syms=1, mod_order=4,
                     ue_stride=64, cb_size=64, cb_stride=64, zeros=0, valid_bits=64, slots=0)
    with pytest.raises(ValueError):
        PipelineGraph([], num_ue=4, num_bs=8, num_sc=16, sc_group=4, num_syms=1, mod_order=4,
                      ue_stride=64, cb_size=64, cb_stride=64, zeros=0, valid_bits=64)
