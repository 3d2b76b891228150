#!/usr/bin/env python3
"""Detected symbols/s of the uplink detection path (ZF + equalize + demod) on
one MI355X.

Workload (default, c3_256x32): one 256-antenna x 32-user cell, 3276 data
subcarriers in groups of 12 (273 groups), 13 data OFDM symbols per slot,
64-QAM at 25 dB SNR over a Kronecker-correlated channel (rho_rx 0.5,
rho_tx 0.3). One step = one slot: ZF weights of its 273 groups + equalize /
demod of its 13 symbols, inputs resident in HBM (four rotating buffer sets so
that the slot is not served from the Infinity Cache). Forms: the plain FP32
drop-in (ZF kernel writing W, then the equalize kernel), the mixed form
(split-BF16 matrix-core ZF and equalize), both slot-pipelined (ZF of step
i+1 on a second stream beside the equalize of step i) and graph-captured
with one or two slots per push, and the fused FP32 kernel where the shape
allows. value = the fastest form whose arithmetic is FP32 throughout and
whose EVM on slot 0 is within 0.1 dB of FP64 zero-forcing's (dtype fp32);
the fastest form of all (the chosen form) and every form's ratio to the
plain drop-in are reported beside it. A sub-record times the 64 x 16 cell
batched over 200 slots (--workload c2_64x16_200slots runs it alone, with
an EVM / BER sweep over SNR and QAM order). Synthetic data generated on the
GPU with a seeded Philox generator.

Prints one JSON line: value = detected symbols/s = U*N*S*slots / step time;
roofline of the dominant kernel (the FP32 equalize/demod, HBM-bound) from HIP
events on the launch stream; the CPU baseline (the oracle's OpenMP
restatement of the same operator math) timed on a bounded sample on the host
cores; EVM of the GPU's equalized symbols next to FP64 zero-forcing's (numpy
complex128) on a sample slot.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "detected symbols/s (ZF+equalize+demod), 1 GPU, at EVM within 0.1 dB of FP64"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # 64x16 cell x 200 slots (the batched GPU configuration)
    "c2_64x16_200slots": dict(U=16, B=64, N=1200, sg=16, S=13, slots=200, Q=4,
                              snr=20.0, channel="iid"),
    # 256x32 cell, one slot
    "c3_256x32": dict(U=32, B=256, N=3276, sg=12, S=13, slots=1, Q=6, snr=25.0,
                      channel="kronecker"),
}


def batch_cells(n_cells=40, seed=7):
    """The uneven 40-cell batch (megastation_amd.synth.uneven_cells)."""
    from megastation_amd import synth
    return synth.uneven_cells(n_cells, seed)


def gen_inputs(cfg, seed=1234):
    """CSI [slots][G][B][U] (= H^H), y [slots][S][B][N], the transmitted
    symbols x [slots][S][U][N] and their indices (uint8) on cuda:0."""
    import torch
    from megastation_amd import synth
    U, B, N, sg, S, slots, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "slots", "Q"))
    G = (N + sg - 1) // sg
    Np = G * sg
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    if cfg["channel"] == "iid":
        H = torch.randn((slots * G, B, U), dtype=torch.complex64, device=dev, generator=gen)
    else:
        rng = np.random.Generator(np.random.Philox(seed))
        Hn = np.stack([synth.channel_kronecker(rng, G, B, U, 0.5, 0.3) for _ in range(slots)])
        H = torch.from_numpy(Hn.reshape(slots * G, B, U)).to(dev)
    table = torch.from_numpy(synth.qam_table(Q)).to(dev)
    idx = torch.randint(0, 1 << Q, (slots, S, U, Np), device=dev, generator=gen)
    x = table[idx]                                             # [slots,S,U,Np]
    X = x.reshape(slots, S, U, G, sg).permute(0, 3, 2, 1, 4).reshape(slots * G, U, S * sg)
    Hr, Hi = H.real.contiguous(), H.imag.contiguous()
    Xr, Xi = X.real.contiguous(), X.imag.contiguous()
    Yr = torch.bmm(Hr, Xr) - torch.bmm(Hi, Xi)
    Yi = torch.bmm(Hr, Xi) + torch.bmm(Hi, Xr)
    del Xr, Xi
    Y = torch.complex(Yr, Yi)
    del Yr, Yi
    sigma = math.sqrt(10.0 ** (-cfg["snr"] / 10.0))
    Y = Y + sigma * torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=gen)
    y = (Y.reshape(slots, G, B, S, sg).permute(0, 3, 2, 1, 4)
         .reshape(slots, S, B, Np)[..., :N].contiguous())
    csi = torch.conj(H).resolve_conj().reshape(slots, G, B, U).contiguous()
    x = x[..., :N].contiguous()
    return csi, y, x, idx[..., :N].to(torch.uint8)


# Forms derived from a sequential one (same kernels, another schedule), and
# the chunk counts of the chunked forms.
DERIVED = ("_pipe", "_graph", "_graph2", "_chunk2", "_chunk4", "_chunk8")
CHUNKS = (2, 4, 8)
# Forms whose arithmetic is FP32 throughout (the reference's precision): the
# headline value is the fastest of these; the mixed forms are the chosen
# precision form, reported beside it with their ratio to the plain drop-in.
FP32_FORMS = ("plain", "plain_pipe", "plain_graph", "plain_graph2", "fused")


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3_256x32",
                    choices=sorted(WORKLOADS) + ["c4_40cells", "c5_learned",
                                                 "d3_precode_256x32", "u3_fft_256x32",
                                                 "l3_ldpc_256x32", "uc3_uplink_256x32"])
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="approximate CPU time budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="c3_256x32: skip the 64x16 x 200-slot sub-record")
    ap.add_argument("--forms", default="",
                    help="single-cell workloads: comma-separated forms to time (default all; "
                         "e.g. 'plain' to profile the sequential FP32 kernels alone)")
    ap.add_argument("--sweep-slots", type=int, default=20,
                    help="c2_64x16_200slots: slots per SNR / QAM point of the EVM / BER sweep "
                         "(0 = no sweep)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of a known size (1 GiB), whose "
                         "FETCH_SIZE / WRITE_SIZE counters scale the equalize traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the detection path shards by subcarrier group with no exchange; "
              "run one process per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c4_40cells":
        return main_batch(args)
    if args.workload == "c5_learned":
        return main_learned(args)
    if args.workload == "d3_precode_256x32":
        return main_precode(args)
    if args.workload == "u3_fft_256x32":
        return main_fft(args)
    if args.workload == "l3_ldpc_256x32":
        return main_ldpc(args)
    if args.workload == "uc3_uplink_256x32":
        return main_uplink(args)

    if args.copy_calibration:
        import torch
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    result = run_cell(args, args.workload, cpu=not args.no_cpu_baseline,
                      sel=tuple(f for f in args.forms.split(",") if f) or None)
    if args.workload == "c3_256x32" and not args.no_sub:
        sub = run_cell(args, "c2_64x16_200slots", sel=("plain", "mixed"), cpu=False)
        result["sub"] = {"c2_64x16_200slots": {
            k: sub[k] for k in ("value", "ms_per_step", "dtype", "config", "roofline", "forms",
                                "evm")}}
    if args.workload == "c2_64x16_200slots" and args.sweep_slots > 0:
        result["sweep"] = snr_sweep(args.sweep_slots)
    print(json.dumps(result), flush=True)


def run_cell(args, name, sel=None, cpu=True):
    """Times the forms of one single-cell workload (all of them, or those in
    sel) and returns the result record (value = fastest FP32 form)."""
    import torch
    import megastation_amd as m
    from megastation_amd import synth

    cfg = WORKLOADS[name]
    U, B, N, sg, S, slots, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "slots", "Q"))
    G = (N + sg - 1) // sg
    lay = synth.standalone_layout(N, Q)
    dev = torch.device("cuda:0")
    print(f"[bench] workload {name}: generating inputs", file=sys.stderr, flush=True)
    csi0, y0, x, _ = gen_inputs(cfg)
    # Successive steps rotate through NR distinct sets of input and output
    # buffers, so that a one-slot working set (about 140 MB at 256x32, inside
    # the 256 MB Infinity Cache) is not served from that cache step after step:
    # four sets for one-slot workloads, one for multi-slot ones (GBs per step).
    NR = 4 if slots == 1 else 1
    csi_r = [csi0] + [csi0.clone() for _ in range(NR - 1)]
    y_r = [y0] + [y0.clone() for _ in range(NR - 1)]
    w_r = [torch.empty_like(csi0) for _ in range(NR)]
    llr_r = [torch.empty(slots * S * U * lay.ue_stride, dtype=torch.int16, device=dev)
             for _ in range(NR)]
    csi, y, w = csi0, y0, w_r[0]
    ring = [0]

    def nxt():
        """Buffer set of the next step (csi, y, w, llr, index)."""
        j = ring[0] % NR
        ring[0] += 1
        return csi_r[j], y_r[j], w_r[j], llr_r[j], j
    stream = torch.cuda.current_stream()

    eq_kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S,
                 num_slots=slots, mod_order=Q, ue_stride=lay.ue_stride,
                 cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                 valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                 y_sym_stride=B * N, y_slot_stride=S * B * N,
                 llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride)

    def step_plain(evs=None):
        """Drop-in form: ZF kernel (W to HBM) then the equalize kernel."""
        c_, y_, w_, l_, _ = nxt()
        if evs:
            evs[0].record(stream)
        m.zf(c_, w_, B, U, slots * G, stream=stream)
        if evs:
            evs[1].record(stream)
        m.equalize_slots(w_, y_, l_, stream=stream, **eq_kw)
        if evs:
            evs[2].record(stream)

    Bp = (B + 15) // 16 * 16
    split_r = [torch.empty(slots * G * U * Bp, dtype=torch.int64, device=dev)
               for _ in range(NR)]
    split = split_r[0]

    def step_mixed(evs=None):
        """Mixed form: split-BF16 matrix-core ZF (split weight words to HBM),
        then the split-BF16 equalize kernel."""
        c_, y_, _, l_, j = nxt()
        if evs:
            evs[0].record(stream)
        m.zf_split(c_, split_r[j], B, U, slots * G, stream=stream)
        if evs:
            evs[1].record(stream)
        m.equalize_slots(split_r[j], y_, l_, form=m.FORM_MIXED, stream=stream,
                         **dict(eq_kw, csi_slot_stride=G * U * Bp))
        if evs:
            evs[2].record(stream)

    def step_fused(evs=None):
        """Fused form: ZF + equalize + demod in one kernel, W kept on chip."""
        c_, y_, _, l_, _ = nxt()
        if evs:
            evs[0].record(stream)
        m.detect_slots(c_, y_, l_, stream=stream, **eq_kw)
        if evs:
            evs[1].record(stream)
            evs[2].record(stream)

    # Slot-pipelined forms (single-slot workloads): the library's slot pipeline
    # (mega_pipeline_*, through megastation_amd.pipeline.SlotPipeline) issues
    # the ZF of step i+1 on its own stream before the equalize of step i, so
    # the two kernels run together (W triple-buffered, every dependency an
    # event, one C call per step).
    from megastation_amd.pipeline import PipelineGraph, SlotPipeline
    pipes = {}

    def get_pipe(mixed):
        if mixed not in pipes:
            pipes[mixed] = SlotPipeline(
                num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
                ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
                zeros=lay.zeros, valid_bits=lay.valid_bits,
                form=m.FORM_MIXED if mixed else m.FORM_FP32, stream=stream)
        return pipes[mixed]

    def step_pipe(mixed, evs=None):
        c_, y_, _, l_, _ = nxt()
        get_pipe(mixed).push(c_, y_, l_)

    # Graph forms (single-slot workloads): K consecutive slots (K = --steps,
    # buffer sets rotating as above) pipelined as SlotPipeline would and
    # captured into one graph (mega_pipeline_graph_*); one launch runs the K
    # steps with no host work between them. Two slots per push: pairs of
    # consecutive slots in contiguous buffers (the four rotating slots as two
    # pairs); the ZF of a pair is one launch over both slots' groups, its
    # equalize one launch.
    pair_r = []

    def pairs():
        if not pair_r:
            for j in range(0, NR, 2):
                pair_r.append((torch.cat([csi_r[j], csi_r[j + 1]]),
                               torch.cat([y_r[j], y_r[j + 1]]),
                               torch.cat([llr_r[j], llr_r[j + 1]])))
        return pair_r

    def make_graph(mixed, per_push=1):
        if per_push == 2:
            pr = pairs()
            pushes = [pr[i % len(pr)] for i in range(max(1, args.steps // 2))]
        else:
            pushes = [(csi_r[i % NR], y_r[i % NR], llr_r[i % NR]) for i in range(args.steps)]
        return PipelineGraph(
            pushes, num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
            ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
            zeros=lay.zeros, valid_bits=lay.valid_bits,
            form=m.FORM_MIXED if mixed else m.FORM_FP32, slots=per_push), per_push * len(pushes)

    def timed_graph(gr, steps):
        """W warm-up launches, then one timed launch of `steps` steps."""
        for _ in range(max(1, args.warmup)):
            gr.launch(stream)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        gr.launch(stream)
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps, None, None

    host_ms = {}  # host time to issue one pipelined step (launches, waits, records)

    def timed(step, pipe=None):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        if pipe is not None:  # the timed region ends when both streams have drained
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            pipe.zs.wait_stream(stream)
            h0 = time.perf_counter()
            for i in range(args.steps):  # K ZFs and K equalizes (the first one pending)
                step()
            host_ms[pipe] = (time.perf_counter() - h0) * 1e3 / args.steps
            stream.wait_stream(pipe.zs)
            t1.record(stream)
            torch.cuda.synchronize()
            pipe.flush()
            torch.cuda.synchronize()
            return t0.elapsed_time(t1) / args.steps, None, None
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)]
               for _ in range(args.steps)]
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for i in range(args.steps):
            step(evs[i])
        t1.record(stream)
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / args.steps
        k1 = statistics.mean(e[0].elapsed_time(e[1]) for e in evs)
        k2 = statistics.mean(e[1].elapsed_time(e[2]) for e in evs)
        return ms, k1, k2

    def want(k):
        return sel is None or k in sel

    forms = {}
    if want("fused") and m.detect_fused_shape(U, B):
        forms["fused"] = timed(step_fused)
    for k, fn in (("plain", step_plain), ("mixed", step_mixed)):
        if want(k):
            forms[k] = timed(fn)
    print(f"[bench] {name}: sequential forms timed ({args.steps} steps)", file=sys.stderr,
          flush=True)
    if slots == 1:
        for k, mixed in (("plain_pipe", False), ("mixed_pipe", True)):
            if want(k):
                forms[k] = timed(lambda evs=None, mx=mixed: step_pipe(mx), pipe=get_pipe(mixed))
        for k, mixed, per in (("plain_graph", False, 1), ("mixed_graph", True, 1),
                              ("plain_graph2", False, 2), ("mixed_graph2", True, 2)):
            if want(k):
                gr, steps = make_graph(mixed, per)
                forms[k] = timed_graph(gr, steps)
                gr.close()
        print(f"[bench] {name}: pipelined and graph forms timed", file=sys.stderr, flush=True)

    # Chunked forms (multi-slot workloads): the job's slots go through the
    # slot pipeline in C pushes of slots / C, so the ZF of chunk j+1 (bound by
    # latency) runs beside the equalize of chunk j (bound by HBM). Same kernels
    # as the mixed form; the LLRs of one chunked step are compared with the
    # mixed form's bit for bit below.
    chunk_pipes = {}

    def get_chunk_pipe(C):
        if C not in chunk_pipes:
            chunk_pipes[C] = SlotPipeline(
                num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, mod_order=Q,
                ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
                zeros=lay.zeros, valid_bits=lay.valid_bits, form=m.FORM_MIXED,
                stream=stream, slots=slots // C)
        return chunk_pipes[C]

    def step_chunk(C):
        p, cs = get_chunk_pipe(C), slots // C
        ls = cs * S * U * lay.ue_stride
        for j in range(C):
            p.push(csi0[j * cs:(j + 1) * cs], y0[j * cs:(j + 1) * cs],
                   llr_r[0][j * ls:(j + 1) * ls])

    if slots > 1:
        for C in CHUNKS:
            if slots % C or not want(f"mixed_chunk{C}"):
                continue
            forms[f"mixed_chunk{C}"] = timed(lambda evs=None, C=C: step_chunk(C),
                                             pipe=get_chunk_pipe(C))
            print(f"[bench] mixed_chunk{C} form timed ({args.steps} steps)", file=sys.stderr,
                  flush=True)
    chunk_same = {}
    for C in chunk_pipes:
        llr_r[0].fill_(-1)
        step_mixed()
        torch.cuda.synchronize()
        ref = llr_r[0].clone()
        llr_r[0].fill_(-1)
        step_chunk(C)
        chunk_pipes[C].flush()
        torch.cuda.synchronize()
        chunk_same[f"mixed_chunk{C}"] = bool(torch.equal(ref, llr_r[0]))
        del ref

    symbols = U * N * S * slots
    io_bytes = 8 * B * N * S * slots + 8 * U * B * G * slots + 2 * U * Q * N * S * slots
    zf_flops = G * slots * (4 * U * (U + 1) * B + 8 * U * U * B + (4 / 3) * U ** 3)
    eq_flops = 8 * U * B * N * S * slots

    # ---- EVM of every form on slot 0 (x_hat side output, untimed) against the
    # transmitted symbols, next to FP64 zero-forcing's ---------------------------
    xh = torch.zeros(S * U * N, dtype=torch.complex64, device=dev)
    kw1 = dict(eq_kw, num_slots=1)
    llr1 = torch.empty(S * U * lay.ue_stride, dtype=torch.int16, device=dev)
    x0 = x[0].cpu().numpy()
    evm = {}
    # the sequential forms, and the base form of every derived one (same
    # kernels) even where only the derived form was timed
    bases = dict.fromkeys(k.rsplit("_", 1)[0] if k.endswith(DERIVED) else k for k in forms)
    for k in bases:
        xh.zero_()
        if k == "fused":
            m.detect_slots(csi[0], y[0], llr1, xhat=xh, stream=stream, **kw1)
        elif k == "plain":
            m.zf(csi[0], w[0], B, U, G, stream=stream)
            m.equalize_slots(w[0], y[0], llr1, xhat=xh, stream=stream, **kw1)
        else:
            m.zf_split(csi[0], split, B, U, G, stream=stream)
            m.equalize_slots(split, y[0], llr1, xhat=xh, form=m.FORM_MIXED,
                             stream=stream, **dict(kw1, csi_slot_stride=G * U * Bp))
        torch.cuda.synchronize()
        evm[k] = synth.evm_db(xh.cpu().numpy().reshape(S, U, N), x0)
    for k in forms:
        if k.endswith(DERIVED):  # same kernels as the sequential form
            evm[k] = evm[k.rsplit("_", 1)[0]]
    evm64 = synth.evm_db(fp64_xhat(cfg, csi[:1], y[:1])[0], x0)
    ok = {k: abs(evm[k] - evm64) < 0.1 and chunk_same.get(k, True) for k in forms}

    # value = the fastest FP32 form within 0.1 dB of FP64's EVM; the chosen
    # (fastest of all) form and every form's ratio to the plain drop-in beside it
    fp32 = [k for k in forms if k in FP32_FORMS and ok.get(k)]
    head_form = min(fp32, key=lambda k: forms[k][0]) if fp32 else "plain"
    anyf = [k for k in forms if ok.get(k)]
    chosen = min(anyf, key=lambda k: forms[k][0]) if anyf else head_form
    ms_step = forms[head_form][0]
    ms_plain = forms["plain"][0] if "plain" in forms else None
    # The dominant kernel's time for the roofline: a batch of back-to-back
    # launches over the rotating buffer sets between two HIP events on the
    # launch stream (an event pair around every launch adds a few
    # microseconds of its own at 256x32), W already computed for every set.
    def batch_avg(launch, n=20):
        for j in range(NR):
            launch(j)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(n):
            launch(i % NR)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    if head_form == "fused":
        dom_kernel, dom_prefix = "detect_fused (ZF + equalize + demod)", "detect_fused"
        dom_ms = batch_avg(lambda j: m.detect_slots(csi_r[j], y_r[j], llr_r[j], stream=stream,
                                                    **eq_kw))
    else:
        dom_kernel, dom_prefix = "equalize_demod (FP32)", "equalize_"
        for j in range(NR):
            m.zf(csi_r[j], w_r[j], B, U, slots * G, stream=stream)
        dom_ms = batch_avg(lambda j: m.equalize_slots(w_r[j], y_r[j], llr_r[j], stream=stream,
                                                      **eq_kw))
    value = symbols / (ms_step * 1e-3)
    achieved = io_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(name, dom_prefix)
    print(f"[bench] {name}: value {head_form} {ms_step:.4f} ms/step, chosen {chosen} "
          f"{forms[chosen][0]:.4f}; " +
          ", ".join(f"{k} {v[0]:.4f}" for k, v in forms.items()), file=sys.stderr, flush=True)

    def form_entry(k):
        ms, t1, t2 = forms[k]
        e = {"symbols_per_s": symbols / (ms * 1e-3), "ms_per_step": ms,
             "arith": "fp32" if k in FP32_FORMS else "bf16x3 split, fp32 accumulate",
             "ratio_to_plain": ms_plain / ms if ms_plain else None,
             "evm_db": evm[k], "evm_delta_db": evm[k] - evm64, "evm_ok": ok[k]}
        if k.endswith("_pipe"):
            e["schedule"] = ("ZF of step i+1 on a second stream, issued before the "
                             "equalize of step i; W triple-buffered; one C call per step")
            e["host_ms_per_step"] = host_ms.get(get_pipe(k == "mixed_pipe"))
        elif k.endswith("_graph"):
            e["schedule"] = (f"the {k[:-6]}_pipe schedule of {args.steps} consecutive steps "
                             "captured into one graph, one launch")
        elif "_chunk" in k:
            C = int(k.rsplit("chunk", 1)[1])
            e["schedule"] = (f"the {slots} slots in {C} pushes of {slots // C} through the "
                             "slot pipeline: ZF of chunk j+1 on a second stream beside the "
                             "equalize of chunk j")
            e["llr_bit_identical_to_mixed"] = chunk_same.get(k)
            e["host_ms_per_step"] = host_ms.get(get_chunk_pipe(C))
        elif k.endswith("_graph2"):
            e["schedule"] = (f"two consecutive slots per push (ZF of both slots' groups in one "
                             f"launch, equalize of both in one), the pipelined schedule of "
                             f"{2 * max(1, args.steps // 2)} steps captured into one graph")
        elif k != "fused":
            e.update({"zf_ms": t1, "equalize_ms": t2,
                      "zf_tflops": zf_flops / (t1 * 1e-3) / 1e12,
                      "equalize_gbs": io_bytes / (t2 * 1e-3) / 1e9,
                      "equalize_hbm_frac": io_bytes / (t2 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "equalize_tflops": eq_flops / (t2 * 1e-3) / 1e12})
        return e

    result = {
        "metric": METRIC, "value": value, "unit": "symbols/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": name, "antennas": B, "users": U,
                   "subcarriers": N, "sc_group": sg, "groups": G, "data_symbols": S,
                   "slots": slots, "qam": 1 << Q, "snr_db": cfg["snr"],
                   "channel": cfg["channel"], "form": head_form,
                   "parallelism": "1 GPU, groups x slots"},
        "roofline": {"bound": "hbm", "kernel": dom_kernel,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["traffic_bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "algorithmic_bytes": io_bytes, "avg_ms": dom_ms},
        "chosen_form": {"form": chosen, "ms_per_step": forms[chosen][0],
                        "symbols_per_s": symbols / (forms[chosen][0] * 1e-3),
                        "ratio_to_plain": ms_plain / forms[chosen][0] if ms_plain else None,
                        "ratio_to_value": ms_step / forms[chosen][0]},
        "forms": {k: form_entry(k) for k in forms},
        "evm": {"form": head_form, "gpu_db": evm[head_form], "fp64_db": evm64,
                "delta_db": evm[head_form] - evm64,
                "within_0p1_db": ok[head_form]},
    }
    if cpu:
        result["cpu_baseline"] = cpu_baseline(cfg, lay, csi, y, x, args.cpu_seconds)
    for p in list(pipes.values()) + list(chunk_pipes.values()):
        p.close()
    return result


def snr_sweep(n_slots):
    """EVM and BER of the plain and mixed forms on the 64 x 16 cell (1200
    subcarriers, groups of 16, 13 symbols, i.i.d. channels) at SNR 10-30 dB
    in 5 dB steps, 16- and 64-QAM, n_slots slots per point, next to FP64
    zero-forcing (numpy complex128) on the same inputs. BER counts hard
    decisions (LLR sign, LLR > 0 = bit 0) against the transmitted bits;
    'vs_fp64' counts them against the hard decisions of the FP64 x_hat."""
    import torch
    import megastation_amd as m
    from megastation_amd import synth
    out = []
    dev = torch.device("cuda:0")
    for Q in (4, 6):
        for snr in (10.0, 15.0, 20.0, 25.0, 30.0):
            cfg = dict(WORKLOADS["c2_64x16_200slots"], slots=n_slots, Q=Q, snr=snr)
            U, B, N, sg, S, slots = (cfg[k] for k in ("U", "B", "N", "sg", "S", "slots"))
            G = (N + sg - 1) // sg
            Bp = (B + 15) // 16 * 16
            lay = synth.standalone_layout(N, Q)
            csi, y, x, idx = gen_inputs(cfg, seed=int(100 * snr) + Q)
            kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=slots,
                      mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
                      cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
                      csi_slot_stride=G * B * U, y_sym_stride=B * N, y_slot_stride=S * B * N,
                      llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride)
            llr = torch.empty(slots * S * U * lay.ue_stride, dtype=torch.int16, device=dev)
            xh = torch.empty(slots * S * U * N, dtype=torch.complex64, device=dev)
            bits = synth.index_bits(idx.cpu().numpy(), Q)              # [slots][S][U][N][Q]
            x64 = fp64_xhat(cfg, csi, y)
            h64 = synth.hard_bits_xhat(x64, Q)
            xn = x.cpu().numpy()
            rec = {"qam": 1 << Q, "snr_db": snr, "slots": slots, "bits": int(bits.size),
                   "fp64": {"evm_db": synth.evm_db(x64, xn),
                            "ber": float(np.mean(h64 != bits))}}
            for form in ("plain", "mixed"):
                if form == "plain":
                    w = torch.empty_like(csi)
                    m.zf(csi, w, B, U, slots * G)
                    m.equalize_slots(w, y, llr, xhat=xh, **kw)
                else:
                    sp = torch.empty(slots * G * U * Bp, dtype=torch.int64, device=dev)
                    m.zf_split(csi, sp, B, U, slots * G)
                    m.equalize_slots(sp, y, llr, xhat=xh, form=m.FORM_MIXED,
                                     **dict(kw, csi_slot_stride=G * U * Bp))
                torch.cuda.synchronize()
                l16 = llr.cpu().numpy().view(np.uint16).reshape(slots, S, U, lay.ue_stride)
                hb = synth.hard_bits_llr(l16[..., :N * Q].reshape(slots, S, U, N, Q))
                xg = xh.cpu().numpy().reshape(slots, S, U, N)
                rec[form] = {"evm_db": synth.evm_db(xg, xn),
                             "ber": float(np.mean(hb != bits)),
                             "vs_fp64": float(np.mean(hb != h64))}
            rec["evm_delta_db"] = {f: rec[f]["evm_db"] - rec["fp64"]["evm_db"]
                                   for f in ("plain", "mixed")}
            out.append(rec)
            print(f"[bench] sweep {1 << Q}-QAM {snr:.0f} dB: BER fp64 {rec['fp64']['ber']:.3g} "
                  f"plain {rec['plain']['ber']:.3g} mixed {rec['mixed']['ber']:.3g}",
                  file=sys.stderr, flush=True)
    return out


def gen_y(A, cfg, seed):
    """y [S][B][N] and x [S][U][N] (cuda) for channel estimates A = H^H
    [G][B][U] (one channel per group), unit-energy QAM, AWGN at cfg's SNR."""
    import torch
    from megastation_amd import synth
    U, B, N, sg, S, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "Q"))
    G = (N + sg - 1) // sg
    Np = G * sg
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    H = torch.conj(A).resolve_conj()                               # [G][B][U]
    table = torch.from_numpy(synth.qam_table(Q)).to(dev)
    x = table[torch.randint(0, 1 << Q, (S, U, Np), device=dev, generator=gen)]
    X = x.reshape(S, U, G, sg).permute(2, 1, 0, 3).reshape(G, U, S * sg)
    Hr, Hi = H.real.contiguous(), H.imag.contiguous()
    Xr, Xi = X.real.contiguous(), X.imag.contiguous()
    Y = torch.complex(torch.bmm(Hr, Xr) - torch.bmm(Hi, Xi), torch.bmm(Hr, Xi) + torch.bmm(Hi, Xr))
    sigma = math.sqrt(10.0 ** (-cfg["snr"] / 10.0))
    Y = Y + sigma * torch.randn(Y.shape, dtype=torch.complex64, device=dev, generator=gen)
    y = Y.reshape(G, B, S, sg).permute(2, 1, 0, 3).reshape(S, B, Np)[..., :N].contiguous()
    return y, x[..., :N].contiguous()


def main_learned(args):
    """The learned form over a slot sequence of the time-varying channel:
    128 x 32 antennas x users, 3276 subcarriers in groups of 12, 64-QAM at
    25 dB, 24-tap TDL (300 ns RMS) with 100 Hz Jakes Doppler, 0.5 ms slots.
    The network is the committed checkpoint (200k slot pairs of two generator
    sequences, trained by the library's kernels, megastation_amd.train_learned;
    trained here on 200k group samples if absent); a step is one slot of a held-out
    sequence: learned ZF (Gram, network, 2 Newton-Schulz steps, residual check,
    plain-ZF fallback) + equalize/demod of 13 symbols. The plain FP32 form
    (ZF + equalize) on the same slots is timed beside it."""
    import torch
    import megastation_amd as m
    from megastation_amd import learned, synth
    cfg = dict(U=32, B=128, N=3276, sg=12, S=13, Q=6, snr=25.0)
    U, B, N, sg, S, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "Q"))
    G = (N + sg - 1) // sg
    lay = synth.standalone_layout(N, Q)
    dev = torch.device("cuda:0")
    from megastation_amd import train_learned
    t0 = time.perf_counter()
    if os.path.exists(train_learned.DEFAULT_OUT):
        # the committed checkpoint (200k generator slot pairs), SHA-256 checked
        net, tinfo = learned.load(train_learned.DEFAULT_OUT)
        training = dict(tinfo, source=os.path.relpath(train_learned.DEFAULT_OUT, ROOT))
    else:
        chans = [learned.GpuChannel(synth.TdlJakes(B, U, N, sg, seed=s)) for s in (11, 12)]
        net = learned.WarmStartNet(U, seed=0)
        hist = learned.train(net, chans, samples=200_000)
        training = {"source": "trained in this run", "train_samples": hist[-1][0],
                    "loss_first": hist[0][1], "loss_last": hist[-1][1]}
    torch.cuda.synchronize()
    training["load_or_train_seconds"] = time.perf_counter() - t0
    print(f"[bench] learned: {training}", file=sys.stderr, flush=True)
    test = learned.GpuChannel(synth.TdlJakes(B, U, N, sg, seed=99))
    n_slots = args.warmup + args.steps
    slots = []
    for k in range(n_slots):
        A = test.slot(500 + k)
        y, x = gen_y(A, cfg, seed=5000 + k)
        slots.append((A, y, x))
    llr = torch.empty(S * U * lay.ue_stride, dtype=torch.int16, device=dev)
    w = torch.empty(G * B * U, dtype=torch.complex64, device=dev)
    eq_kw = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=1,
                 mod_order=Q, ue_stride=lay.ue_stride, cb_size=lay.cb_size,
                 cb_stride=lay.cb_stride, zeros=lay.zeros, valid_bits=lay.valid_bits,
                 csi_slot_stride=G * B * U, y_sym_stride=B * N, y_slot_stride=S * B * N,
                 llr_sym_stride=U * lay.ue_stride, llr_slot_stride=S * U * lay.ue_stride)
    stream = torch.cuda.current_stream()

    def run(form):
        z = learned.LearnedZF(U, B, G, net) if form == "learned" else None
        fb = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k, (A, y, _) in enumerate(slots):
            if k == args.warmup:
                torch.cuda.synchronize()
                e0.record(stream)
            if z is not None:
                W, _ = z.step(A, stream=stream)
            else:
                W = w
                m.zf(A, W, B, U, G, stream=stream)
            m.equalize_slots(W, y, llr, stream=stream, **eq_kw)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        rate = None
        if z is not None:  # fallback share over the timed slots, in an untimed replay
            z2 = learned.LearnedZF(U, B, G, net)
            for k, (A, _, _) in enumerate(slots):
                z2.step(A, stream=stream)
                if k >= args.warmup:
                    fb.append(z2.last["flag"].cpu())
            rate = float(torch.cat(fb).float().mean())
        return ms, rate, z

    ms_plain, _, _ = run("plain")
    ms_learned, fb_rate, z = run("learned")
    # EVM of the last slot's learned W against FP64 zero-forcing's
    A, y, x = slots[-1]
    xh = torch.zeros(S * U * N, dtype=torch.complex64, device=dev)
    W, _ = z.step(A)  # one more step of the same slot sequence state
    m.equalize_slots(W, y, llr, xhat=xh, **eq_kw)
    torch.cuda.synchronize()
    x0 = x.cpu().numpy()
    evm_gpu = synth.evm_db(xh.cpu().numpy().reshape(S, U, N), x0)
    evm64 = fp64_evm(cfg, A[None], y[None], x0)
    symbols = U * N * S
    io = 8 * B * N * S + 3 * 8 * U * B * G + 2 * U * Q * N * S
    result = {
        "metric": METRIC, "value": symbols / (ms_learned * 1e-3), "unit": "symbols/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_learned,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (network bf16)", "data": "synthetic",
        "config": {"workload": "c5_learned", "antennas": B, "users": U, "subcarriers": N,
                   "sc_group": sg, "groups": G, "data_symbols": S, "qam": 1 << Q,
                   "snr_db": cfg["snr"], "channel": "TDL-24 300ns, Jakes 100 Hz, 0.5 ms slots",
                   "form": "learned", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "whole slot (learned ZF + equalize)",
                     "achieved": io / (ms_learned * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": io / (ms_learned * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": io, "avg_ms": ms_learned},
        "forms": {"learned": {"ms_per_step": ms_learned, "fallback_rate": fb_rate},
                  "plain": {"ms_per_step": ms_plain,
                            "symbols_per_s": symbols / (ms_plain * 1e-3)}},
        "training": training,
        "evm": {"gpu_db": evm_gpu, "fp64_db": evm64, "delta_db": evm_gpu - evm64,
                "within_0p1_db": abs(evm_gpu - evm64) < 0.1},
    }
    print(json.dumps(result), flush=True)


def main_precode(args):
    """Downlink precode of a 256x32 cell (mega_precode_syms): 3276 subcarriers
    in 273 groups of 12 (one B x U precoder each), 13 OFDM symbols per step,
    written at the FFT-shifted columns of 4096-point antenna rows (start 410).
    A step precodes one slot; successive steps rotate four buffer sets (about
    140 MB each) so that they are not served from the Infinity Cache. The
    CPU baseline is the oracle's single-threaded C port on one symbol."""
    import torch
    import megastation_amd as m
    B, U, N, sg, S, row, start = 256, 32, 3276, 12, 13, 4096, 410
    G = N // sg
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(77)

    def cplx(n):
        return torch.complex(torch.randn(n, device=dev, generator=gen),
                             torch.randn(n, device=dev, generator=gen))
    NR = 4
    P = [cplx(G * U * B) for _ in range(NR)]
    X = [cplx(S * U * N) for _ in range(NR)]
    O = [torch.zeros(S * B * row, dtype=torch.complex64, device=dev) for _ in range(NR)]
    stream = torch.cuda.current_stream()
    kw = dict(num_bs=B, num_ue=U, num_sc=N, sc_group=sg, bs_stride=row, bs_start=start,
              num_syms=S, stream=stream)

    def step(j):
        m.precode(P[j % NR], X[j % NR], O[j % NR], **kw)

    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for j in range(args.warmup):
        step(j)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for j in range(args.steps):
        evs[j][0].record(stream)
        step(j)
        evs[j][1].record(stream)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    k_ms = statistics.mean(a.elapsed_time(b) for a, b in evs)
    symbols = U * G * sg * S
    io = 8 * B * U * G + S * (8 * U * N + 8 * B * G * sg)
    flops = 8 * B * U * G * sg * S
    result = {
        "metric": "precoded user symbols/s (downlink precode), 1 GPU",
        "value": symbols / (ms * 1e-3), "unit": "symbols/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "d3_precode_256x32", "antennas": B, "users": U,
                   "subcarriers": N, "sc_group": sg, "groups": G, "ofdm_symbols": S,
                   "row": row, "bs_start": start, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "precode_kernel",
                     "achieved": io / (k_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": io / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("d3_precode_256x32", "precode_kernel") or
                                 {}).get("traffic_bytes"),
                     "algorithmic_bytes": io, "avg_ms": k_ms,
                     "tflops": flops / (k_ms * 1e-3) / 1e12},
    }
    if not args.no_cpu_baseline:
        import oracle
        if not os.path.exists(oracle.LIB_PATH):
            oracle.build()
        p0 = P[0].cpu().numpy().reshape(G, U, B)
        x0 = X[0][:U * N].cpu().numpy().reshape(U, N)
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            oracle.precode(p0, x0, B, U, N, sg, row, start)
            ts.append(time.perf_counter() - c0)
        sec = statistics.median(ts)
        result["cpu_baseline"] = {"value": U * G * sg / sec, "unit": "symbols/s", "cores": 1,
                                  "kind": "port",
                                  "sample": "1 of 13 symbols, median of 3, oracle C port "
                                            "(single thread)"}
    print(json.dumps(result), flush=True)


def main_fft(args):
    """Uplink FFT stage of a 256x32 cell (mega_pilot_fft, mega_uplink_fft):
    4096-point symbols of 16-bit (re, im) samples, 3276 subcarriers from bin
    offset 410; one step = one slot: three pilot symbols (32 users in groups
    of 12, users 0-11, 12-23, 24-31) into the CSI [273][256][32] in one
    launch, and 13 data symbols into y [13][256][3276] in one launch. Four rotating buffer
    sets. The CPU baseline is numpy's FFT (the oracle's) on one symbol."""
    import torch
    import megastation_amd as m
    B, U, ca, N, sg, S, P, start = 256, 32, 4096, 3276, 12, 13, 3, 410
    G = (N + sg - 1) // sg
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(99)
    NR = 4

    def samples(n):
        return torch.randint(-8000, 8000, (n,), dtype=torch.int16, device=dev, generator=gen)
    D = [samples(S * B * ca * 2) for _ in range(NR)]
    PL = [samples(P * B * ca * 2) for _ in range(NR)]
    Y = [torch.empty(S * B * N, dtype=torch.complex64, device=dev) for _ in range(NR)]
    C = [torch.empty(G * B * U, dtype=torch.complex64, device=dev) for _ in range(NR)]
    psign = torch.from_numpy(m.pilot_sign(N)).to(dev)
    stream = torch.cuda.current_stream()
    kw = dict(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, stream=stream)

    def step(j, evs=None):
        j %= NR
        if evs:
            evs[0].record(stream)
        m.pilot_fft(PL[j], C[j], psign, ue_num=U, sc_group=sg, num_syms=P, **kw)
        if evs:
            evs[1].record(stream)
        m.uplink_fft(D[j], Y[j], num_syms=S, **kw)
        if evs:
            evs[2].record(stream)

    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for j in range(args.warmup):
        step(j)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for j in range(args.steps):
        step(j, evs[j])
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    pil_ms = statistics.mean(e[0].elapsed_time(e[1]) for e in evs)
    up_ms = statistics.mean(e[1].elapsed_time(e[2]) for e in evs)
    sym = B * (S + P)  # antenna-symbols transformed per step
    up_bytes = S * B * (4 * ca + 8 * N)
    result = {
        "metric": "antenna-symbols/s through the uplink FFT stage (pilot CSI + data y), 1 GPU",
        "value": sym / (ms * 1e-3), "unit": "antenna-symbols/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (int16 samples)", "data": "synthetic",
        "config": {"workload": "u3_fft_256x32", "antennas": B, "users": U, "fft_size": ca,
                   "subcarriers": N, "sc_group": sg, "data_symbols": S, "pilot_symbols": P,
                   "ofdm_start": start, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "fft_kernel (uplink, 13 symbols)",
                     "achieved": up_bytes / (up_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": up_bytes / (up_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("u3_fft_256x32", "fft_kernel<short, 0, 12>") or
                                 {}).get("traffic_bytes"),
                     "algorithmic_bytes": up_bytes, "avg_ms": up_ms},
        "forms": {"pilot_ms": pil_ms, "uplink_ms": up_ms},
    }
    if not args.no_cpu_baseline:
        from oracle import fft as offt
        iq = D[0][:B * ca * 2].cpu().numpy()
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            offt.uplink_fft(iq, start, ca, N, B)
            ts.append(time.perf_counter() - c0)
        result["cpu_baseline"] = {"value": B / statistics.median(ts), "unit": "antenna-symbols/s",
                                  "cores": 1, "kind": "port",
                                  "sample": "1 of 13 data symbols (256 antennas), median of 3, "
                                            "numpy FFT (the oracle)"}
    print(json.dumps(result), flush=True)


def uplink_slot_samples(seed=1234):
    """One 256 x 32 slot as received int16 samples: 3 pilot symbols (pilot
    symbol s carries users 12 s .. 12 s + 11 on the subcarriers of every
    group, the channel divided by the conjugate pilot sign, as the reference's
    pilot storer expects) and 13 data symbols (64-QAM over a Kronecker
    channel at 25 dB), bins 410.. of a 4096-point spectrum, taken to the time
    domain with numpy and scaled to 40 % of int16 full scale. Returns
    (pilot samples [3][B][4096][2], data samples [13][B][4096][2], the slot
    (synth.Slot), the common scale of the FFT outputs)."""
    from megastation_amd import synth
    import megastation_amd as m
    B, U, ca, N, sg, S, P, start, Q = 256, 32, 4096, 3276, 12, 13, 3, 410, 6
    G = (N + sg - 1) // sg
    rng = np.random.Generator(np.random.Philox(seed))
    sl = synth.make_slot(rng, U, B, N, sg, S, Q, 25.0, "kronecker")
    p = m.pilot_sign(N)
    cols = (np.arange(N) + start + ca // 2) % ca
    X = np.zeros((P + S, B, ca), np.complex64)
    for s_ in range(P):
        u = s_ * sg + np.arange(N) % sg
        ok = u < U
        i = np.arange(N)[ok]
        X[s_][:, cols[i]] = (sl.H[i // sg, :, u[ok]] / np.conj(p[i])[:, None]).T
    X[P:, :, cols] = sl.y
    xt = np.fft.ifft(X, axis=2)
    gain = 0.4 / max(np.abs(xt.real).max(), np.abs(xt.imag).max())
    iq = np.empty((P + S, B, ca, 2), np.int16)
    iq[..., 0] = np.round(xt.real * (32768 * gain))
    iq[..., 1] = np.round(xt.imag * (32768 * gain))
    return iq[:P], iq[P:], sl, gain


def main_uplink(args):
    """The uplink slot of a 256 x 32 cell from received samples to LLRs: 3
    pilot symbols through mega_pilot_fft (CSI, A = H^H), mega_zf (FP32 W),
    then the 13 data symbols through the uplink data path -- FFT and
    equalize + demod (64-QAM, LLRs in the standalone layout). Forms: staged
    (mega_uplink_fft writes the slot's y [13][256][3276] into that slot's
    own buffer, then one mega_equalize_slots) and fused (mega_uplink_run: the
    same two launches through the handle's one reused y buffer, which stays
    in the Infinity Cache; LLRs equal the staged form's bit for bit). value =
    detected symbols/s of the faster FP32 form; EVM of the chain against FP64
    ZF with the true channel. Four rotating buffer sets of samples, CSI, W, y
    (staged) and LLRs."""
    import torch
    import megastation_amd as m
    from megastation_amd import synth
    B, U, ca, N, sg, S, P, start, Q = 256, 32, 4096, 3276, 12, 13, 3, 410, 6
    G = (N + sg - 1) // sg
    dev = torch.device("cuda:0")
    print("[bench] uplink_256x32: building the slot's samples", file=sys.stderr, flush=True)
    pil, dat, sl, scale = uplink_slot_samples()
    NR = 4
    PL = [torch.from_numpy(pil).to(dev) for _ in range(NR)]
    D = [torch.from_numpy(dat).to(dev) for _ in range(NR)]
    C = [torch.zeros(G * B * U, dtype=torch.complex64, device=dev) for _ in range(NR)]
    W = [torch.empty_like(C[0]) for _ in range(NR)]
    Y = [torch.empty(S * B * N, dtype=torch.complex64, device=dev) for _ in range(NR)]
    lay = synth.standalone_layout(N, Q)
    L = [torch.empty(S * U * lay.ue_stride, dtype=torch.int16, device=dev) for _ in range(NR)]
    psign = torch.from_numpy(m.pilot_sign(N)).to(dev)
    stream = torch.cuda.current_stream()
    fk = dict(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, stream=stream)
    eq = dict(num_ue=U, num_bs=B, num_sc=N, sc_group=sg, num_syms=S, num_slots=1, mod_order=Q,
              ue_stride=lay.ue_stride, cb_size=lay.cb_size, cb_stride=lay.cb_stride,
              zeros=lay.zeros, valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
              y_sym_stride=B * N, y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
              llr_slot_stride=S * U * lay.ue_stride)
    up = m.UplinkDetect(ofdm_start=start, ofdm_ca=ca, ofdm_num=N, bs_num=B, num_syms=S,
                        num_ue=U, sc_group=sg, mod_order=Q, ue_stride=lay.ue_stride,
                        cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                        valid_bits=lay.valid_bits)

    def front(j):
        m.pilot_fft(PL[j], C[j], psign, ue_num=U, sc_group=sg, num_syms=P, **fk)
        m.zf(C[j], W[j], B, U, G, stream=stream)

    def staged(j, evs=None, xhat=None):
        front(j)
        if evs:
            evs[0].record(stream)
        m.uplink_fft(D[j], Y[j], num_syms=S, **fk)
        if evs:
            evs[1].record(stream)
        m.equalize_slots(W[j], Y[j], L[j], xhat=xhat, stream=stream, **eq)

    def fused(j, evs=None, xhat=None):
        front(j)
        up.run(D[j], W[j], L[j], xhat=xhat, stream=stream)

    def timed(step):
        for j in range(args.warmup):
            step(j % NR)
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)]
               for _ in range(args.steps)]
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        for j in range(args.steps):
            step(j % NR, evs[j])
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / args.steps, evs

    ms_staged, evs = timed(staged)
    fft_ms = statistics.mean(e[0].elapsed_time(e[1]) for e in evs)
    ms_fused, _ = timed(fused)
    print(f"[bench] uplink_256x32: staged {ms_staged:.4f} ms, fused {ms_fused:.4f} ms, "
          f"data FFT {fft_ms:.4f} ms", file=sys.stderr, flush=True)
    # LLRs of the two forms, and the chain's EVM
    xh = torch.zeros(S * U * N, dtype=torch.complex64, device=dev)
    staged(0, xhat=xh)
    torch.cuda.synchronize()
    l_staged = L[0].clone()
    fused(0, xhat=xh)
    torch.cuda.synchronize()
    same = bool(torch.equal(l_staged, L[0]))
    x_chain = xh.cpu().numpy().reshape(S, U, N)
    cfg = dict(U=U, B=B, N=N, sg=sg, S=S)
    x64 = fp64_xhat(cfg, torch.from_numpy(sl.csi[None]), torch.from_numpy(sl.y[None]))[0]
    evm_gpu, evm64 = synth.evm_db(x_chain, sl.x), synth.evm_db(x64, sl.x)
    head = "fused" if ms_fused <= ms_staged else "staged"
    ms = min(ms_fused, ms_staged)
    symbols = U * N * S
    fft_bytes = S * B * (4 * ca + 8 * N)
    result = {
        "metric": "detected symbols/s from received samples (pilot FFT + ZF + data FFT + "
                  "equalize + demod), 1 GPU, at EVM within 0.1 dB of FP64",
        "value": symbols / (ms * 1e-3), "unit": "symbols/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (int16 samples)", "data": "synthetic",
        "config": {"workload": "uc3_uplink_256x32", "antennas": B, "users": U, "fft_size": ca,
                   "subcarriers": N, "sc_group": sg, "data_symbols": S, "pilot_symbols": P,
                   "qam": 1 << Q, "snr_db": 25.0, "channel": "kronecker", "form": head,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "fft_kernel (uplink, 13 symbols, staged form)",
                     "achieved": fft_bytes / (fft_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fft_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("u3_fft_256x32", "fft_kernel<short, 0, 12>") or
                                 {}).get("traffic_bytes"),
                     "algorithmic_bytes": fft_bytes, "avg_ms": fft_ms},
        "forms": {"staged": {"ms_per_step": ms_staged, "symbols_per_s": symbols / (ms_staged * 1e-3)},
                  "fused": {"ms_per_step": ms_fused, "symbols_per_s": symbols / (ms_fused * 1e-3),
                            "y_buffer_bytes": S * B * N * 8,
                            "llr_bit_identical_to_staged": same},
                  "fused_gain": ms_staged / ms_fused},
        "evm": {"gpu_db": evm_gpu, "fp64_db": evm64, "delta_db": evm_gpu - evm64,
                "within_0p1_db": abs(evm_gpu - evm64) < 0.1},
    }
    if not args.no_cpu_baseline:
        from oracle import fft as offt
        iq = dat[0].reshape(-1)
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            offt.uplink_fft(iq, start, ca, N, B)
            ts.append(time.perf_counter() - c0)
        det = cpu_baseline(dict(cfg, slots=1, Q=Q), lay,
                           torch.from_numpy(np.conj(sl.H)[None].astype(np.complex64)),
                           torch.from_numpy(sl.y[None]), None, args.cpu_seconds)
        t_slot = (S + P) * statistics.median(ts) + symbols / det["value"]
        result["cpu_baseline"] = {
            "value": symbols / t_slot, "unit": "symbols/s", "cores": det["cores"],
            "kind": "port",
            "sample": f"numpy FFT (the oracle) of one symbol of 256 antennas (median of 3, "
                      f"1 thread) times 16 symbols, plus the OpenMP detection restatement of "
                      f"the slot ({det['cores']} threads)"}
    up.close()
    print(json.dumps(result), flush=True)


def main_ldpc(args):
    """LDPC decoding of the code blocks of one 256 x 32 slot (mega_ldpc_decode):
    32 users x 13 symbols, one rate-1/3 BG1 code block each (Zc = 288, 46
    parity rows: the LDPC layout the equalizer writes for 3276 subcarriers of
    64-QAM), binary16 LLRs in 128-byte padded rows, 8 layered min-sum
    iterations. LLRs: the all-zero codeword (a codeword of every linear code;
    the decoder's work does not depend on the bits: fixed iterations, no
    early stop) through BPSK + AWGN at 1 dB. Four rotating buffer sets. The
    CPU baseline is the oracle's single-threaded C decoder on a sample of
    the code blocks."""
    import torch
    import megastation_amd as m
    from megastation_amd import synth
    U, S, bg, zc, mb, iters = 32, 13, 1, 288, 46, 8
    lay = synth.ldpc_uplink_layout(3276, 6)
    assert lay.zeros == 2 * zc and lay.cb_size == (20 + mb) * zc
    kb = 22
    n_cb = U * S
    nv = (kb + mb) * zc
    words = (kb * zc + 31) // 32
    rng = np.random.default_rng(3)
    sigma = math.sqrt(0.5 / 10 ** (1.0 / 10))
    llr = np.zeros((n_cb, lay.cb_stride), np.float16)
    llr[:, :nv] = (2 * (1.0 + sigma * rng.standard_normal((n_cb, nv))) /
                   sigma ** 2).astype(np.float16)
    llr[:, :2 * zc] = 0
    dev = torch.device("cuda:0")
    NR = 4
    L = [torch.from_numpy(llr.reshape(-1)).to(dev) for _ in range(NR)]
    O = [torch.zeros(n_cb * words, dtype=torch.int32, device=dev) for _ in range(NR)]
    stream = torch.cuda.current_stream()
    kw = dict(base_graph=bg, lifting_size=zc, parity_nodes=mb, max_iterations=iters,
              num_cb=n_cb, llr_stride=lay.cb_stride, out_stride=words, stream=stream)

    def step(j):
        m.ldpc_decode(L[j % NR], O[j % NR], **kw)
    for j in range(args.warmup):
        step(j)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for j in range(args.steps):
        step(j)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    dec = np.unpackbits(O[0].cpu().numpy().view(np.uint8).reshape(n_cb, -1), axis=1,
                        bitorder="little")[:, :kb * zc]
    ber_ch = float(np.mean(llr[:, 2 * zc:kb * zc].astype(np.float32) < 0))
    io = n_cb * (nv * 2 + words * 4)
    result = {
        "metric": "decoded information bits/s (LDPC, layered min-sum), 1 GPU",
        "value": n_cb * kb * zc / (ms * 1e-3), "unit": "bits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (binary16 LLRs)", "data": "synthetic",
        "config": {"workload": "l3_ldpc_256x32", "code_blocks": n_cb, "base_graph": bg,
                   "lifting_size": zc, "parity_rows": mb, "iterations": iters,
                   "snr_db_bpsk": 1.0, "parallelism": "1 GPU, one workgroup per code block"},
        "roofline": {"bound": "hbm", "kernel": "ldpc_decode",
                     "achieved": io / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": io / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": io, "avg_ms": ms},
        "decode": {"code_blocks_per_s": n_cb / (ms * 1e-3),
                   "channel_ber": ber_ch, "decoded_ber": float(np.mean(dec != 0))},
    }
    if not args.no_cpu_baseline:
        from oracle import ldpc as oldpc
        sample = 16
        l32 = llr[:sample, :nv].astype(np.float32)
        ts = []
        for _ in range(3):
            c0 = time.perf_counter()
            oldpc.decode(l32, bg, zc, mb, iters, 0.8125)
            ts.append(time.perf_counter() - c0)
        result["cpu_baseline"] = {"value": sample * kb * zc / statistics.median(ts),
                                  "unit": "bits/s", "cores": 1, "kind": "port",
                                  "sample": f"{sample} of {n_cb} code blocks, median of 3, "
                                            "oracle C layered min-sum (single thread)"}
    print(json.dumps(result), flush=True)


def main_batch(args):
    """The uneven 40-cell batch through the batching form (mega_batch_*): one
    hipGraph launch per step (ZF classes, then the persistent equalize
    classes), tile plans from the autotuner."""
    import torch
    import megastation_amd as m
    from megastation_amd import synth
    dev = torch.device("cuda:0")
    cfgs = batch_cells()
    print(f"[bench] c4_40cells: generating {len(cfgs)} cells", file=sys.stderr, flush=True)
    cells = []
    for k, cfg in enumerate(cfgs):
        U, B, N, sg, S, Q = (cfg[x] for x in ("U", "B", "N", "sg", "S", "Q"))
        G = (N + sg - 1) // sg
        csi, y, x, _ = gen_inputs(cfg, seed=1000 + k)
        lay = synth.standalone_layout(N, Q)
        cells.append(dict(U=U, B=B, N=N, sg=sg, S=S, Q=Q, layout=lay, csi=csi[0], y=y[0], x=x[0],
                          w=torch.empty(G * B * U, dtype=torch.complex64, device=dev),
                          llr=torch.empty(S * U * lay.ue_stride, dtype=torch.int16, device=dev)))
    t0 = time.perf_counter()
    batch = m.Batch(cells)
    plan_s = time.perf_counter() - t0
    print(batch.plan(), file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        batch.run(stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        batch.run(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # The same 40 cells through the single-cell drop-in entry points, one
    # cell after another (mega_zf, then mega_equalize_slots for its 13
    # symbols): what a caller without the batching form would launch.

    def per_cell():
        for c in cells:
            U, B, N, sg, S, Q, lay = (c[x] for x in ("U", "B", "N", "sg", "S", "Q", "layout"))
            G = (N + sg - 1) // sg
            m.zf(c["csi"], c["w"], B, U, G, stream=stream)
            m.equalize_slots(c["w"], c["y"], c["llr"], num_ue=U, num_bs=B, num_sc=N,
                             sc_group=sg, num_syms=S, num_slots=1, mod_order=Q,
                             ue_stride=lay.ue_stride, cb_size=lay.cb_size,
                             cb_stride=lay.cb_stride, zeros=lay.zeros,
                             valid_bits=lay.valid_bits, csi_slot_stride=G * B * U,
                             y_sym_stride=B * N, y_slot_stride=S * B * N,
                             llr_sym_stride=U * lay.ue_stride,
                             llr_slot_stride=S * U * lay.ue_stride, stream=stream)
    for _ in range(args.warmup):
        per_cell()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        per_cell()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_cell = e0.elapsed_time(e1) / args.steps
    symbols = sum(c["U"] * c["N"] * c["S"] for c in cells)
    io = sum(8 * c["B"] * c["N"] * c["S"] + 3 * 8 * c["U"] * c["B"] * ((c["N"] + c["sg"] - 1) // c["sg"])
             + 2 * c["U"] * c["Q"] * c["N"] * c["S"] for c in cells)
    # EVM of cell 0 (x_hat through the single-cell entry points of the same
    # W) against FP64 zero-forcing's
    c0 = cells[0]
    U, B, N, sg, S, Q = (c0[x] for x in ("U", "B", "N", "sg", "S", "Q"))
    G = (N + sg - 1) // sg
    xh = torch.zeros(S * U * N, dtype=torch.complex64, device=dev)
    llr1 = torch.empty_like(c0["llr"])
    lay = c0["layout"]
    m.equalize_slots(c0["w"], c0["y"], llr1, num_ue=U, num_bs=B, num_sc=N, sc_group=sg,
                     num_syms=S, num_slots=1, mod_order=Q, ue_stride=lay.ue_stride,
                     cb_size=lay.cb_size, cb_stride=lay.cb_stride, zeros=lay.zeros,
                     valid_bits=lay.valid_bits, csi_slot_stride=G * B * U, y_sym_stride=B * N,
                     y_slot_stride=S * B * N, llr_sym_stride=U * lay.ue_stride,
                     llr_slot_stride=S * U * lay.ue_stride, xhat=xh)
    torch.cuda.synchronize()
    x0 = c0["x"].cpu().numpy()
    evm_gpu = synth.evm_db(xh.cpu().numpy().reshape(S, U, N), x0)
    evm64 = fp64_evm(dict(U=U, B=B, N=N, sg=sg, S=S, Q=Q), c0["csi"][None], c0["y"][None], x0)
    achieved = io / (ms * 1e-3) / 1e9
    result = {
        "metric": METRIC, "value": symbols / (ms * 1e-3), "unit": "symbols/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "c4_40cells", "cells": len(cells),
                   "shapes": [f"{c['B']}x{c['U']}/N{c['N']}/sg{c['sg']}/{1 << c['Q']}QAM"
                              for c in cells],
                   "form": "batch (one hipGraph launch)", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "whole graph (ZF + equalize classes)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": io, "avg_ms": ms},
        "forms": {"batch_graph": {"ms_per_step": ms, "symbols_per_s": symbols / (ms * 1e-3)},
                  "per_cell": {"ms_per_step": ms_cell, "symbols_per_s": symbols / (ms_cell * 1e-3),
                               "schedule": "mega_zf + mega_equalize_slots per cell, cells in "
                                           "sequence on one stream"},
                  "batch_gain": ms_cell / ms},
        "plan": batch.plan(), "plan_seconds": plan_s,
        "evm": {"cell": 0, "gpu_db": evm_gpu, "fp64_db": evm64, "delta_db": evm_gpu - evm64,
                "within_0p1_db": abs(evm_gpu - evm64) < 0.1},
    }
    batch.close()
    print(json.dumps(result), flush=True)


def measured_traffic(workload, kernel_prefix):
    """HBM bytes per launch of the dominant kernel from the last rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this workload (copy-calibrated), recorded
    in profiles/traffic.json by megastation_amd.profiling; None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        data = json.load(fh).get(workload, {})
    for k, v in data.items():
        # the FP32 equalize kernels (equalize_mfma / _stream / _wide), not the
        # mixed form's
        if k.startswith(kernel_prefix) and not k.startswith("equalize_mixed"):
            return {"traffic_bytes": v["traffic_bytes"],
                    "source": f"profiles/traffic.json: {k}, rocprofv3 FETCH_SIZE/WRITE_SIZE "
                              "scaled by a 1 GiB copy"}
    return None


def cpu_baseline(cfg, lay, csi, y, x, budget_s):
    """The oracle's OpenMP restatement (FP32 Gram -> Cholesky -> W, then
    equalize + demod per group) timed on the host cores over a bounded sample
    of the same slots."""
    import oracle
    from megastation_amd import synth
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    U, B, N, sg, S, slots, Q = (cfg[k] for k in ("U", "B", "N", "sg", "S", "slots", "Q"))
    G = (N + sg - 1) // sg
    threads = len(os.sched_getaffinity(0))
    n = 1
    c_h = csi[:n].cpu().numpy()
    y_h = y[:n].cpu().numpy()
    t = time.perf_counter()
    oracle.detect_cpu(c_h, y_h, n, S, U, B, N, sg, Q, lay.ue_stride, lay.cb_size,
                      lay.cb_stride, lay.zeros, lay.valid_bits, threads)
    one = time.perf_counter() - t
    n = int(max(1, min(slots, budget_s / 6 / max(one, 1e-6))))
    c_h = csi[:n].cpu().numpy()
    y_h = y[:n].cpu().numpy()
    llr_h = np.zeros((n, S, U, lay.ue_stride), dtype=np.uint16)
    oracle.detect_cpu(c_h, y_h, n, S, U, B, N, sg, Q, lay.ue_stride, lay.cb_size,
                      lay.cb_stride, lay.zeros, lay.valid_bits, threads, llr=llr_h)
    times = []
    for _ in range(5):
        t = time.perf_counter()
        _, nt = oracle.detect_cpu(c_h, y_h, n, S, U, B, N, sg, Q, lay.ue_stride,
                                  lay.cb_size, lay.cb_stride, lay.zeros,
                                  lay.valid_bits, threads, llr=llr_h)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    base = {"value": U * N * S * n / med, "unit": "symbols/s", "cores": nt,
            "kind": "port",
            "sample": f"{n} of {slots} slots ({U * N * S * n} symbols), median of 5 "
                      f"runs after 1 warm-up, OpenMP over groups, {nt} threads"}
    return base


def fp64_xhat(cfg, csi, y):
    """x_hat of FP64 zero-forcing, complex128 [slots][S][U][N], for csi
    [slots][G][B][U] and y [slots][S][B][N] (torch, any device): textbook ZF
    in numpy (W_g = (A_g A_g^H)^-1 A_g by LAPACK solve, then x = W_g y per
    subcarrier), independent of the library. The reference every form's EVM
    and hard decisions are held to."""
    U, B, N, sg, S = (cfg[k] for k in ("U", "B", "N", "sg", "S"))
    G = (N + sg - 1) // sg
    n = csi.shape[0]
    A = csi.cpu().numpy().astype(np.complex128).reshape(n, G, B, U).transpose(0, 1, 3, 2)
    W = np.linalg.solve(A @ A.conj().transpose(0, 1, 3, 2), A)   # [n][G][U][B]
    yp = np.zeros((n, S, B, G * sg), np.complex128)
    yp[..., :N] = y.cpu().numpy().reshape(n, S, B, N)
    Y = yp.reshape(n, S, B, G, sg).transpose(0, 3, 2, 1, 4).reshape(n, G, B, S * sg)
    X = W @ Y                                                    # [n][G][U][S*sg]
    return X.reshape(n, G, U, S, sg).transpose(0, 3, 2, 1, 4).reshape(n, S, U, G * sg)[..., :N]


def fp64_evm(cfg, csi, y, x0):
    """EVM (dB) of FP64 zero-forcing of slot 0 against the transmitted
    symbols x0 [S][U][N]."""
    from megastation_amd import synth
    return synth.evm_db(fp64_xhat(cfg, csi[:1], y[:1])[0], x0)


if __name__ == "__main__":
    main()
