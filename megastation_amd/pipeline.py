"""Slot pipeline of one cell: ZF of slot i+1 overlapped with the equalize of
slot i on a second stream.

The reference's baseband pipeline runs the beamformer of a frame, then the
equalize calls of its symbols, one frame after another on its streams
(src/baseband/process.cc:172-185, src/scheduler/gpu_engine.cc:289-343). Here
consecutive slots of a cell overlap: push(slot i) issues the ZF of slot i on
the pipeline's ZF stream, *then* the equalize of slot i-1 on the caller's
stream, so that both kernels are resident together (issued the other way
round, the ZF waits for the equalize waves to drain). The weights are
triple-buffered; every dependency is an event: ZF(i) -> equalize(i), and
equalize(i) -> ZF(i+3), which reuses its weight buffer. Results are those of
the sequential path (same kernels on the same inputs), bit for bit.

Both forms: FORM_FP32 (mega_zf + mega_equalize_slots) and FORM_MIXED
(mega_zf_split + the split-BF16 equalize). One slot per push; csi is
[G][B][U] complex64 (A = H^H), y [S][B][N] complex64, llr the slot's LLR rows.
"""
from __future__ import annotations

from . import FORM_FP32, FORM_MIXED, equalize_slots, zf, zf_split


class SlotPipeline:
    def __init__(self, *, num_ue, num_bs, num_sc, sc_group, num_syms, mod_order, ue_stride,
                 cb_size, cb_stride, zeros, valid_bits, form: int = FORM_MIXED,
                 stream=None, depth: int = 3):
        import torch
        if form not in (FORM_FP32, FORM_MIXED):
            raise ValueError(f"unsupported form {form}")
        if depth < 2:
            raise ValueError("depth must be at least 2 (the ZF runs one slot ahead)")
        self.U, self.B, self.N, self.sg, self.S = num_ue, num_bs, num_sc, sc_group, num_syms
        self.G = (num_sc + sc_group - 1) // sc_group
        self.form = form
        self.Bp = (num_bs + 15) // 16 * 16
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.zs = torch.cuda.Stream()
        dev = torch.device("cuda", torch.cuda.current_device())
        if form == FORM_MIXED:
            self.wbuf = [torch.empty(self.G * self.U * self.Bp, dtype=torch.int64, device=dev)
                         for _ in range(depth)]
        else:
            self.wbuf = [torch.empty(self.G * self.B * self.U, dtype=torch.complex64, device=dev)
                         for _ in range(depth)]
        self.eq_kw = dict(num_ue=num_ue, num_bs=num_bs, num_sc=num_sc, sc_group=sc_group,
                          num_syms=num_syms, num_slots=1, mod_order=mod_order,
                          ue_stride=ue_stride, cb_size=cb_size, cb_stride=cb_stride,
                          zeros=zeros, valid_bits=valid_bits,
                          csi_slot_stride=(self.G * self.U * self.Bp if form == FORM_MIXED
                                           else self.G * self.B * self.U),
                          y_sym_stride=num_bs * num_sc, y_slot_stride=num_syms * num_bs * num_sc,
                          llr_sym_stride=num_ue * ue_stride,
                          llr_slot_stride=num_syms * num_ue * ue_stride)
        self.depth = depth
        self.i = 0             # slots pushed
        self.zf_ev = [None] * depth
        self.eq_done = [None] * depth
        self.pending = None    # (y, llr, xhat) of the slot whose equalize is not issued

    def _zf(self, csi, k):
        import torch
        if self.eq_done[k] is not None:
            self.zs.wait_event(self.eq_done[k])
        n = self.G
        if self.form == FORM_MIXED:
            zf_split(csi, self.wbuf[k], self.B, self.U, n, stream=self.zs)
        else:
            zf(csi, self.wbuf[k], self.B, self.U, n, stream=self.zs)
        ev = torch.cuda.Event()
        ev.record(self.zs)
        self.zf_ev[k] = ev

    def _equalize(self, k, y, llr, xhat):
        import torch
        self.stream.wait_event(self.zf_ev[k])
        equalize_slots(self.wbuf[k], y, llr, xhat=xhat, form=self.form, stream=self.stream,
                       **self.eq_kw)
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.eq_done[k] = ev

    def push(self, csi, y, llr, xhat=None):
        """Issue the ZF of this slot, then the equalize of the previous one.
        The buffers of a slot must stay valid until its equalize has run
        (after the next push or flush, on the caller's stream)."""
        k = self.i % self.depth
        self._zf(csi, k)
        if self.pending is not None:
            self._equalize((self.i - 1) % self.depth, *self.pending)
        self.pending = (y, llr, xhat)
        self.i += 1

    def flush(self):
        """Issue the equalize of the last pushed slot."""
        if self.pending is not None:
            self._equalize((self.i - 1) % self.depth, *self.pending)
            self.pending = None
