"""Slot pipeline of one cell (mega_pipeline_*): ZF of slot i+1 overlapped with
the equalize of slot i on a second stream.

The reference's baseband pipeline runs the beamformer of a frame, then the
equalize calls of its symbols, one frame after another on its streams
(src/baseband/process.cc:172-185, src/scheduler/gpu_engine.cc:289-343). Here
consecutive slots of a cell overlap: push(slot i) issues the ZF of slot i on
the pipeline's ZF stream, *then* the equalize of slot i-1 on the caller's
stream, so that both kernels are resident together (issued the other way
round, the ZF waits for the equalize waves to drain). The scheduling lives in
the library (one C call per slot; streams and events created once); weights
are kept in `depth` buffers it owns. Results are those of the sequential path
(same kernels on the same inputs), bit for bit.

PipelineGraph (mega_pipeline_graph_*) binds n consecutive slots' buffers and
captures their pipelined schedule into one graph: one launch runs all n
slots, with no host work per slot.

Both forms: FORM_FP32 (mega_zf + mega_equalize_slots) and FORM_MIXED
(mega_zf_split + the split-BF16 equalize). One slot per push by default (`slots=n`: n consecutive dense slots per push, their ZF and equalize one launch each); csi is
[G][B][U] complex64 (A = H^H), read on the pipeline's ZF stream `zs` (produce
it there, or make `zs` wait for its producer); y [S][B][N] complex64 and llr,
the slot's LLR rows ([S][U][ue_stride] halves), on the caller's stream.
"""
from __future__ import annotations

import collections
import ctypes

from . import FORM_FP32, FORM_MIXED, EqualizeDesc, _check, _ptr, _stream, lib


def _shape_desc(num_ue, num_bs, num_sc, sc_group, num_syms, mod_order, ue_stride, cb_size,
                cb_stride, zeros, valid_bits, form, slots=1):
    if slots < 1:
        raise ValueError("slots per push must be at least 1")
    d = EqualizeDesc()
    d.num_ue, d.num_bs, d.num_sc, d.sc_group = num_ue, num_bs, num_sc, sc_group
    d.num_syms, d.num_slots, d.mod_order, d.form = num_syms, slots, mod_order, form
    d.ue_stride, d.cb_size, d.cb_stride = ue_stride, cb_size, cb_stride
    d.zeros, d.valid_bits = zeros, valid_bits
    d.y_sym_stride = num_bs * num_sc
    d.llr_sym_stride = num_ue * ue_stride
    return d


def _check_form(form, depth=3):
    if form not in (FORM_FP32, FORM_MIXED):
        raise ValueError(f"unsupported form {form}")
    if depth < 2:
        raise ValueError("depth must be at least 2 (the ZF runs one slot ahead)")


class SlotPipeline:
    def __init__(self, *, num_ue, num_bs, num_sc, sc_group, num_syms, mod_order, ue_stride,
                 cb_size, cb_stride, zeros, valid_bits, form: int = FORM_MIXED,
                 stream=None, depth: int = 3, slots: int = 1):
        import torch
        _check_form(form, depth)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        d = _shape_desc(num_ue, num_bs, num_sc, sc_group, num_syms, mod_order, ue_stride,
                        cb_size, cb_stride, zeros, valid_bits, form, slots)
        self._h = ctypes.c_void_p()
        _check(lib().mega_pipeline_create(ctypes.byref(d), depth, _stream(self.stream),
                                          ctypes.byref(self._h)))
        self.zs = torch.cuda.ExternalStream(lib().mega_pipeline_zf_stream(self._h))
        self.depth = depth
        # the tensors of the last depth+1 slots stay referenced: by then a
        # slot's equalize (after its ZF) is issued on the caller's stream
        self._keep = collections.deque(maxlen=depth + 1)

    def push(self, csi, y, llr, xhat=None):
        """Issue the ZF of this slot, then the equalize of the previous one.
        The buffers of a slot must stay valid until its equalize has run
        (after the next push or flush, on the caller's stream)."""
        self._keep.append((csi, y, llr, xhat))
        _check(lib().mega_pipeline_push(self._h, _ptr(csi), _ptr(y), _ptr(llr), _ptr(xhat)))

    def flush(self):
        """Issue the equalize of the last pushed slot."""
        _check(lib().mega_pipeline_flush(self._h))

    def close(self):
        """Wait for the pipeline's work and free its buffers, streams and events."""
        if self._h is not None and self._h.value:
            _check(lib().mega_pipeline_destroy(self._h))
        self._h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PipelineGraph:
    """n consecutive pushes of one cell, buffers bound, pipelined as
    SlotPipeline would and captured into one graph; launch() runs them all.
    pushes: a list of (csi, y, llr) or (csi, y, llr, xhat) tuples of device
    tensors (xhat given for all or none), each holding `slots` dense slots;
    they stay referenced by the handle."""

    def __init__(self, pushes, *, num_ue, num_bs, num_sc, sc_group, num_syms, mod_order,
                 ue_stride, cb_size, cb_stride, zeros, valid_bits, form: int = FORM_MIXED,
                 slots: int = 1):
        _check_form(form)
        if not pushes:
            raise ValueError("no slots")
        with_x = len(pushes[0]) > 3 and pushes[0][3] is not None
        n = len(pushes)
        arr = ctypes.c_void_p * n
        self._keep = [tuple(s_) for s_ in pushes]
        csi = arr(*(_ptr(s_[0]) for s_ in pushes))
        y = arr(*(_ptr(s_[1]) for s_ in pushes))
        llr = arr(*(_ptr(s_[2]) for s_ in pushes))
        xh = arr(*(_ptr(s_[3]) for s_ in pushes)) if with_x else None
        d = _shape_desc(num_ue, num_bs, num_sc, sc_group, num_syms, mod_order, ue_stride,
                        cb_size, cb_stride, zeros, valid_bits, form, slots)
        self._h = ctypes.c_void_p()
        _check(lib().mega_pipeline_graph_create(ctypes.byref(d), n, csi, y, llr, xh,
                                                ctypes.byref(self._h)))
        self.n = n

    def launch(self, stream=None):
        """All n slots, one launch on stream (default: the current stream)."""
        _check(lib().mega_pipeline_graph_launch(self._h, _stream(stream)))

    def close(self):
        """Wait for the last launch and free the graph and its buffers."""
        if self._h is not None and self._h.value:
            _check(lib().mega_pipeline_graph_destroy(self._h))
        self._h = None
        self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
