"""megastation_amd — MI355X (gfx950) uplink detection path of MegaStation.

Zero-forcing beamforming weights per subcarrier group and the fused
equalize + max-log demod GEMV, as hand-written HIP kernels behind the C ABI of
``lib/libmegadetect.so`` (declared in ``include/megadetect.h``).

This module is the Python mirror of the reference operator API
(``src/operators/beamform/beamform.h:27-84``,
``src/operators/sgemv/equalize.h:25-74``): ``Beamform`` with
``beamform_uplink`` / ``beamformer`` / ``beamform_downlink`` and
``Equalize.equalize`` with the reference's argument meaning. Device buffers are
torch tensors (plumbing for memory and streams); every computation runs in the
HIP library, and importing the library fails loudly when it is not built.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmegadetect.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "megadetect.h")

MEGA_OK = 0
MEGA_ERR_INVALID_ARG = -1
MEGA_ERR_UNSUPPORTED = -2
MEGA_ERR_BATCH = -3
MEGA_ERR_HIP = -4

FORM_FP32 = 0
FORM_MIXED = 1

_lib = None


class MegaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class EqualizeDesc(ctypes.Structure):
    """Mirror of ``mega_equalize_desc`` (include/megadetect.h)."""
    _fields_ = [
        ("num_ue", ctypes.c_uint32), ("num_bs", ctypes.c_uint32),
        ("num_sc", ctypes.c_uint32), ("sc_group", ctypes.c_uint32),
        ("num_syms", ctypes.c_uint32), ("num_slots", ctypes.c_uint32),
        ("mod_order", ctypes.c_uint32), ("form", ctypes.c_int32),
        ("ue_stride", ctypes.c_uint64), ("cb_size", ctypes.c_uint64),
        ("cb_stride", ctypes.c_uint64), ("zeros", ctypes.c_uint64),
        ("valid_bits", ctypes.c_uint64),
        ("csi", ctypes.c_void_p), ("csi_slot_stride", ctypes.c_uint64),
        ("y", ctypes.c_void_p), ("y_sym_stride", ctypes.c_uint64),
        ("y_slot_stride", ctypes.c_uint64),
        ("llr", ctypes.c_void_p), ("llr_sym_stride", ctypes.c_uint64),
        ("llr_slot_stride", ctypes.c_uint64),
        ("xhat", ctypes.c_void_p),
    ]


class DetectDesc(ctypes.Structure):
    """Mirror of ``mega_detect_desc`` (include/megadetect.h)."""
    _fields_ = [("eq", EqualizeDesc), ("w_out", ctypes.c_void_p),
                ("info_dev", ctypes.c_void_p), ("work", ctypes.c_void_p),
                ("route_dev", ctypes.c_void_p)]


class CellDesc(ctypes.Structure):
    """Mirror of ``mega_cell_desc`` (include/megadetect.h)."""
    _fields_ = [
        ("num_ue", ctypes.c_uint32), ("num_bs", ctypes.c_uint32),
        ("num_sc", ctypes.c_uint32), ("sc_group", ctypes.c_uint32),
        ("num_syms", ctypes.c_uint32), ("mod_order", ctypes.c_uint32),
        ("ue_stride", ctypes.c_uint64), ("cb_size", ctypes.c_uint64),
        ("cb_stride", ctypes.c_uint64), ("zeros", ctypes.c_uint64),
        ("valid_bits", ctypes.c_uint64),
        ("csi", ctypes.c_void_p), ("w", ctypes.c_void_p), ("y", ctypes.c_void_p),
        ("y_sym_stride", ctypes.c_uint64), ("llr", ctypes.c_void_p),
        ("llr_sym_stride", ctypes.c_uint64), ("xhat", ctypes.c_void_p),
        ("info_dev", ctypes.c_void_p),
    ]


def declared_symbols() -> list[str]:
    """Names of the functions include/megadetect.h declares."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mega_\w+)\s*\(",
                                 text, re.M)))


class PrecodeDesc(ctypes.Structure):
    """mega_precode_desc (include/megadetect.h)."""
    _fields_ = [("num_bs", ctypes.c_uint32), ("num_ue", ctypes.c_uint32),
                ("num_sc", ctypes.c_uint32), ("sc_group", ctypes.c_uint32),
                ("num_syms", ctypes.c_uint32), ("bs_stride", ctypes.c_uint64),
                ("bs_start", ctypes.c_uint64), ("precode", ctypes.c_void_p),
                ("uevec", ctypes.c_void_p), ("ue_sym_stride", ctypes.c_uint64),
                ("bsvec", ctypes.c_void_p), ("bs_sym_stride", ctypes.c_uint64)]


class LdpcDesc(ctypes.Structure):
    """mega_ldpc_desc (include/megadetect.h)."""
    _fields_ = [("base_graph", ctypes.c_uint32), ("lifting_size", ctypes.c_uint32),
                ("parity_nodes", ctypes.c_uint32), ("max_iterations", ctypes.c_uint32),
                ("normalization", ctypes.c_float), ("llr_bits", ctypes.c_int32),
                ("llr", ctypes.c_void_p), ("llr_stride", ctypes.c_uint64),
                ("num_cb", ctypes.c_uint32), ("out", ctypes.c_void_p),
                ("out_stride", ctypes.c_uint64)]


class FftDesc(ctypes.Structure):
    """mega_fft_desc (include/megadetect.h)."""
    _fields_ = [("ofdm_start", ctypes.c_uint32), ("ofdm_ca", ctypes.c_uint32),
                ("ofdm_num", ctypes.c_uint32), ("bs_num", ctypes.c_uint32),
                ("num_syms", ctypes.c_uint32), ("quant_bits", ctypes.c_int32),
                ("in_", ctypes.c_void_p), ("in_sym_stride", ctypes.c_uint64),
                ("out", ctypes.c_void_p), ("out_sym_stride", ctypes.c_uint64)]


def lib() -> ctypes.CDLL:
    """Load libmegadetect.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m megastation_amd.build`"
                           " (the detection path has no fallback)")
    # torch (the device memory and streams the library works on) ships its
    # own HIP and HSA runtimes under the same sonames as the system's. Loading
    # torch first makes the dynamic linker bind this library to torch's
    # copies, so the process holds one HIP runtime; loading this library
    # first would bind torch to the system copies instead.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    P, u8, u32, u64, i32 = (ctypes.c_void_p, ctypes.c_uint8, ctypes.c_uint32,
                            ctypes.c_uint64, ctypes.c_int)
    sig = {
        "mega_version": ([], ctypes.c_char_p),
        "mega_last_error": ([], ctypes.c_char_p),
        "mega_tuning_reload": ([], i32),
        "mega_beamform_create": ([u32, u32, u32, ctypes.POINTER(P)], i32),
        "mega_beamform_destroy": ([P], i32),
        "mega_beamform_uplink": ([P, P, u32, P], i32),
        "mega_beamformer": ([P, P, P, u32, P], i32),
        "mega_beamform_downlink": ([P, P, u32, u32, u32, P], i32),
        "mega_beamform_info": ([P, P, u32, P], i32),
        "mega_zf": ([u32, u32, u32, P, P, P, i32, P], i32),
        "mega_zf_split": ([u32, u32, u32, P, P, P, P, P, P], i32),
        "mega_equalize": ([u64, u64, u64, P, P, P, u64, u64, u8, u64, u64, u64,
                           u64, P], i32),
        "mega_equalize_slots": ([ctypes.POINTER(EqualizeDesc), P], i32),
        "mega_detect_slots": ([ctypes.POINTER(DetectDesc), P], i32),
        "mega_detect_slots_noalloc": ([ctypes.POINTER(DetectDesc), P], i32),
        "mega_detect_work_bytes": ([ctypes.POINTER(DetectDesc)], u64),
        "mega_demod_host": ([P, u64, u8, P], i32),
        "mega_precode": ([u64, u64, u64, P, P, P, u64, u64, u64, P], i32),
        "mega_uplink_fft": ([ctypes.POINTER(FftDesc), P], i32),
        "mega_uplink_create": ([ctypes.POINTER(FftDesc), ctypes.POINTER(EqualizeDesc), u32,
                                ctypes.POINTER(ctypes.c_void_p)], i32),
        "mega_uplink_run": ([P, P, P, P, P, P], i32),
        "mega_uplink_destroy": ([P], i32),
        "mega_pilot_fft": ([ctypes.POINTER(FftDesc), u32, u32, u32, P, P], i32),
        "mega_pilot_sign": ([u32, P], i32),
        "mega_downlink_ifft": ([ctypes.POINTER(FftDesc), P], i32),
        "mega_precode_syms": ([ctypes.POINTER(PrecodeDesc), P], i32),
        "mega_ldpc_decode": ([ctypes.POINTER(LdpcDesc), P], i32),
        "mega_batch_create": ([ctypes.POINTER(CellDesc), u32, ctypes.c_char_p,
                               ctypes.POINTER(P)], i32),
        "mega_batch_run": ([P, P], i32),
        "mega_batch_plan": ([P], ctypes.c_char_p),
        "mega_batch_destroy": ([P], i32),
        "mega_pipeline_create": ([ctypes.POINTER(EqualizeDesc), u32, P, ctypes.POINTER(P)], i32),
        "mega_pipeline_push": ([P, P, P, P, P], i32),
        "mega_pipeline_flush": ([P], i32),
        "mega_pipeline_zf_stream": ([P], P),
        "mega_pipeline_destroy": ([P], i32),
        "mega_pipeline_graph_create": ([ctypes.POINTER(EqualizeDesc), u32, ctypes.POINTER(P),
                                        ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                                        ctypes.POINTER(P)], i32),
        "mega_pipeline_graph_launch": ([P, P], i32),
        "mega_pipeline_graph_destroy": ([P], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise MegaError(rc, lib().mega_last_error().decode())
    return rc


def last_error() -> str:
    return lib().mega_last_error().decode()


def tuning_reload() -> None:
    """Make changed MEGA_* tuning variables take effect (the library parses
    them once and keeps them)."""
    lib().mega_tuning_reload()


def _stream(stream):
    """hipStream_t handle of a torch stream (default: the current stream)."""
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(getattr(stream, "cuda_stream", stream) or 0)


def _ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    return ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# Beamform (src/operators/beamform/beamform.h:27-84)
# ---------------------------------------------------------------------------

class Beamform:
    """ZF beamformer for ``batch_count`` subcarrier groups of a
    ``num_bs`` x ``num_ue`` cell. Buffers are complex64 device tensors in the
    ``[G][B][U]`` CSI layout (precoder ``[G][U][B]``)."""

    def __init__(self, num_bs: int, num_ue: int, batch_count: int):
        self.num_bs, self.num_ue, self.batch_count = num_bs, num_ue, batch_count
        h = ctypes.c_void_p()
        _check(lib().mega_beamform_create(num_bs, num_ue, batch_count, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mega_beamform_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def beamform_uplink(self, csi, batch: int | None = None, stream=None):
        """In place: csi_g <- (A_g A_g^H)^-1 A_g."""
        batch = csi.shape[0] if batch is None else batch
        _check(lib().mega_beamform_uplink(self._h, _ptr(csi), batch, _stream(stream)))

    def beamformer(self, csi, precode, batch: int | None = None, stream=None):
        batch = csi.shape[0] if batch is None else batch
        _check(lib().mega_beamformer(self._h, _ptr(csi), _ptr(precode), batch,
                                     _stream(stream)))

    def info(self, batch: int | None = None, stream=None):
        """Per-group Cholesky status of the last call (synchronises)."""
        import numpy as np
        batch = self.batch_count if batch is None else batch
        out = np.zeros(batch, dtype=np.int32)
        _check(lib().mega_beamform_info(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                        batch, _stream(stream)))
        return out

    @staticmethod
    def beamform_downlink(csi, precode, num_ue: int, num_bs: int, batch: int,
                          stream=None):
        _check(lib().mega_beamform_downlink(_ptr(csi), _ptr(precode), num_ue,
                                            num_bs, batch, _stream(stream)))


def zf(csi_in, w_out, num_bs: int, num_ue: int, batch: int, info=None,
       form: int = FORM_FP32, stream=None):
    """Out-of-place ZF (w_out may be csi_in)."""
    _check(lib().mega_zf(num_bs, num_ue, batch, _ptr(csi_in), _ptr(w_out),
                         _ptr(info), form, _stream(stream)))


def zf_split(csi_in, w_split, num_bs: int, num_ue: int, batch: int, w_out=None,
             info=None, route=None, stream=None):
    """Mixed-form ZF: split BF16 weight words (8 bytes per weight, CSI layout)
    for equalize_slots(form=FORM_MIXED, csi=w_split)."""
    _check(lib().mega_zf_split(num_bs, num_ue, batch, _ptr(csi_in), _ptr(w_split),
                               _ptr(w_out), _ptr(info), _ptr(route), _stream(stream)))


# ---------------------------------------------------------------------------
# Equalize (src/operators/sgemv/equalize.h:25-74)
# ---------------------------------------------------------------------------

class Equalize:
    @staticmethod
    def equalize_raw(row, col, batch_count, csi, bsvec, uevec, ue_stride,
                     csi_skip, mod_order, cb_size, cb_stride, zeros, valid_bits,
                     stream=None):
        """equalize.cu:20-51 argument order: U, B, N, W, y, llr, ..."""
        _check(lib().mega_equalize(row, col, batch_count, _ptr(csi), _ptr(bsvec),
                                   _ptr(uevec), ue_stride, csi_skip, mod_order,
                                   cb_size, cb_stride, zeros, valid_bits,
                                   _stream(stream)))

    @staticmethod
    def equalize(csi, bsvec, uevec, csi_skip, mod_order, cb_size, zeros,
                 cb_blocks, stream=None):
        """Matrix overload (equalize.cu:53-67) for torch tensors:
        csi [G][B][U] complex64, bsvec [B][N] complex64, uevec
        [U*cb_blocks][cb_stride] float16 (cb_stride = uevec row stride)."""
        U, B = csi.shape[-1], csi.shape[-2]
        N = bsvec.shape[-1]
        cb_stride = uevec.stride(0)
        ue_stride = cb_stride * cb_blocks
        Equalize.equalize_raw(U, B, N, csi, bsvec, uevec, ue_stride, csi_skip,
                              mod_order, cb_size, cb_stride, zeros,
                              cb_size * cb_blocks, stream)


def equalize_slots(csi, y, llr, *, num_ue, num_bs, num_sc, sc_group, num_syms,
                   num_slots, mod_order, ue_stride, cb_size, cb_stride, zeros,
                   valid_bits, csi_slot_stride, y_sym_stride, y_slot_stride,
                   llr_sym_stride, llr_slot_stride, xhat=None,
                   form: int = FORM_FP32, stream=None):
    d = EqualizeDesc(num_ue, num_bs, num_sc, sc_group, num_syms, num_slots,
                     mod_order, form, ue_stride, cb_size, cb_stride, zeros,
                     valid_bits, _ptr(csi).value, csi_slot_stride, _ptr(y).value,
                     y_sym_stride, y_slot_stride, _ptr(llr).value, llr_sym_stride,
                     llr_slot_stride, _ptr(xhat).value)
    _check(lib().mega_equalize_slots(ctypes.byref(d), _stream(stream)))


def precode(prec, ue, out, *, num_bs, num_ue, num_sc, sc_group, bs_stride, bs_start,
            num_syms: int = 1, ue_sym_stride: int = 0, bs_sym_stride: int = 0, stream=None):
    """Downlink precode (mega_precode_syms): prec [G][U][B] complex64, ue
    [S][U][N], out [S][B][bs_stride]; symbol t at ue + t*ue_sym_stride and
    out + t*bs_sym_stride (complex elements; default: dense)."""
    d = PrecodeDesc(num_bs, num_ue, num_sc, sc_group, num_syms, bs_stride, bs_start,
                    _ptr(prec).value, _ptr(ue).value, ue_sym_stride or num_ue * num_sc,
                    _ptr(out).value, bs_sym_stride or num_bs * bs_stride)
    _check(lib().mega_precode_syms(ctypes.byref(d), _stream(stream)))


def _fft_desc(iq, out, ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, in_sym_stride,
              out_sym_stride):
    import torch
    bits = {torch.int16: 16, torch.int8: 8}.get(getattr(iq, "dtype", None))
    if bits is None:
        raise ValueError("samples must be an int16 or int8 tensor of (re, im) pairs")
    return FftDesc(ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, bits, _ptr(iq).value,
                   in_sym_stride, _ptr(out).value, out_sym_stride)


def uplink_fft(iq, y, *, ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms: int = 1,
               in_sym_stride: int = 0, out_sym_stride: int = 0, stream=None):
    """UplinkFFT (mega_uplink_fft): iq [S][B][ca][2] int16/int8 -> y
    [S][B][ofdm_num] complex64."""
    d = _fft_desc(iq, y, ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, in_sym_stride,
                  out_sym_stride)
    _check(lib().mega_uplink_fft(ctypes.byref(d), _stream(stream)))


class UplinkDetect:
    """The uplink data path of one slot (mega_uplink_*): received samples
    [S][B][ofdm_ca] (int16 or int8 pairs) -> FFT -> equalize + demod -> LLRs
    through y buffers the handle owns: one FFT and one equalize launch over
    a reused slot of y (chunk_syms 0), or chunks of chunk_syms symbols
    through a two-chunk ring, the FFT of one chunk beside the equalize of the
    previous. The equalize shape is mega_equalize_slots' for one slot; form
    FP32 takes W in the CSI layout, MIXED mega_zf_split's planes."""

    def __init__(self, *, ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, quant_bits=16,
                 in_sym_stride=0, num_ue, sc_group, mod_order, ue_stride, cb_size, cb_stride,
                 zeros, valid_bits, llr_sym_stride=0, form: int = FORM_FP32,
                 chunk_syms: int = 0):
        f = FftDesc(ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, quant_bits, None,
                    in_sym_stride, None, 0)
        e = EqualizeDesc(num_ue, bs_num, ofdm_num, sc_group, num_syms, 1, mod_order, form,
                         ue_stride, cb_size, cb_stride, zeros, valid_bits, None, 0, None, 0, 0,
                         None, llr_sym_stride, 0, None)
        h = ctypes.c_void_p()
        _check(lib().mega_uplink_create(ctypes.byref(f), ctypes.byref(e), chunk_syms,
                                        ctypes.byref(h)))
        self._h = h

    def run(self, samples, w, llr, xhat=None, stream=None):
        _check(lib().mega_uplink_run(self._h, _ptr(samples), _ptr(w), _ptr(llr), _ptr(xhat),
                                     _stream(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib().mega_uplink_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pilot_fft(iq, csi, pilot, *, ofdm_start, ofdm_ca, ofdm_num, bs_num, ue_num, sc_group,
              ue_start: int = 0, num_syms: int = 1, stream=None):
    """PilotFFT (mega_pilot_fft): pilot symbols iq [S][B][ca][2] -> CSI
    [G][B][U] complex64 (A = H^H), symbol s carrying users ue_start +
    s*sc_group + (i % sc_group); pilot: device complex64 [ofdm_num]."""
    d = _fft_desc(iq, csi, ofdm_start, ofdm_ca, ofdm_num, bs_num, num_syms, 0, 0)
    _check(lib().mega_pilot_fft(ctypes.byref(d), ue_num, sc_group, ue_start, _ptr(pilot),
                                _stream(stream)))


def downlink_ifft(x, iq, *, ofdm_ca, bs_num, num_syms: int = 1, in_sym_stride: int = 0,
                  out_sym_stride: int = 0, stream=None):
    """DownlinkIFFT (mega_downlink_ifft): complex64 rows x [S][B][ca] ->
    int16/int8 (re, im) pairs iq [S][B][ca][2]."""
    d = _fft_desc(iq, None, 0, ofdm_ca, ofdm_ca, bs_num, num_syms, in_sym_stride,
                  out_sym_stride)
    d.in_, d.out = _ptr(x).value, _ptr(iq).value
    _check(lib().mega_downlink_ifft(ctypes.byref(d), _stream(stream)))


def ldpc_decode(llr, out, *, base_graph, lifting_size, parity_nodes, max_iterations, num_cb,
                llr_stride, out_stride, normalization: float = 0.0, stream=None):
    """LDPC decoding of num_cb code blocks (mega_ldpc_decode): llr a float16 or
    float32 device tensor, out an int32 tensor of bit words."""
    import torch
    bits = 16 if llr.dtype in (torch.float16, torch.int16) else 32
    d = LdpcDesc(base_graph, lifting_size, parity_nodes, max_iterations, normalization, bits,
                 _ptr(llr).value, llr_stride, num_cb, _ptr(out).value, out_stride)
    _check(lib().mega_ldpc_decode(ctypes.byref(d), _stream(stream)))


def pilot_sign(ofdm_num: int):
    """PilotSign::init's table (mega_pilot_sign, host): numpy complex64."""
    import numpy as np
    out = np.zeros(2 * ofdm_num, np.float32)
    _check(lib().mega_pilot_sign(ofdm_num, out.ctypes.data_as(ctypes.c_void_p)))
    return out.view(np.complex64)


def detect_fused_shape(num_ue: int, num_bs: int) -> bool:
    """Whether mega_detect_slots runs one fused kernel for this cell shape
    (otherwise it runs ZF then equalize through w_out)."""
    return num_ue <= 16 and num_bs <= 256


def detect_slots(csi, y, llr, *, num_ue, num_bs, num_sc, sc_group, num_syms,
                 num_slots, mod_order, ue_stride, cb_size, cb_stride, zeros,
                 valid_bits, csi_slot_stride, y_sym_stride, y_slot_stride,
                 llr_sym_stride, llr_slot_stride, xhat=None, w_out=None,
                 info=None, form: int = FORM_FP32, work=None, route=None,
                 stream=None, noalloc: bool = False):
    """Fused ZF + equalize + demod: csi holds the channel estimates A = H^H.
    form=FORM_MIXED needs `work` (detect_work_bytes(...) bytes of device
    memory, e.g. a torch uint8 tensor); `route` (int32, num_slots*G) gets 1
    for groups computed on the FP64 route. noalloc=True calls
    mega_detect_slots_noalloc instead: no `work` (the library's static
    scratch of DETECT_SCRATCH_BYTES per GPU holds the split weights)."""
    e = EqualizeDesc(num_ue, num_bs, num_sc, sc_group, num_syms, num_slots,
                     mod_order, form, ue_stride, cb_size, cb_stride, zeros,
                     valid_bits, _ptr(csi).value, csi_slot_stride, _ptr(y).value,
                     y_sym_stride, y_slot_stride, _ptr(llr).value, llr_sym_stride,
                     llr_slot_stride, _ptr(xhat).value)
    d = DetectDesc(e, _ptr(w_out).value, _ptr(info).value, _ptr(work).value,
                   _ptr(route).value)
    fn = lib().mega_detect_slots_noalloc if noalloc else lib().mega_detect_slots
    _check(fn(ctypes.byref(d), _stream(stream)))


DETECT_SCRATCH_BYTES = 32 << 20  # MEGA_DETECT_SCRATCH_BYTES


def detect_work_bytes(*, num_ue, num_bs, num_sc, sc_group, num_slots,
                      form: int = FORM_FP32) -> int:
    """Device scratch bytes detect_slots needs (mega_detect_work_bytes)."""
    e = EqualizeDesc()
    e.num_ue, e.num_bs, e.num_sc, e.sc_group = num_ue, num_bs, num_sc, sc_group
    e.num_syms, e.num_slots, e.form = 1, num_slots, form
    d = DetectDesc()
    d.eq = e
    return int(lib().mega_detect_work_bytes(ctypes.byref(d)))


class Batch:
    """The batching form (mega_batch_*): many cells of uneven shapes planned,
    autotuned and captured into one hipGraph; run() replays it.

    cells: dicts with U, B, N, sg, S, Q, layout (synth.LlrLayout) and torch
    device tensors csi ([G][B][U] complex64, A = H^H), w (same size), y
    ([S][B][N]), llr (int16, S*U*ue_stride) and optionally xhat
    ([S][U][N] complex64) and info (int32, G)."""

    def __init__(self, cells, tune_path: str | None = None):
        self._keep = cells
        arr = (CellDesc * len(cells))()
        for d, c in zip(arr, cells):
            lay = c["layout"]
            d.num_ue, d.num_bs, d.num_sc = c["U"], c["B"], c["N"]
            d.sc_group, d.num_syms, d.mod_order = c["sg"], c["S"], c["Q"]
            d.ue_stride, d.cb_size, d.cb_stride = lay.ue_stride, lay.cb_size, lay.cb_stride
            d.zeros, d.valid_bits = lay.zeros, lay.valid_bits
            d.csi, d.w, d.y = _ptr(c["csi"]).value, _ptr(c["w"]).value, _ptr(c["y"]).value
            d.y_sym_stride = c["B"] * c["N"]
            d.llr = _ptr(c["llr"]).value
            d.llr_sym_stride = c["U"] * lay.ue_stride
            d.xhat = _ptr(c.get("xhat")).value
            d.info_dev = _ptr(c.get("info")).value
        self._h = ctypes.c_void_p()
        _check(lib().mega_batch_create(arr, len(cells),
                                       tune_path.encode() if tune_path else None,
                                       ctypes.byref(self._h)))

    def run(self, stream=None):
        _check(lib().mega_batch_run(self._h, _stream(stream)))

    def plan(self) -> str:
        return lib().mega_batch_plan(self._h).decode()

    def close(self):
        if self._h:
            lib().mega_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def demod_host(xhat, mod_order: int):
    """Host demod through the library (numpy complex64 in, uint16 LLR bits
    out, shape xhat.shape + (Q,))."""
    import numpy as np
    x = np.require(np.asarray(xhat, dtype=np.complex64), requirements="C")
    out = np.zeros(x.shape + (mod_order,), dtype=np.uint16)
    _check(lib().mega_demod_host(x.ctypes.data_as(ctypes.c_void_p), x.size,
                                 mod_order, out.ctypes.data_as(ctypes.c_void_p)))
    return out
