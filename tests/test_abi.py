"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, rejects bad arguments before touching the device, and its
host demod is bit-identical to the oracle."""
import ctypes
import subprocess

import numpy as np
import pytest

import megastation_amd as m


def test_library_exports_every_declared_symbol():
    L = m.lib()
    names = m.declared_symbols()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", m.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported


def test_version_and_no_error_initially():
    assert b"gfx950" in m.lib().mega_version()


def test_unsupported_modulation_is_rejected_before_launch():
    L = m.lib()
    buf = (ctypes.c_uint16 * 64)()
    rc = L.mega_equalize(4, 8, 16, ctypes.c_void_p(1), ctypes.c_void_p(1),
                         ctypes.cast(buf, ctypes.c_void_p), 32, 16, 3, 32, 32, 0,
                         32, None)
    assert rc == m.MEGA_ERR_UNSUPPORTED
    assert "modulation order 3" in m.last_error()
    assert all(v == 0 for v in buf)


def test_bad_beamform_arguments():
    L = m.lib()
    h = ctypes.c_void_p()
    assert L.mega_beamform_create(0, 4, 2, ctypes.byref(h)) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_beamform_create(256, 129, 2, ctypes.byref(h)) == m.MEGA_ERR_UNSUPPORTED
    # 65 to 128 users: the FP32 form only
    assert L.mega_zf(256, 65, 1, ctypes.c_void_p(1), ctypes.c_void_p(1), None, m.FORM_MIXED,
                     None) == m.MEGA_ERR_UNSUPPORTED
    assert "mixed form" in m.last_error()
    assert L.mega_zf(256, 129, 1, ctypes.c_void_p(1), ctypes.c_void_p(1), None, m.FORM_FP32,
                     None) == m.MEGA_ERR_UNSUPPORTED
    assert L.mega_beamform_uplink(None, None, 1, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_zf(8, 4, 1, None, None, None, 0, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_zf(8, 4, 1, ctypes.c_void_p(1), ctypes.c_void_p(1), None, 7,
                     None) == m.MEGA_ERR_UNSUPPORTED
    assert L.mega_beamform_destroy(None) == m.MEGA_OK


def test_bad_pipeline_arguments():
    L = m.lib()
    d = m.EqualizeDesc()
    d.num_ue, d.num_bs, d.num_sc, d.sc_group, d.num_syms, d.num_slots = 4, 8, 16, 4, 1, 1
    d.mod_order, d.form, d.ue_stride, d.cb_size, d.cb_stride = 4, m.FORM_MIXED, 64, 64, 64
    d.valid_bits = 64
    h = ctypes.c_void_p()
    assert L.mega_pipeline_create(ctypes.byref(d), 1, None, ctypes.byref(h)) == \
        m.MEGA_ERR_INVALID_ARG
    assert "depth" in m.last_error()
    assert L.mega_pipeline_create(ctypes.byref(d), 3, None, None) == m.MEGA_ERR_INVALID_ARG
    d.form = 7
    assert L.mega_pipeline_create(ctypes.byref(d), 3, None, ctypes.byref(h)) == \
        m.MEGA_ERR_UNSUPPORTED
    d.form, d.mod_order = m.FORM_FP32, 5
    assert L.mega_pipeline_create(ctypes.byref(d), 3, None, ctypes.byref(h)) == \
        m.MEGA_ERR_UNSUPPORTED
    d.mod_order, d.num_sc = 4, 0
    assert L.mega_pipeline_create(ctypes.byref(d), 3, None, ctypes.byref(h)) == \
        m.MEGA_ERR_INVALID_ARG
    assert h.value is None
    assert L.mega_pipeline_push(None, None, None, None, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pipeline_flush(None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pipeline_zf_stream(None) is None
    assert L.mega_pipeline_destroy(None) == m.MEGA_OK
    d.num_sc = 16
    arr = (ctypes.c_void_p * 2)(1, 1)
    none = (ctypes.c_void_p * 2)(1, None)
    assert L.mega_pipeline_graph_create(ctypes.byref(d), 0, arr, arr, arr, None,
                                        ctypes.byref(h)) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pipeline_graph_create(ctypes.byref(d), 2, arr, arr, None, None,
                                        ctypes.byref(h)) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pipeline_graph_create(ctypes.byref(d), 2, arr, none, arr, None,
                                        ctypes.byref(h)) == m.MEGA_ERR_INVALID_ARG
    assert "slot 1" in m.last_error()
    assert L.mega_pipeline_graph_launch(None, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pipeline_graph_destroy(None) == m.MEGA_OK


def test_bad_precode_arguments():
    L = m.lib()
    assert L.mega_precode(8, 65, 16, 1, 1, 1, 32, 16, 0, None) == m.MEGA_ERR_UNSUPPORTED
    assert L.mega_precode(8, 4, 16, None, 1, 1, 32, 16, 0, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_precode(8, 4, 16, 1, 1, 1, 0, 16, 0, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_precode(1 << 33, 4, 16, 1, 1, 1, 32, 16, 0, None) == m.MEGA_ERR_UNSUPPORTED
    # no full group (N < sg): nothing to write, no buffer needed
    assert L.mega_precode(8, 4, 8, None, None, None, 32, 16, 0, None) == m.MEGA_OK
    assert L.mega_precode_syms(None, None) == m.MEGA_ERR_INVALID_ARG
    assert ctypes.sizeof(m.PrecodeDesc) == 80


@pytest.mark.parametrize("n", [16, 32, 1200, 3276])
def test_pilot_sign_matches_oracle_bit_for_bit(n):
    """mega_pilot_sign (host) against the oracle's restatement of
    PilotSign::init on the same libm sincosf."""
    from oracle import fft as offt
    a = m.pilot_sign(n)
    b = offt.pilot_sign(n)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_bad_fft_arguments():
    L = m.lib()
    d = m.FftDesc(8, 32, 16, 8, 1, 16, 1, 0, 1, 0)
    assert L.mega_uplink_fft(None, None) == m.MEGA_ERR_INVALID_ARG
    d.ofdm_ca = 48  # not a power of two
    assert L.mega_uplink_fft(ctypes.byref(d), None) == m.MEGA_ERR_UNSUPPORTED
    d.ofdm_ca = 8192  # above 4096
    assert L.mega_uplink_fft(ctypes.byref(d), None) == m.MEGA_ERR_UNSUPPORTED
    d.ofdm_ca, d.ofdm_num = 32, 40  # more subcarriers than bins
    assert L.mega_uplink_fft(ctypes.byref(d), None) == m.MEGA_ERR_UNSUPPORTED
    d.ofdm_num, d.quant_bits = 16, 12
    assert L.mega_uplink_fft(ctypes.byref(d), None) == m.MEGA_ERR_UNSUPPORTED
    d.quant_bits, d.in_ = 16, None
    assert L.mega_uplink_fft(ctypes.byref(d), None) == m.MEGA_ERR_INVALID_ARG
    d.in_ = 1
    assert L.mega_pilot_fft(ctypes.byref(d), 4, 4, 0, None, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pilot_fft(ctypes.byref(d), 4, 0, 0, 1, None) == m.MEGA_ERR_INVALID_ARG
    assert L.mega_pilot_sign(4, None) == m.MEGA_ERR_INVALID_ARG
    assert ctypes.sizeof(m.FftDesc) == 56


def test_equalize_desc_layout_matches_header():
    # offsets of the C struct on LP64 (natural alignment)
    D = m.EqualizeDesc
    assert D.ue_stride.offset == 32
    assert D.csi.offset == 72
    assert D.xhat.offset == 136
    assert ctypes.sizeof(D) == 144


@pytest.mark.parametrize("Q", [2, 4, 6, 8])
def test_host_demod_matches_oracle(orc, Q):
    rng = np.random.default_rng(Q)
    x = (rng.standard_normal(4000) * 1.3 + 1j * rng.standard_normal(4000) * 1.3)
    x = np.concatenate([x, [0, 1e-9, 700 + 0j, -700j, 1e-3 - 1e-3j]]).astype(np.complex64)
    assert np.array_equal(m.demod_host(x, Q), orc.demod(x, Q))
    with pytest.raises(m.MegaError):
        m.demod_host(x[:4], 5)


def test_bad_ldpc_arguments_rejected_before_launch():
    """mega_ldpc_decode checks the base graph, lifting size, parity rows,
    iterations, LLR type, scale and strides before any device work."""
    L = m.lib()
    base = dict(base_graph=1, lifting_size=64, parity_nodes=46, max_iterations=8,
                normalization=0.0, llr_bits=32, llr=1, llr_stride=68 * 64, num_cb=1, out=1,
                out_stride=44)

    def call(**kw):
        d = m.LdpcDesc(**dict(base, **kw))
        return L.mega_ldpc_decode(ctypes.byref(d), None)
    for kw, code, msg in ((dict(base_graph=0), m.MEGA_ERR_UNSUPPORTED, "base_graph"),
                          (dict(lifting_size=17), m.MEGA_ERR_UNSUPPORTED, "lifting size 17"),
                          (dict(parity_nodes=3), m.MEGA_ERR_UNSUPPORTED, "parity_nodes"),
                          (dict(base_graph=2, parity_nodes=43), m.MEGA_ERR_UNSUPPORTED,
                           "parity_nodes"),
                          (dict(max_iterations=101), m.MEGA_ERR_UNSUPPORTED, "max_iterations"),
                          (dict(llr_bits=8), m.MEGA_ERR_UNSUPPORTED, "llr_bits"),
                          (dict(normalization=1.5), m.MEGA_ERR_INVALID_ARG, "normalization"),
                          (dict(llr=None), m.MEGA_ERR_INVALID_ARG, "NULL"),
                          (dict(llr_stride=68 * 64 - 1), m.MEGA_ERR_INVALID_ARG, "llr_stride"),
                          (dict(out_stride=43), m.MEGA_ERR_INVALID_ARG, "out_stride")):
        assert call(**kw) == code, kw
        assert msg in m.last_error(), (kw, m.last_error())
    assert call(num_cb=0) == m.MEGA_OK          # nothing to decode
