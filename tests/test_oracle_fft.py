"""The uplink FFT oracle (oracle/fft.py) against the reference's own FFT test
fixtures (test/fft/test_fft.cc) and the properties that fix its index math."""
import numpy as np

from oracle import fft as offt


def test_zadoff_chu_fixtures(golden):
    """test_fft.cc:10-54: zadoff_chu_sequence(32) (first 31 values) and
    zadoff_chu_sequence(16) shifted by pi/4, 1e-5."""
    z = offt.zadoff_chu(32)
    assert np.abs(z[:31] - golden["zadoff_chu_32"]).max() < 1e-5
    zs = offt.cyclic_shift(offt.zadoff_chu(16), np.pi / 4)
    assert np.abs(zs - golden["zadoff_chu_16_shift"]).max() < 1e-5


def test_zadoff_chu_length_rule():
    # largest prime <= n, capped at the reference table's last prime
    assert offt.zc_length(32) == 31 and offt.zc_length(16) == 13
    assert offt.zc_length(2039) == 2039 and offt.zc_length(3276) == 2039
    assert offt.zc_length(1200) == 1193
    # constant amplitude, period L
    z = offt.zadoff_chu(1200)
    assert np.allclose(np.abs(z), 1.0, atol=1e-6)
    assert np.array_equal(z[:7], z[1193:1200])


def test_uplink_fft_fixture(golden):
    """test_fft.cc:105-160: UplinkFFT(8, 32, 16, 8); the reference output is
    stored [subcarrier][antenna], 1e-5."""
    y = offt.uplink_fft(golden["fft_uplink_in"], 8, 32, 16, 8)
    ref = golden["fft_uplink_out_kb"].reshape(16, 8).T
    assert np.abs(y - ref).max() < 1e-5


def test_pilot_fft_fixture(golden):
    """test_fft.cc:56-103: PilotFFT(8, 32, 16, 8, 4, 4) with the pilot sign of
    16 subcarriers; CSI [G=4][bs=8][ue=4], 1e-4."""
    p = offt.pilot_sign(16)
    csi = offt.pilot_fft(golden["fft_pilot_in"], 8, 32, 16, 8, 4, 4, p)
    assert np.abs(csi.reshape(-1) - golden["fft_pilot_csi"]).max() < 1e-4


def test_pilot_store_recovers_channel():
    """A known channel built in the frequency domain: antenna b's pilot
    subcarrier i carries H[i/sg][b][i%sg] / conj(pilot[i]), so the stored
    CSI conj(X) * pilot must be conj(H) (the A = H^H layout ZF reads), with
    users past ue skipped and ue_start applied; int16 quantisation without
    clipping."""
    rng = np.random.default_rng(5)
    bs, ca, num, sg, ue, start = 6, 64, 40, 8, 5, 3
    G = num // sg
    p = offt.pilot_sign(num)
    H = (rng.standard_normal((G, bs, sg)) + 1j * rng.standard_normal((G, bs, sg)))
    cols = (np.arange(num) + start + ca // 2) % ca
    X = np.zeros((bs, ca), np.complex128)
    for i in range(num):
        X[:, cols[i]] = H[i // sg, :, i % sg] * np.conj(1.0 / p[i])
    x = np.fft.ifft(X, axis=1)
    q = np.round(np.stack([x.real, x.imag], -1) * 32768).astype(np.int64)
    assert np.abs(q).max() < 32767
    iq = q.astype(np.int16)
    csi = offt.pilot_fft(iq, start, ca, num, bs, ue, sg, p, ue_start=1)
    for i in range(num):
        u = 1 + i % sg
        if u >= ue:
            continue
        np.testing.assert_allclose(csi[i // sg, :, u], np.conj(H[i // sg, :, i % sg]),
                                   atol=2e-3)
    # users past ue are never written (out stays zero there)
    assert not csi[:, :, 0].any()


def test_downlink_ifft_fixture(golden):
    """test_fft.cc:162-206: DownlinkIFFT(8, 32, 16, 8), int16 output compared
    exactly."""
    out = offt.downlink_ifft(golden["ifft_in"], 32, 8)
    assert np.array_equal(out.reshape(-1), golden["ifft_out"])


def test_downlink_ifft_inverts_uplink():
    """Round trip: samples through the uplink transform (all bins) and back
    through the downlink one give the samples again (quantisation aside)."""
    rng = np.random.default_rng(2)
    bs, ca = 3, 256
    iq = rng.integers(-6000, 6000, size=(bs, ca, 2)).astype(np.int16)
    X = np.fft.fft(offt._load(iq, bs, ca), axis=1).astype(np.complex64)
    back = offt.downlink_ifft(X, ca, bs)
    assert np.abs(back.astype(np.int32) - iq.astype(np.int32)).max() <= 1
