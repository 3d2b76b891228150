"""CPU checks of the comparison helpers the GPU parity tests rely on."""
import numpy as np

from gpu_helpers import assert_imag_counts, rel_fro


def test_rel_fro_counts_imaginary_parts():
    rng = np.random.default_rng(3)
    z = rng.standard_normal((6, 5)) + 1j * rng.standard_normal((6, 5))
    assert rel_fro(z, z) == 0.0
    # the conjugate differs only in its imaginary parts: a comparison that
    # dropped them would report 0 here
    assert rel_fro(np.conj(z), z) > 0.5
    w = z.copy()
    w.imag += 1e-3
    assert 1e-4 < rel_fro(w, z) < 1e-2
    assert rel_fro(z.astype(np.complex64), z) < 1e-6
    assert_imag_counts(z)


def test_rel_fro_real_inputs():
    a = np.array([3.0, 4.0])
    assert rel_fro(a + np.array([0.0, 0.5]), a) == 0.1
    assert rel_fro(np.zeros(2), np.zeros(2)) == 0.0


def test_hard_decisions_match_demod_signs(orc):
    """synth.hard_bits_xhat (float64 thresholds, used for BER) gives every
    constellation point its own bits and agrees with the sign of the
    oracle's binary16 demod except where binary16 rounding of an LLR within
    1/8 of the boundary can flip it; hard_bits_llr reads LLR signs."""
    from megastation_amd import synth
    for Q in (2, 4, 6, 8):
        t = synth.qam_table(Q)
        assert np.array_equal(synth.hard_bits_xhat(t, Q),
                              synth.index_bits(np.arange(1 << Q), Q))
        rng = np.random.default_rng(Q)
        z = (rng.standard_normal(20000) + 1j * rng.standard_normal(20000)).astype(np.complex64)
        llr = orc.demod(z, Q).reshape(-1, Q)
        bad = synth.hard_bits_llr(llr) != synth.hard_bits_xhat(z, Q)
        mag = np.abs(llr.view(np.float16).astype(np.float64))
        assert not (bad & (mag > 0.125)).any(), Q
    assert np.array_equal(synth.hard_bits_llr(np.array([0x3C00, 0xBC00, 0x0000, 0x8000])),
                          [0, 1, 0, 1])


def test_learned_checkpoint_roundtrip_and_tamper(tmp_path):
    """The learned form's checkpoint format: arrays and info come back
    exactly; a changed byte of any weight, a changed info field or a wrong
    shape is refused by the SHA-256 / shape check (CPU only)."""
    import pytest
    from megastation_amd import learned
    rng = np.random.default_rng(0)
    dims = [12, 8, 8, 6]
    arrs = {f"w{l}": rng.standard_normal((dims[l], dims[l + 1])).astype(np.float32)
            for l in range(3)} | {f"b{l}": rng.standard_normal(dims[l + 1]).astype(np.float32)
                                 for l in range(3)}
    p = str(tmp_path / "net.npz")
    dig = learned.write_checkpoint(p, arrs, {"U": 2, "dims": dims, "slots": 5})
    got, info = learned.read_checkpoint(p)
    assert info == {"U": 2, "dims": dims, "slots": 5} and len(dig) == 64
    for k in arrs:
        assert np.array_equal(got[k], arrs[k])
    # altered weights with the old digest
    with np.load(p, allow_pickle=False) as z:
        parts = {k: z[k] for k in z.files}
    parts["w1"] = parts["w1"].copy()
    parts["w1"].view(np.uint32)[0, 0] ^= 1
    bad = str(tmp_path / "bad.npz")
    np.savez(bad, **parts)
    with pytest.raises(learned.CheckpointError):
        learned.read_checkpoint(bad)
    # altered info with the old digest
    parts["w1"].view(np.uint32)[0, 0] ^= 1
    parts["meta"] = np.array('{"U": 3, "dims": [12, 8, 8, 6], "slots": 5}')
    np.savez(bad, **parts)
    with pytest.raises(learned.CheckpointError):
        learned.read_checkpoint(bad)
    # shapes that do not match dims
    learned.write_checkpoint(bad, arrs, {"U": 2, "dims": [12, 8, 9, 6]})
    with pytest.raises(learned.CheckpointError):
        learned.read_checkpoint(bad)
