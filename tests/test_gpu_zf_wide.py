"""The three-kernel FP32 ZF (Gram items, per-group sweep, apply over antenna
ranges; zf_wide.hip) against the one-workgroup-per-group kernel (zf.hip) and
the FP64 oracle, through the C ABI on cuda:0.

Every arithmetic step of the three-kernel form is the one-workgroup kernel's,
in the same order, so W and the group status must be the same bits; both are
also held to the FP64 oracle's W within 1e-4 (relative Frobenius per group),
the bar of the FP32 form (src/operators/beamform/beamform.cu:74-122 computes
the same W with cuBLAS / cuSOLVER in FP32)."""
import numpy as np
import pytest

import megastation_amd as m
import oracle
from megastation_amd import synth

from gpu_helpers import dev_c64, host_c64, torch

pytestmark = pytest.mark.gpu


def _zf(a, U, B, G):
    t = torch()
    d_in = dev_c64(a)
    d_out = t.zeros_like(d_in)
    d_info = t.full((G,), -7, dtype=t.int32, device="cuda:0")
    m.zf(d_in, d_out, B, U, G, info=d_info)
    t.cuda.synchronize()
    assert np.array_equal(host_c64(d_in), a)       # out of place: input intact
    return host_c64(d_out), d_info.cpu().numpy()


def _channel(U, B, G, kind, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    H = (synth.channel_iid(rng, G, B, U) if kind == "iid"
         else synth.channel_kronecker(rng, G, B, U, 0.5, 0.3))
    return H


@pytest.mark.parametrize("U,B,G,kind", [(32, 256, 273, "kronecker"), (20, 100, 50, "iid"),
                                        (64, 128, 30, "iid"), (17, 33, 5, "iid"),
                                        (40, 70, 3, "iid"), (32, 64, 900, "iid")])
def test_three_kernel_zf_equals_one_workgroup_kernel(U, B, G, kind, tuning):
    print(f"case: zf three kernels U={U} B={B} G={G} {kind}", flush=True)
    a = np.conj(_channel(U, B, G, kind, U * 7 + B + G)).astype(np.complex64)
    tuning(MEGA_ZF_WIDE=1)
    w1, i1 = _zf(a, U, B, G)
    tuning(MEGA_ZF_WIDE=2)
    w3, i3 = _zf(a, U, B, G)
    assert np.array_equal(w1.view(np.uint32), w3.view(np.uint32))
    assert np.array_equal(i1, i3) and not i3.any()
    _, w64, info64 = oracle.zf(a, U, B, G, "f64")
    ref = w64.reshape(G, B * U)
    rel = np.linalg.norm(w3.reshape(G, B * U) - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-4, rel.max()
    assert np.array_equal(i3, info64)


def test_three_kernel_zf_failed_groups_and_in_place(tuning):
    """Rank-deficient groups (a repeated user, an all-zero channel) get W = 0
    and status k+1 for the failing pivot k, as the oracle says, in both forms;
    the Beamform handle's in-place call (its own workspace) writes the same
    W over the channel estimates."""
    print("case: zf three kernels, failed groups, in place", flush=True)
    U, B, G = 24, 64, 12
    H = _channel(U, B, G, "iid", 31)
    H[3, :, 9] = H[3, :, 2]
    H[7] = 0
    a = np.conj(H).astype(np.complex64)
    _, _, info64 = oracle.zf(a, U, B, G, "f64")
    assert info64[3] > 0 and info64[7] > 0
    tuning(MEGA_ZF_WIDE=1)
    w1, i1 = _zf(a, U, B, G)
    tuning(MEGA_ZF_WIDE=2)
    w3, i3 = _zf(a, U, B, G)
    assert np.array_equal(w1.view(np.uint32), w3.view(np.uint32))
    assert np.array_equal(i1, i3) and np.array_equal(i3, info64)
    w3 = w3.reshape(G, B * U)
    assert np.all(w3[3] == 0) and np.all(w3[7] == 0)
    t = torch()
    bf = m.Beamform(B, U, G)
    d = dev_c64(a)
    bf.beamform_uplink(d, G)
    t.cuda.synchronize()
    assert np.array_equal(host_c64(d).reshape(G, B * U).view(np.uint32), w3.view(np.uint32))
    assert np.array_equal(bf.info(), info64)
    bf.close()
