"""The C++ drop-in classes (include/mega/) exercised by tests/cpp/test_dropin,
a restatement of the reference's own unit tests (test_matrix.cc,
test_cuphy_tensor.cc, test_modulation.cc, test_sgemv.cc, test_beamform.cc)."""
import os
import subprocess

import numpy as np
import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def _run(*args):
    assert os.path.exists(BIN), "tests/cpp/test_dropin not built: run __graft_entry__.build()"
    r = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout


def test_cpp_dropin_host():
    _run("host")


@pytest.mark.gpu
def test_cpp_dropin_gpu(golden, tmp_path):
    for k, v in golden.items():
        np.ascontiguousarray(v).tofile(tmp_path / f"{k}.bin")
    _run("gpu", str(tmp_path))
