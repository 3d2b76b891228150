import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "reference_fixtures.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture
def tuning(monkeypatch):
    """Sets MEGA_* tuning overrides for one test, tuning(MEGA_EQ_MODE=1, ...):
    the library parses them once, so it is told to re-read them now and
    again after the test has restored the environment."""
    import megastation_amd as m

    def set_(**kv):
        for k, v in kv.items():
            monkeypatch.setenv(k, str(v))
        m.tuning_reload()
    yield set_
    monkeypatch.undo()
    m.tuning_reload()
