import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(autouse=True)
def _np_seed_guard():
    # tests that use the legacy global RNG seed it themselves; restore afterwards
    import numpy as np
    st = np.random.get_state()
    yield
    np.random.set_state(st)
