import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def _device_count():
    try:
        from tensornetworks_amd import lib
    except ImportError:
        return 0
    return lib.tn_device_count()


@pytest.fixture(scope="session")
def ctx():
    if _device_count() == 0:
        pytest.skip("no HIP device visible")
    from tensornetworks_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture
def restore_mlp(ctx):
    """For tests that load their own preconditioner network into the session
    context: the default network (init_params(0), unit normalisation) is put
    back afterwards, also when the test fails."""
    yield ctx
    import numpy as np
    from tensornetworks_amd import precond
    ctx.set_mlp(precond.init_params(0), np.array([0, 0, 0, 0, 1, 1, 1, 1], dtype=np.float32))
