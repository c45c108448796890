import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "differentiable-flowmap_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdfm.so")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _noconv_is_an_error(request):
    """In GPU tests a projection that stops at max_iters above its tolerance
    fails the test (the library warns by default)."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    import warnings

    from dfm.ops import NoConvergenceWarning

    with warnings.catch_warnings():
        warnings.simplefilter("error", NoConvergenceWarning)
        yield


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def pytest_sessionfinish(session, exitstatus):
    """Write the bounds checked through gpu_util.within() and the values they
    saw to $DFM_PARITY_LOG (JSON), when that is set."""
    path = os.environ.get("DFM_PARITY_LOG")
    mod = sys.modules.get("gpu_util")
    if not path or mod is None:
        return
    import json

    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        json.dump(mod.PARITY, f, indent=1)
