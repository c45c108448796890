"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py): a change to the oracle that moves
its outputs fails here before it can silently move the GPU parity bar."""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _gen():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _close(a, b, rtol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


@pytest.mark.parametrize("name", ["lfm2d_32", "lfm2d_48x40", "lfm3d_16"])
def test_oracle_reproduces_lfm_fixture(name):
    gen = _gen()
    ref = np.load(os.path.join(GOLDEN, name + ".npz"))
    got = gen.lfm_case(*gen.CASES[name])
    assert set(got) == set(ref.files)
    for k in ref.files:
        assert _close(got[k], ref[k], 1e-10), k


def test_oracle_reproduces_projection_and_vortex_fixtures():
    gen = _gen()
    for fname, fn in (("projection3d_24x20x16", gen.projection_case), ("vortex2d_40x36", gen.vortex_case)):
        ref = np.load(os.path.join(GOLDEN, fname + ".npz"))
        got = fn()
        for k in ref.files:
            assert _close(got[k], ref[k], 1e-10), (fname, k)


def test_projection_fixture_is_divergence_free_and_idempotent():
    """Closed-form checks of the stored projection: zero divergence with the
    wall faces zeroed, and projecting again changes nothing."""
    from oracle import lfm_oracle as O

    ref = np.load(os.path.join(GOLDEN, "projection3d_24x20x16.npz"))
    pu = [ref["Pu_%d" % a] for a in range(3)]
    assert np.abs(O.divergence(pu)).max() < 1e-10 * max(np.abs(c).max() for c in pu)
    again = [c.copy() for c in pu]
    O.project_exact(again)
    assert max(np.abs(a - b).max() for a, b in zip(again, pu)) < 1e-10
