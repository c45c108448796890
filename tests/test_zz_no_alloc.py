"""dfm_bfecc_u_noalloc (the error field in the library's own device array)
against dfm_bfecc_u with a caller-provided error field, through the C-ABI:
bitwise equal on two calls in a row, the second on a smaller grid, plus the
size limit."""
import numpy as np
import pytest
import torch

from dfm import _lib as L
from dfm.ops import FlowMap, MacField, Ops

pytestmark = pytest.mark.gpu


def _maps_and_field(n, seed, device):
    dim = len(n)
    dx = 1.0 / n[1]
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    vel = MacField.empty(n, device)
    init = MacField.empty(n, device)
    for c in vel.c + init.c:
        c.copy_(torch.rand(c.shape, generator=gen, device=device) - 0.5)
    ops = Ops(n, dx)
    fwd, bwd = FlowMap(n, device), FlowMap(n, device)
    for m, h in ((fwd, 2.0 * dx), (bwd, -2.0 * dx)):
        ops.reset_identity(m)
        ops.march(vel, m, h, 2 if dim == 2 else 4)
    return ops, init, fwd, bwd


@pytest.mark.parametrize("sizes", [((48, 40, 32), (24, 20, 16)), ((96, 80), (32, 40))])
def test_noalloc_bfecc_matches_bfecc_twice(sizes, device):
    lib = L.lib()
    stream = L.stream_handle()
    for k, n in enumerate(sizes):
        print(f"noalloc bfecc call {k}: grid {n}", flush=True)
        ops, init, fwd, bwd = _maps_and_field(n, 10 + k, device)
        err = MacField.empty(n, device)
        want, got = MacField.empty(n, device), MacField.empty(n, device)
        L.check(lib.dfm_bfecc_u(ops.g, init.s, fwd.s, bwd.s, err.s, want.s, stream), "bfecc_u")
        L.check(lib.dfm_bfecc_u_noalloc(ops.g, init.s, fwd.s, bwd.s, got.s, stream), "bfecc_u_noalloc")
        torch.cuda.synchronize()
        for a, b in zip(got.c, want.c):
            assert torch.equal(a, b), n
        assert float(sum(c.abs().sum() for c in got.c)) > 0


def test_noalloc_bfecc_rejects_grids_beyond_the_scratch(device):
    lib = L.lib()
    # 3 * 129 * 128 * 128 = 6,340,608 face values fit; 160^3 (12.4M) does not
    big = L.make_grid((160, 160, 160), 1 / 160)
    m, f = L.FlowMapS(), L.Mac()
    assert lib.dfm_bfecc_u_noalloc(big, f, m, m, f, None) == -1  # DFM_E_ARG, before any pointer is used
    assert "DFM_NOALLOC_MAX_VALUES" in lib.dfm_last_error().decode()
    faces = sum(int(np.prod([128 + (b == a) for b in range(3)])) for a in range(3))
    assert faces <= L.NOALLOC_MAX_VALUES
