"""The PoissonSolver-compatible entry (the interface MGPCG_3 calls,
3D/fast_mgpcg.py:192,258-273) on the GPU: the 8^3-tiled layout round trip,
the operator check, and MGPCG_3.Poisson's sequence (apply_bc, b = -div,
tile, setup/build/solve, untile, subtract_grad_p, apply_bc) against the exact
projection."""
import numpy as np
import pytest
import torch

from gpu_util import rel_l2_fields, round32

from oracle import lfm_oracle as O

pytestmark = pytest.mark.gpu


def to_tiles(a):
    """rearrange_tensor_for_cuda (3D/fast_mgpcg.py:121-145) in numpy."""
    t = [s // 8 for s in a.shape]
    return np.ascontiguousarray(a.reshape(t[0], 8, t[1], 8, t[2], 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1))


def from_tiles(f, n):
    t = [s // 8 for s in n]
    return np.ascontiguousarray(f.reshape(t[0], t[1], t[2], 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(n))


def reference_operator(n):
    """MGPCG_3.init's arrays (construct_adiag / construct_ax, 3D/fast_mgpcg.py:48-74)."""
    idx = np.meshgrid(*[np.arange(s) for s in n], indexing="ij")
    diag = np.full(n, 6.0, np.float32)
    offs = []
    for a in range(3):
        diag -= ((idx[a] == 0) | (idx[a] == n[a] - 1)).astype(np.float32)
        offs.append(np.where(idx[a] == n[a] - 1, 0.0, -1.0).astype(np.float32))
    return diag, offs, np.ones(n, np.uint8)


def dev(a):
    return torch.tensor(np.ascontiguousarray(a), device="cuda")


@pytest.mark.parametrize("n", [(32, 24, 40), (64, 64, 64)])
def test_poisson_solver_shim_projection(n, device):
    from poisson_solver_permenant_stream import PoissonSolver

    from dfm import _lib as L

    rng = np.random.default_rng(3)
    u = round32([rng.standard_normal(O.face_shape(n, a)) for a in range(3)])
    # tile layout round trip through the C-ABI, and agreement with the numpy tiling
    f = rng.standard_normal(n).astype(np.float32)
    g = L.make_grid(n, 1.0 / n[1])
    dense, tiled, back = dev(f), torch.empty(f.size, device=device), torch.empty(n, device=device)
    L.check(L.lib().dfm_tile8_convert(g, dense.data_ptr(), tiled.data_ptr(), 1, L.stream_handle()))
    L.check(L.lib().dfm_tile8_convert(g, tiled.data_ptr(), back.data_ptr(), 0, L.stream_handle()))
    assert np.array_equal(tiled.cpu().numpy(), to_tiles(f))
    assert np.array_equal(back.cpu().numpy(), f)
    # MGPCG_3.Poisson with this solver
    diag, offs, dof = reference_operator(n)
    w = [c.copy() for c in u]
    O.apply_bc(w)
    b = (-O.divergence(w)).astype(np.float32)
    x = torch.zeros(int(np.prod(n)), device=device)
    ps = PoissonSolver([s // 8 for s in n], level_num=3, bottom_smoothing=40)
    ps.setup(a_x=dev(to_tiles(offs[0])), a_y=dev(to_tiles(offs[1])), a_z=dev(to_tiles(offs[2])), b=dev(to_tiles(b)),
             a_diag=dev(to_tiles(diag)), is_dof=dev(to_tiles(dof)), x_initial=x, is_uniform=False, uniform_coef=1.0,
             neg_bc_type_vec=[1, 1, 1], pos_bc_type_vec=[1, 1, 1])
    ps.build()
    ps.solve(x)
    p = from_tiles(x.cpu().numpy(), n).astype(np.float64)
    O.subtract_grad_p(w, p)
    O.apply_bc(w)
    ref = [c.copy() for c in u]
    O.project_exact(ref)
    assert rel_l2_fields(w, ref) < 2e-6
    # the pressure itself, up to the Neumann null space (a constant)
    pe = O.neumann_solve_dct(b.astype(np.float64) - b.astype(np.float64).mean())
    assert np.linalg.norm((p - p.mean()) - pe) <= 1e-4 * np.linalg.norm(pe)
    # a second solve with the same set-up gives the same pressure
    y = torch.zeros_like(x)
    ps.solve(y)
    assert torch.equal(x, y)


def test_poisson_solver_shim_rejects_other_operators(device):
    from poisson_solver_permenant_stream import PoissonSolver

    n = (16, 16, 24)
    diag, offs, dof = reference_operator(n)
    x = torch.zeros(int(np.prod(n)), device=device)
    args = dict(a_x=dev(to_tiles(offs[0])), a_y=dev(to_tiles(offs[1])), a_z=dev(to_tiles(offs[2])),
                b=dev(np.zeros(x.numel(), np.float32)), a_diag=dev(to_tiles(diag)), is_dof=dev(to_tiles(dof)),
                x_initial=x)
    ps = PoissonSolver([2, 2, 3])
    ps.setup(**args)
    ps.build()
    bad = to_tiles(offs[1]).copy()
    bad[77] = -2.0
    ps.setup(**dict(args, a_y=dev(bad)))
    with pytest.raises(NotImplementedError, match="1 cells differ"):
        ps.build()
    nodof = to_tiles(dof).copy()
    nodof[:5] = 0
    ps.setup(**dict(args, is_dof=dev(nodof)))
    with pytest.raises(NotImplementedError, match="5 cells differ"):
        ps.build()
    with pytest.raises(NotImplementedError):
        ps.setup(**args, neg_bc_type_vec=[2, 1, 1])
    with pytest.raises(ValueError):
        ps.setup(**dict(args, b=dev(np.zeros(10, np.float32))))
