"""The C-ABI library loads and exports every symbol include/dfm.h declares;
argument validation returns the documented error codes without touching the
device (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dfm.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import torch  # noqa: F401  binds the HIP runtime first

    from dfm import _lib as L

    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libdfm.so is not built (run __graft_entry__.build())")
    return L.lib()


def test_header_declares_expected_families():
    syms = declared_symbols()
    for want in ["dfm_march", "dfm_pullback_u", "dfm_sl_advect", "dfm_bfecc_u", "dfm_accumulate",
                 "dfm_adj_conv_source", "dfm_adj_scalar_source", "dfm_mgpcg_project", "dfm_mgpcg_create"]:
        assert want in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_header():
    from dfm import _lib as L

    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_argument_validation_codes(lib):
    from dfm import _lib as L

    g = L.make_grid((16, 16), 1 / 16)
    bad = L.make_grid((16, 16), 1 / 16)
    bad.dim = 5
    u = L.Mac()
    m = L.FlowMapS()
    assert lib.dfm_version() >= 1
    assert lib.dfm_march(None, u, m, 0.1, 2, None) == -2
    assert lib.dfm_march(bad, u, m, 0.1, 2, None) == -1
    assert "dim" in lib.dfm_last_error().decode()
    assert lib.dfm_march(g, u, m, 0.1, 2, None) == -2  # NULL velocity components
    small = L.make_grid((2, 16), 1 / 16)
    assert lib.dfm_reset_identity(small, m, None) == -1
    # the samplers' 32-bit addressing: face grids below 2^30 values, 3D (x, y)
    # planes below 2^24 (4095^2 cells -> 4096 x 4095 x-faces per plane), and
    # in 2D each extent + 1 below 2^24 (16777216 x 4: 84M values, under 2^30)
    for n in [(4095, 4095, 4), (1024, 1024, 1024), (32768, 32768), (16777216, 4), (4, 16777215)]:
        assert lib.dfm_march(L.make_grid(n, 1e-3), u, m, 0.1, 2, None) == -1, n
        assert "too large" in lib.dfm_last_error().decode()
    for n in [(4094, 4094, 4), (16777214, 4), (4, 16777214)]:
        assert lib.dfm_march(L.make_grid(n, 1e-3), u, m, 0.1, 2, None) == -2, n  # size accepted
    h = ctypes.c_void_p()
    assert lib.dfm_mgpcg_create(g, 20, 2, 50, ctypes.byref(h)) == -1  # too many levels
    odd = L.make_grid((24, 20), 1 / 24)
    assert lib.dfm_mgpcg_create(odd, 4, 2, 50, ctypes.byref(h)) == -1  # 20 not divisible by 8
    assert lib.dfm_mgpcg_project(None, u, 1e-6, 1, 10, None) == -2


def test_force_and_tiled_argument_validation(lib):
    from dfm import _lib as L

    g2 = L.make_grid((16, 16), 1 / 16)
    u = L.Mac()
    assert lib.dfm_gaussian_force(g2, None, None, None, -1, u, None) == -1  # negative count
    assert lib.dfm_gaussian_force(g2, None, None, None, 3, u, None) == -2  # NULL parameters
    assert lib.dfm_gaussian_force(g2, None, None, None, 0, u, None) == -2  # NULL force field
    g3 = L.make_grid((16, 16, 20), 1 / 16)
    assert lib.dfm_tile8_convert(g3, None, None, 1, None) == -1  # 20 is not a multiple of 8
    assert lib.dfm_tile8_convert(g2, None, None, 1, None) == -1  # 2D
    g8 = L.make_grid((16, 16, 24), 1 / 16)
    assert lib.dfm_tile8_convert(g8, None, None, 1, None) == -2
    n = ctypes.c_int(0)
    assert lib.dfm_poisson_operator_check(g8, None, None, None, None, None, 1, ctypes.byref(n), None) == -2


def test_poisson_solver_shim_host_checks():
    import torch

    from poisson_solver_permenant_stream import PoissonSolver

    ps = PoissonSolver([2, 2, 3])
    assert ps.n == (16, 16, 24)
    t = torch.zeros(16 * 16 * 24)
    kw = dict(a_x=t, a_y=t, a_z=t, b=t, a_diag=t, is_dof=t.to(torch.uint8), x_initial=t)
    with pytest.raises(NotImplementedError):
        ps.setup(**kw, neg_bc_type_vec=[1, 2, 1])
    with pytest.raises(TypeError):  # host tensors
        ps.setup(**kw)
    with pytest.raises(RuntimeError):
        ps.build()
    with pytest.raises(ValueError):
        PoissonSolver([2, 2])


def test_interp_kernel_argument_validation(lib):
    """The sampling-kernel variants accept kern 1, 2, 3 only (DFM_E_ARG, no
    device work, even when there is nothing to advect)."""
    from dfm import _lib as L

    g = L.make_grid((16, 16), 1 / 16)
    fake = ctypes.c_void_p(256)
    mac = L.Mac()
    mac.c[0] = mac.c[1] = fake.value
    m = L.FlowMapS()
    for a in range(2):
        m.psi_f[a] = m.T_f[a] = fake.value
    for kern in (0, 4, -1):
        assert lib.dfm_pullback_u_k(g, kern, mac, m, 1.0, 0.0, None, mac, None) == -1
        assert "interpolation kernel" in lib.dfm_last_error().decode()
        assert lib.dfm_pullback_s_k(g, kern, fake, fake, 1.0, 0.0, None, fake, None) == -1
        assert lib.dfm_bfecc_u_k(g, kern, mac, m, m, mac, mac, None) == -1
        assert lib.dfm_bfecc_s_k(g, kern, fake, fake, fake, fake, fake, None) == -1
        assert lib.dfm_sl_advect_k(g, kern, mac, 0.1, 2, 0, None, None, 0, None, None, None) == -1
    assert lib.dfm_sl_advect_k(g, 3, mac, 0.1, 2, 0, None, None, 0, None, None, None) == 0  # nothing to advect
