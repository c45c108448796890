"""ctypes binding of libdfm.so (see include/dfm.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime
torch already mapped (both carry the soname libamdhip64.so.7).  There is no
fallback: if libdfm.so is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (must be imported before the HIP library is loaded)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdfm.so")

c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p


class Grid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("n", ctypes.c_int32 * 3), ("dx", c_float)]


class Mac(ctypes.Structure):
    _fields_ = [("c", c_void_p * 3)]


class FlowMapS(ctypes.Structure):
    _fields_ = [("psi", c_void_p), ("psi_f", c_void_p * 3), ("T_f", c_void_p * 3)]


P = ctypes.POINTER

# name -> (restype, argtypes)
SIGNATURES = {
    "dfm_last_error": (ctypes.c_char_p, []),
    "dfm_version": (c_int, []),
    "dfm_reset_identity": (c_int, [P(Grid), P(FlowMapS), c_void_p]),
    "dfm_march": (c_int, [P(Grid), P(Mac), P(FlowMapS), c_float, c_int, c_void_p]),
    "dfm_march_capped": (c_int, [P(Grid), P(Mac), P(FlowMapS), c_float, c_int, c_int, c_void_p]),
    "dfm_pullback_u": (c_int, [P(Grid), P(Mac), P(FlowMapS), c_float, c_float, P(Mac), P(Mac), c_void_p]),
    "dfm_pullback_s": (c_int, [P(Grid), c_void_p, c_void_p, c_float, c_float, c_void_p, c_void_p, c_void_p]),
    "dfm_bfecc_u": (c_int, [P(Grid), P(Mac), P(FlowMapS), P(FlowMapS), P(Mac), P(Mac), c_void_p]),
    "dfm_bfecc_u_noalloc": (c_int, [P(Grid), P(Mac), P(FlowMapS), P(FlowMapS), P(Mac), c_void_p]),
    "dfm_bfecc_s": (c_int, [P(Grid), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dfm_bfecc_u_clamp": (c_int, [P(Grid), P(Mac), P(FlowMapS), P(FlowMapS), P(Mac), P(Mac), P(Mac), c_void_p]),
    "dfm_bfecc_s_clamp": (c_int, [P(Grid), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dfm_accumulate": (c_int, [P(Grid), P(Mac), P(FlowMapS), P(Mac), P(Mac), P(Mac), c_int, c_float, P(Mac),
                               c_float, c_void_p]),
    "dfm_sl_advect": (c_int, [P(Grid), P(Mac), c_float, c_int, c_int, P(Mac), P(Mac), c_int, P(c_void_p),
                              P(c_void_p), c_void_p]),
    "dfm_batch_create": (c_int, [c_void_p, c_int, c_int, c_int, ctypes.c_double, c_int, P(c_void_p)]),
    "dfm_batch_destroy": (c_int, [c_void_p]),
    "dfm_batch_field": (c_int, [c_void_p, c_int, c_int, P(c_void_p), P(c_void_p)]),
    "dfm_batch_copy": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "dfm_batch_set_option": (c_int, [c_void_p, c_int, c_int]),
    "dfm_batch_init": (c_int, [c_void_p, c_void_p]),
    "dfm_batch_forward": (c_int, [c_void_p, c_int, c_void_p]),
    "dfm_batch_init_gradient": (c_int, [c_void_p, c_void_p, c_void_p]),
    "dfm_batch_backward": (c_int, [c_void_p, c_int, c_void_p]),
    "dfm_batch_stats": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    "dfm_batch_levels": (c_int, [c_void_p, c_int, P(c_int), P(c_int)]),
    "dfm_pullback_u_k": (c_int, [P(Grid), c_int, P(Mac), P(FlowMapS), c_float, c_float, P(Mac), P(Mac), c_void_p]),
    "dfm_pullback_s_k": (c_int, [P(Grid), c_int, c_void_p, c_void_p, c_float, c_float, c_void_p, c_void_p, c_void_p]),
    "dfm_bfecc_u_k": (c_int, [P(Grid), c_int, P(Mac), P(FlowMapS), P(FlowMapS), P(Mac), P(Mac), c_void_p]),
    "dfm_bfecc_s_k": (c_int, [P(Grid), c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dfm_sl_advect_k": (c_int, [P(Grid), c_int, P(Mac), c_float, c_int, c_int, P(Mac), P(Mac), c_int, P(c_void_p),
                                P(c_void_p), c_void_p]),
    "dfm_pullback_u_vjp": (c_int, [P(Grid), c_int, P(Mac), P(Mac), P(FlowMapS), c_float, P(Mac), P(FlowMapS),
                                   c_void_p]),
    "dfm_march_vjp": (c_int, [P(Grid), P(Mac), P(FlowMapS), c_float, c_int, P(FlowMapS), P(Mac), c_void_p]),
    "dfm_pullback_s_vjp": (c_int, [P(Grid), c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p,
                                   c_void_p]),
    "dfm_sl_advect_vjp": (c_int, [P(Grid), c_int, P(Mac), c_float, c_int, c_int, P(Mac), P(Mac), P(Mac), c_int, P(c_void_p),
                                  P(c_void_p), P(c_void_p), P(Mac), c_void_p]),
    "dfm_flowmap_set_option": (c_int, [c_int, c_int]),
    "dfm_flowmap_get_option": (c_int, [c_int, P(c_int)]),
    "dfm_adj_conv_source": (c_int, [P(Grid), P(Mac), P(Mac), c_int, P(Mac), c_void_p]),
    "dfm_adj_scalar_source": (c_int, [P(Grid), c_int, P(c_void_p), P(c_void_p), P(Mac), c_void_p]),
    "dfm_axpy_mac": (c_int, [P(Grid), P(Mac), c_float, P(Mac), c_float, P(Mac), P(Mac), c_void_p]),
    "dfm_mgpcg_create": (c_int, [P(Grid), c_int, c_int, c_int, P(c_void_p)]),
    "dfm_mgpcg_destroy": (c_int, [c_void_p]),
    "dfm_mgpcg_project": (c_int, [c_void_p, P(Mac), c_double, c_int, c_int, c_void_p]),
    "dfm_mgpcg_project_begin": (c_int, [c_void_p, P(Mac), c_double, c_int, c_int, c_void_p]),
    "dfm_mgpcg_project_end": (c_int, [c_void_p]),
    "dfm_mgpcg_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_int, c_void_p]),
    "dfm_mgpcg_status": (c_int, [c_void_p, P(c_int), P(c_double)]),
    "dfm_mgpcg_pressure": (c_int, [c_void_p, P(c_void_p)]),
    "dfm_mgpcg_set_option": (c_int, [c_void_p, c_int, c_int]),
    "dfm_mgpcg_get_option": (c_int, [c_void_p, c_int, P(c_int)]),
    "dfm_mgpcg_set_mask": (c_int, [c_void_p, c_void_p, c_void_p]),
    "dfm_mask_velocity": (c_int, [P(Grid), c_void_p, c_void_p, P(Mac), c_void_p]),
    "dfm_mask_scalar": (c_int, [P(Grid), c_void_p, c_void_p, c_int, c_void_p]),
    "dfm_vortex_velocity_2d": (c_int, [P(Grid), c_void_p, c_int, P(Mac), c_void_p]),
    "dfm_vortex_velocity_2d_vjp": (c_int, [P(Grid), c_void_p, c_int, P(Mac), c_void_p, c_void_p]),
    "dfm_central_vector": (c_int, [P(Grid), P(Mac), c_void_p, c_void_p]),
    "dfm_split_central_vector": (c_int, [P(Grid), c_void_p, P(Mac), c_void_p]),
    "dfm_max_speed": (c_int, [P(Grid), P(Mac), c_void_p, c_void_p]),
    "dfm_count_nonfinite": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dfm_tile8_convert": (c_int, [P(Grid), c_void_p, c_void_p, c_int, c_void_p]),
    "dfm_poisson_operator_check": (c_int, [P(Grid), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, P(c_int),
                                           c_void_p]),
    "dfm_gaussian_force": (c_int, [P(Grid), c_void_p, c_void_p, c_void_p, c_int, P(Mac), c_void_p]),
    "dfm_gaussian_force_adjoint": (c_int, [P(Grid), c_void_p, c_void_p, c_void_p, c_int, P(Mac), c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p]),
}

MG_BF16_COARSE = 1
MG_TAIL = 2
MG_MAX_REFINE = 3
MG_CHECK_EVERY = 4
MG_GRAPH = 5
MG_BOTTOM_EXACT = 6
MG_TAIL_LEVEL = 7
MG_FUSED_SWEEPS = 8
MG_SPLIT_LAYOUT = 9
MG_BF16_LEVEL = 10
MG_LEVELS = 11
MG_ROW_KERNELS = 12

# dfm_batch fields and options
BATCH_U, BATCH_S, BATCH_TARGET, BATCH_ADJ_U, BATCH_ADJ_S = 0, 1, 2, 3, 4
BATCH_GRAPH, BATCH_FIRST_ITERS, BATCH_CHUNK, BATCH_TAIL_CELLS, BATCH_BOTTOM_EXACT, BATCH_TILE = 0, 1, 2, 3, 4, 5


class BatchInstance(ctypes.Structure):
    _fields_ = [("n", c_int * 2), ("dx", c_float), ("dt", c_float)]

DFM_E_NOCONV = -4
DFM_E_NONFINITE = -5
NOALLOC_MAX_VALUES = 1 << 23

_lib = None
_lib_lock = threading.Lock()


class DfmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libdfm.so (raises if it has not been built); thread-safe."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m dfm.build, or __graft_entry__.build())")
        lb = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lb, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lb
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().dfm_last_error().decode(errors="replace")
        raise DfmError(rc, f"{what}: {msg}" if what else msg)
    return rc


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def ptr(t):
    return None if t is None else t.data_ptr()


def make_grid(n, dx):
    g = Grid()
    g.dim = len(n)
    nn = list(n) + [1] * (3 - len(n))
    for a in range(3):
        g.n[a] = int(nn[a])
    g.dx = float(dx)
    return g


def make_mac(comps):
    m = Mac()
    for a in range(3):
        m.c[a] = comps[a].data_ptr() if a < len(comps) and comps[a] is not None else None
    return m


def make_map(psi, psi_f, T_f):
    m = FlowMapS()
    m.psi = psi.data_ptr() if psi is not None else None
    for a in range(3):
        m.psi_f[a] = psi_f[a].data_ptr() if psi_f is not None and a < len(psi_f) and psi_f[a] is not None else None
        m.T_f[a] = T_f[a].data_ptr() if T_f is not None and a < len(T_f) and T_f[a] is not None else None
    return m


def ptr_array(tensors):
    arr = (c_void_p * max(1, len(tensors)))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr
