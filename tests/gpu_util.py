"""Helpers shared by the GPU parity tests (not collected by pytest)."""
import os

import numpy as np
import torch

from dfm.ops import FlowMap, MacField


# Whole-trajectory parity bounds against the FP64 oracle: forward fields at the
# FP32 forward bar (rel-L2 1e-5); adjoint fields and gradients at the same
# level, two orders inside the 1e-3 gradient bar (measured: forward <= 6e-7,
# adjoint <= 1.5e-6 on every trajectory test).
FWD_TOL = 1e-5
ADJ_TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def rel_l2_fields(a_list, b_list):
    num = sum(np.sum((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2) for a, b in zip(a_list, b_list))
    den = sum(np.sum(np.asarray(b, np.float64) ** 2) for b in b_list)
    return float(np.sqrt(num / max(den, 1e-60)))


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def mac_to_dev(arrs, dev):
    return MacField([torch.tensor(f32(a), device=dev) for a in arrs])


def mac_to_np(m):
    return [t.detach().cpu().numpy().astype(np.float64) for t in m.c]


def fmap_to_dev(om, dev, with_cells=True):
    """oracle FlowMap (AoS float64) -> device FlowMap (SoA float32)."""
    fm = FlowMap(om.n, dev, with_cells=with_cells)
    if with_cells:
        fm.psi.copy_(torch.tensor(f32(np.moveaxis(om.psi, -1, 0)), device=dev))
    for a in range(len(om.n)):
        fm.psi_f[a].copy_(torch.tensor(f32(np.moveaxis(om.psi_f[a], -1, 0)), device=dev))
        fm.T_f[a].copy_(torch.tensor(f32(np.moveaxis(om.T_f[a], -1, 0)), device=dev))
    return fm


def fmap_to_np(fm):
    """device FlowMap -> (psi, psi_f list, T_f list) in AoS float64."""
    mv = lambda t: np.moveaxis(t.detach().cpu().numpy().astype(np.float64), 0, -1)  # noqa: E731
    return (mv(fm.psi) if fm.psi is not None else None, [mv(t) for t in fm.psi_f], [mv(t) for t in fm.T_f])


def round32(x):
    """Round float64 data to float32 values (so the oracle sees the GPU's inputs)."""
    if isinstance(x, list):
        return [round32(c) for c in x]
    return np.asarray(x, dtype=np.float32).astype(np.float64)


# Every bound checked through within() is recorded with the value it saw; the
# session writes them to the JSON file named by DFM_PARITY_LOG (tests/conftest.py),
# so the measured errors behind each tolerance can be read off one GPU run.
PARITY = []


def within(tag, value, bound):
    """assert value < bound, recording (test, tag, value, bound)."""
    test = os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0]
    PARITY.append({"test": test, "tag": tag, "value": float(value), "bound": float(bound)})
    assert value < bound, f"{tag}: {value:.3e} >= {bound:.1e}"
