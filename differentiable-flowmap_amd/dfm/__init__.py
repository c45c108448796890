"""dfm -- MI355X-native leapfrog flow-map fluid simulation and its adjoint.

Hot path in HIP (libdfm.so, C-ABI in include/dfm.h); host orchestration in
Python mirroring the reference's simulator classes.
"""
from .ops import FlowMap, MacField, MGPCG, Ops, face_shape  # noqa: F401
from .simulator import LFM_Diff_Simulator, LFM_Simulator  # noqa: F401

__all__ = ["LFM_Simulator", "LFM_Diff_Simulator", "MGPCG", "Ops", "MacField", "FlowMap", "face_shape"]
