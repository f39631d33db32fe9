"""grainrecon_amd -- MI355X-native forward-completeness scoring for LabDCT grain mapping.

The hot path of GrainRecon's grain mapping (cuda_forward_comp,
cuda_grow_indexed_region, cuda_forward_comp_compete) as a gfx950 HIP library
with a C ABI (include/grscore.h), driven from Python through ctypes.
"""
from .geometry import Geometry, cubic_hkl, euler2u, form_B, get_det_R, labdct_geometry, quaternion2U
from .engine import Engine, cuda_forward_comp, cuda_forward_comp_compete, cuda_grow_indexed_region

__all__ = [
    "Engine", "Geometry", "cubic_hkl", "euler2u", "form_B", "get_det_R", "labdct_geometry", "quaternion2U",
    "cuda_forward_comp", "cuda_forward_comp_compete", "cuda_grow_indexed_region",
]
