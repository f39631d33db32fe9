"""The compete engine of GrainGrow_engine_compete_comp.m on the device stages.

GrainGrow_engine_compete_comp.m:9-74 with grow_unindexed_compete_comp_cuda.m:
7-49: every grain id's orientation is the EulerZXZ of its first voxel (find
order); the voxels with GrainId == 0 inside the mask are taken in find order
and in segments of nr_voxel_per_iter; each segment gets its neighbour grain
lists from the input GrainId (find_neighbor_ORs on DS_in, grs_find_neighbor_ors)
and the first maximum of C over them (cuda_forward_comp_compete,
grs_forward_comp_compete); each voxel's Completeness, GrainId, Dismedian,
Ninter, VisitFlag = 0.5, EulerZXZ, Rodrigues (angle2rod of the winner's
angles) and IPF001 (the angles over their norm) are written into the output.
"""
from __future__ import annotations

import numpy as np

from .engine import euler_to_rod_quat


def grain_grow_compete_engine(eng, ds, p1, search_radius: int = 20, nr_voxel_per_iter: int = 30000):
    """ds: an object with GrainId, Mask, Completeness, Dismedian, Ninter,
    VisitFlag (nx, ny, nz) and EulerZXZ, Rodrigues, IPF001 (3, nx, ny, nz),
    indexed [x-1, y-1, z-1]; updated in place.  p1 = [RecVolumePixel(:,1)'
    simap_data_flag VoxSize tomo_scale.Center' tomo_scale.Dimension] (11).
    Returns the number of voxels revised."""
    gid_in = ds.GrainId.copy()
    G = int(gid_in.max()) if gid_in.size else 0
    flat = gid_in.reshape(-1, order="F")
    eul_flat = ds.EulerZXZ.reshape(3, -1, order="F")
    euler = np.zeros((max(G, 1), 3))
    ids, first = np.unique(flat, return_index=True)
    for i, k in zip(ids, first):
        if 1 <= i <= G and i == int(i):
            euler[int(i) - 1] = eul_flat[:, k]
    lin = np.flatnonzero((flat == 0) & (ds.Mask.reshape(-1, order="F") == 1))
    if lin.size == 0:
        return 0
    un = np.stack(np.unravel_index(lin, gid_in.shape, order="F"), 1) + 1
    for s0 in range(0, un.shape[0], nr_voxel_per_iter):
        wait = un[s0:s0 + nr_voxel_per_iter]
        id_all, ind = eng.find_neighbor_ors(gid_in, wait, search_radius)
        vox5 = np.column_stack([wait, ind]).astype(float)
        out = eng.forward_comp_compete(vox5, id_all, euler, p1)
        x, y, z = wait[:, 0] - 1, wait[:, 1] - 1, wait[:, 2] - 1
        ds.Completeness[x, y, z] = out[:, 5]
        ds.GrainId[x, y, z] = out[:, 4]
        ds.Dismedian[x, y, z] = out[:, 7]
        ds.Ninter[x, y, z] = out[:, 6]
        ds.VisitFlag[x, y, z] = 0.5
        eu = out[:, 1:4]
        ds.EulerZXZ[:, x, y, z] = eu.T
        rod, _, _ = euler_to_rod_quat(eu)
        ds.Rodrigues[:, x, y, z] = rod.T
        with np.errstate(invalid="ignore", divide="ignore"):
            nrm = np.sqrt((eu[:, 0] * eu[:, 0] + eu[:, 1] * eu[:, 1]) + eu[:, 2] * eu[:, 2])
            ds.IPF001[:, x, y, z] = (eu / nrm[:, None]).T
    return int(un.shape[0])
