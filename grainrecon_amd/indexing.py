"""The indexing engine of GrainIndexing_engine_gpu_comp.m on the device pipeline.

GrainIndexing_engine_gpu_comp.m:40-124 visits the seed positions in order;
a seed whose voxel has Completeness < TrustComp and VisitFlag ~= 1 is indexed
(ForwardCalc_OR_space_comp_cuda, sortrows, fit_OR_constrain_with_local_comp_cuda)
and then either marked failed (VisitFlag = 0.8) or written into the DS
volumes.  The indexing of a seed does not read the volumes, and a write only
raises Completeness and sets VisitFlag, so the test that admits a seed can
only turn from true to false as earlier seeds are committed.  The engine
therefore indexes a batch of admitted seeds in one grs_index_seeds call and
commits them in seed order, re-testing each seed against the volumes as they
are at its turn; a seed rejected at its turn is dropped exactly as the
sequential loop would skip it.  The result equals the sequential loop's for
every batch size.
"""
from __future__ import annotations

import math

import numpy as np


def pos2ind(pos, center, vox_size, rvp_first, dims, tomo_dim, simap: bool = False):
    """pos2ind.m:3-36: 1-based voxel indices of a position in mm, clamped to
    the volume (dims = size of the reconstructed box, tomo_dim =
    tomo_scale.Dimension, rvp_first = RecVolumePixel(:,1), center =
    tomo_scale.Center)."""
    p = np.asarray(pos, dtype=float)
    c = np.asarray(center, dtype=float)
    td = np.asarray(tomo_dim, dtype=float)
    if simap:
        v = np.array([(-p[0] - c[0]) / vox_size + td[0] / 2, (-p[1] - c[1]) / vox_size + td[1] / 2,
                      (p[2] - c[2]) / vox_size + td[2] / 2])
    else:
        v = (p - c) / vox_size + td / 2
    v = v - np.asarray(rvp_first, dtype=float) + 1
    out = [int(math.ceil(x)) for x in v]
    for k in range(3):
        out[k] = min(max(out[k], 1), int(dims[k]))
    return tuple(out)


class DS:
    """The DS volumes the indexing engine updates (x fastest, 1-based access
    through [x-1, y-1, z-1])."""

    def __init__(self, dims):
        nx, ny, nz = (int(v) for v in dims)
        self.Completeness = np.zeros((nx, ny, nz))
        self.GrainId = np.zeros((nx, ny, nz))
        self.Rodrigues = np.zeros((3, nx, ny, nz))
        self.IPF001 = np.zeros((3, nx, ny, nz))
        self.EulerZXZ = np.zeros((3, nx, ny, nz))
        self.Dismedian = np.zeros((nx, ny, nz))
        self.Ninter = np.zeros((nx, ny, nz))
        self.VisitFlag = np.zeros((nx, ny, nz))


def grain_indexing_engine(eng, pos_list, ds: DS, *, trust_comp: float, min_comp: float, max_dmedian: float,
                          first_grain_id: int, center, vox_size: float, rvp_first, tomo_dim, nr_seed: int,
                          simap: bool = False, batch: int = 32, **fit_opts):
    """GrainIndexing_engine_gpu_comp.m:40-124 against the orientation set
    already resident on `eng` (Engine.set_orientation_set).  Updates ds in
    place; returns (indexing_final (P, 18), pos_list_new (P, 3))."""
    pos_list = np.asarray(pos_list, dtype=float).reshape(-1, 3)
    P = pos_list.shape[0]
    out = np.zeros((P, 18))
    dims = ds.Completeness.shape
    inds = [pos2ind(pos_list[i], center, vox_size, rvp_first, dims, tomo_dim, simap) for i in range(P)]

    def admitted(i):
        x, y, z = inds[i]
        return ds.Completeness[x - 1, y - 1, z - 1] < trust_comp and ds.VisitFlag[x - 1, y - 1, z - 1] != 1

    i0 = 0
    while i0 < P:
        cand = []
        i = i0
        while i < P and len(cand) < batch:
            if admitted(i):
                cand.append(i)
            i += 1
        i0 = i
        if not cand:
            continue
        res = eng.index_seeds(pos_list[cand], nr_seed=nr_seed, **fit_opts)
        for k, s in enumerate(cand):
            if not admitted(s):
                continue
            row = res["indexing"][k].copy()
            row[0] = s + 1
            out[s] = row
            x, y, z = (v - 1 for v in inds[s])
            C, dmed = row[6], row[7]
            if C < min_comp or C < ds.Completeness[x, y, z] or dmed > max_dmedian:
                ds.VisitFlag[x, y, z] = 0.8
            else:
                eu = row[8:11]
                ds.GrainId[x, y, z] = s + 1 + first_grain_id - 1
                ds.Rodrigues[:, x, y, z] = row[11:14]
                ds.IPF001[:, x, y, z] = eu / math.sqrt((eu[0] * eu[0] + eu[1] * eu[1]) + eu[2] * eu[2])
                ds.EulerZXZ[:, x, y, z] = eu
                ds.VisitFlag[x, y, z] = 1.0
                ds.Completeness[x, y, z] = C
                ds.Dismedian[x, y, z] = dmed
                ds.Ninter[x, y, z] = row[5]
    return out, pos_list.copy()
