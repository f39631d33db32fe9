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


def _sort_desc(C):
    """Row order of sortrows(X, col, 'descend'): stable, NaN first."""
    C = np.asarray(C, dtype=float)
    idx = np.arange(C.size)
    nan = np.isnan(C)
    rest = idx[~nan]
    return np.concatenate([idx[nan], rest[np.argsort(-C[rest], kind="stable")]])


def _first_max(v):
    """[~, i] = max(v): NaN ignored, first maximum (0-based; all NaN -> 0)."""
    v = np.asarray(v, dtype=float)
    ok = ~np.isnan(v)
    if not ok.any():
        return 0
    return int(np.flatnonzero(ok & (v == v[ok].max()))[0])


def _grid_rows(eng, q_rows, groups, nr, gridsize):
    """q_all of forward_calc_gridding_OR.m:7-15: the local sampling of row j
    at rows (j-1)*len_j+1 .. j*len_j with group id groups[j] (MATLAB grows
    the array on assignment, later rows overwrite earlier ones, unwritten
    rows stay zero with group 0)."""
    locals_ = eng.local_sampling(np.asarray(q_rows, dtype=float).reshape(-1, 4), nr, gridsize)
    lens = [l.shape[0] for l in locals_]
    K = len(lens)
    R = max([K * lens[0]] + [(j + 1) * lens[j] for j in range(K)])
    q_all = np.zeros((R, 4))
    gid = np.zeros(R)
    for j in range(K):
        q_all[j * lens[j]:(j + 1) * lens[j]] = locals_[j]
        gid[j * lens[j]:(j + 1) * lens[j]] = groups[j]
    return q_all, gid, lens


def _score9(eng, pos, q_all, gid):
    ns, ni, C, med = eng.forward_comp(pos, q_all)
    return np.column_stack([ns, ni, C, med.astype(float), q_all, gid])


def index_seed_fit_all(eng, pos, q_set, nr_seed: int, *, n_fit: int = 5, nr: int = 8, gridsize: float = 0.004,
                       regrid_nr: int = 5, regrid_gridsize: float = 0.001, iterations: int = 2,
                       half_width: float = 0.02, max_fun_evals: int = 2000, max_iter: int = 2000,
                       tol_fun: float = 1e-10, tol_x: float = 1e-4) -> dict:
    """fit_OR_constrain_with_local_comp_cuda_2.m:8-201 with fit_all_flag = 1
    (Nr_fit = 5), the seed-indexing branch GrainMapping_engine_gpu_comp.m:
    119-127 takes for seeds away from indexed voxels.  Every scoring, local
    sampling, fit and verification runs on the engine; the host keeps the
    reference's row bookkeeping:
      index_seeds = the first min(Nr_seed, N) rows of sortrows(Output,3,
      'descend') over the orientation set q_set (engine :100-109);
      local sampling (nr, gridsize) of every seed with group id j (:22-30),
      scored; per group the first maximum of C over the blocks of the last
      seed's sampling length (:73-84); sortrows(.,3,'descend');
      `iterations` times forward_calc_gridding_OR(rows, regrid_nr,
      regrid_gridsize) (:86-90): sampling around every row with its group id,
      scored, stably sorted by group id, per-group maxima over blocks of the
      last row's sampling length, sorted by C descending;
      fminsearchbnd from the first n_fit rows' (q1, q2, q3) (:151-170), the
      orientation bookkeeping, the verification counts and spot pairs
      (forward_comp_pair_spots.m, so the frame spots must be set); the fit of
      highest C (:193-197).
    The reference asks forward_calc_gridding_OR for a third output
    (nr_per_OR1) that the function does not declare; its own block size (the
    last row's sampling length) is used, as the recomputation at :136-147
    implies.  Returns a dict: indexing (18,), candidate (group id of the
    chosen fit), pairs (HittedSpots rows of the chosen fit), fits (per fit:
    x0, x, fval, n_fun_evals, n_iter, group, C), output_max (K, 9)."""
    pos = np.asarray(pos, dtype=float).reshape(3)
    q_set = np.asarray(q_set, dtype=float).reshape(-1, 4)
    ns, ni, C, med = eng.forward_comp(pos, q_set)
    K = min(int(nr_seed), q_set.shape[0])
    if n_fit > K:
        raise ValueError(f"{n_fit} fits need at least {n_fit} seeds (got {K})")
    seeds = _sort_desc(C)[:K]
    q_all, gid, lens = _grid_rows(eng, q_set[seeds], np.arange(1, K + 1, dtype=float), nr, gridsize)
    out9 = _score9(eng, pos, q_all, gid)
    nper = lens[-1]
    omax = np.zeros((K, 9))
    for j in range(1, K + 1):
        if (gid == j).any():
            lo = nper * (j - 1)
            omax[j - 1] = out9[lo + _first_max(out9[lo:nper * j, 2])]
    omax = omax[_sort_desc(omax[:, 2])]
    for _ in range(int(iterations)):
        q2, gid2, lens2 = _grid_rows(eng, omax[:, 4:8], omax[:, 8], regrid_nr, regrid_gridsize)
        srt = _score9(eng, pos, q2, gid2)
        srt = srt[np.argsort(srt[:, 8], kind="stable")]
        n1 = lens2[-1]
        om = np.zeros((K, 9))
        for j in range(1, K + 1):
            if (srt[:, 8] == j).any():
                lo = n1 * (j - 1)
                om[j - 1] = srt[lo + _first_max(srt[lo:n1 * j, 2])]
        omax = om[_sort_desc(om[:, 2])]
    x0 = omax[:int(n_fit), 5:8].copy()
    x, fval, nf, it = eng.fit_or_local(np.repeat(pos[None, :], x0.shape[0], 0), x0, half_width, max_fun_evals,
                                       max_iter, tol_fun, tol_x)
    from .engine import fit_orientations

    o10, U, ok = fit_orientations(x)
    vns, vni, vmed, pairs = eng.pair_spots(np.repeat(pos[None, :], x0.shape[0], 0), U)
    Cfit = np.where((ok != 0) & (vns > 0), vni / np.maximum(vns, 1), np.nan)
    fits = [dict(x0=x0[j], x=x[j], fval=fval[j], n_fun_evals=int(nf[j]), n_iter=int(it[j]), group=int(omax[j, 8]),
                 C=float(Cfit[j])) for j in range(x0.shape[0])]
    best = _first_max(Cfit)
    if ok[best]:
        row = np.concatenate([[0.0], pos, [vns[best], vni[best], Cfit[best], float(vmed[best])], o10[best]])
    else:
        row = np.concatenate([[0.0], pos, np.full(14, np.nan)])
    return dict(indexing=row, candidate=int(omax[best, 8]), pairs=pairs[best], fits=fits, output_max=omax)
