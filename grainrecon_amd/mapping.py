"""The seed -> index -> fit -> grow -> re-centre loop of
GrainMapping_engine_gpu_comp.m:54-331 on the device stages.

Every numeric step runs in the engine: seed indexing (grs_index_seeds,
fit_OR_constrain_with_local_comp_cuda; with check_fit_all the
fit_OR_constrain_with_local_comp_cuda_2 variant of
indexing.index_seed_fit_all for seeds more than 2 voxels from every voxel
indexed before the call), the spot pairing of the fitted
orientation (grs_pair_spots), the grain centre-of-mass fit
(grs_fit_grain_com), the re-scoring after a centre update (the counts of
grs_pair_spots), the region growth on the resident Completeness / Mask /
Dismedian / Ninter volumes (grs_grow_region).  This module keeps the
reference's per-seed control flow and the host-side DS volumes (GrainId,
Rodrigues, IPF001, EulerZXZ, VisitFlag).  When the spot half length is at
most 2 pixels the reference grows by connected voxels on the CPU
(grow_indexed_region.m); grow_connected keeps that loop on the host and
scores each step's new neighbours on the engine.
"""
from __future__ import annotations

import math

import numpy as np

from .engine import euler2u_deg, euler_to_rod_quat, fit_orientations
from .fit import SpotFit, levenberg_marquardt
from .geometry import quaternion2U
from .indexing import index_seed_fit_all, pos2ind

REFINE_NELDER_MEAD = "nelder-mead"
REFINE_ADJOINT = "adjoint"


# smoothing length (pixels) of the Charbonnier loss of the adjoint fits: the
# loss is the Euclidean distance the reference's objectives average
# (fit_grain_COM.m:84-86), quadratic within about this distance
LM_DELTA_PX = 0.5


def _lm_grain(eng, geo7, rows, rod, pos, fix_position):
    """Levenberg-Marquardt on the unrounded centroid residual of one grain's
    paired spots (grs_spot_normal: the hand-written adjoint Jacobian reduced
    on the device), geometry held, Charbonnier loss (LM_DELTA_PX) so that a
    spot paired with a neighbour grain's spot weighs like one distance, not
    its square: rows are grs_pair_spots rows (hkl 12:15, rotation angle 16,
    spot centroid 17:19).  Returns ([rod(3) pos(3)], FitResult)."""
    n = rows.shape[0]
    fit = SpotFit(eng, rows[:, 12:15], rows[:, 16], rows[:, 17:19], np.zeros(n, np.int32), 1)
    gr6 = np.concatenate([np.asarray(rod, float), np.asarray(pos, float)])[None, :]
    r = levenberg_marquardt(fit, geo7, gr6, fix_geometry=True, fix_position=fix_position, delta=LM_DELTA_PX)
    return r.gr6[0], r


def _rod_to_x(rod):
    """The vector part of the unit quaternion (q0 > 0) of a Rodrigues vector."""
    rod = np.asarray(rod, float)
    return rod / math.sqrt(1.0 + float(rod @ rod))


# at least this many paired spots (2 residuals each) before a 3- or
# 6-parameter least-squares fit is attempted
_LM_MIN_SPOTS = 4


def index_seed_adjoint(eng, geo7, pos, nr_seed, seed_number):
    """fit_OR_constrain_with_local_comp_cuda.m with its fminsearchbnd of
    1 - completeness over the quaternion's vector part (:115-126) replaced:
    the local sample of largest completeness (the reference's start, x0 of
    grs_index_seeds, whose Nelder-Mead is cut to its first simplex here) is
    paired with the spots (forward_comp_pair_spots.m), its Rodrigues vector is
    fitted to those spots' centroids by Levenberg-Marquardt with the position
    held (once more if the fitted orientation pairs other spots), and the
    orientation and verification counts follow as in :128-153.  Returns (indexing row (18,), candidate, fit info dict)."""
    res = eng.index_seeds(pos[None, :], nr_seed=nr_seed, max_fun_evals=1, max_iter=1)
    x0 = res["x0"][0]
    q0 = math.sqrt(max(0.0, 1.0 - float(x0 @ x0)))
    U0 = quaternion2U(np.concatenate([[q0], x0]))
    _, _, _, pr = eng.pair_spots(pos[None, :], U0[None, :, :])
    rows = pr[0]
    info = {"n_pairs": int(rows.shape[0]), "lm_iterations": 0, "rms_px": math.nan}
    x = x0
    if rows.shape[0] >= _LM_MIN_SPOTS and q0 > 0.0:
        rod = x0 / q0
        for _ in range(2):
            g, r = _lm_grain(eng, geo7, rows, rod, pos, fix_position=True)
            info.update(lm_iterations=info["lm_iterations"] + r.iterations, rms_px=r.rms)
            rod = g[:3]
            # the pairs of the fitted orientation; one more fit if they changed
            _, _, _, pr = eng.pair_spots(pos[None, :], quaternion2U(np.concatenate([[1.0], rod]) /
                                                                     math.sqrt(1.0 + float(rod @ rod)))[None])
            if pr[0].shape[0] == rows.shape[0] and np.array_equal(pr[0][:, 17:19], rows[:, 17:19]):
                break
            rows = pr[0]
            info["n_pairs"] = int(rows.shape[0])
            if rows.shape[0] < _LM_MIN_SPOTS:
                break
        x = _rod_to_x(rod)
    o10, U, _ = fit_orientations(x[None, :])
    ns, ni, med, _ = eng.pair_spots(pos[None, :], U)
    row = np.zeros(18)
    row[0] = seed_number
    row[1:4] = pos
    row[4], row[5] = ns[0], ni[0]
    row[6] = ni[0] / ns[0] if ns[0] else math.nan
    row[7] = float(med[0])
    row[8:18] = o10[0]
    return row, float(np.asarray(res["candidate"]).reshape(-1)[0]), info


def fit_grain_com_adjoint(eng, geo7, pairs, rod):
    """fit_grain_COM.m:5-29 with its fminsearchbnd of the mean rounded pixel
    distance over [position, Euler angles] replaced by Levenberg-Marquardt on
    the unrounded centroid residual over [Rodrigues vector, position], from
    the same start (the pair rows' position, :17, and the indexed
    orientation).  Returns FitOutput (6,) = [position, Euler ZXZ degrees] and
    the LM result (None below the spot minimum, where the start is
    returned)."""
    pos0 = pairs[0, 9:12]
    if pairs.shape[0] < _LM_MIN_SPOTS:
        o10, _, _ = fit_orientations(_rod_to_x(rod)[None, :])
        return np.concatenate([pos0, o10[0, :3]]), None
    g, r = _lm_grain(eng, geo7, pairs, rod, pos0, fix_position=False)
    o10, _, _ = fit_orientations(_rod_to_x(g[:3])[None, :])
    return np.concatenate([g[3:], o10[0, :3]]), r


class MappingDS:
    """DS of GrainMapping_engine_gpu_comp.m:24-34 (volumes indexed
    [x-1, y-1, z-1]); Completeness, Mask, Dismedian and Ninter are mirrored on
    the engine for the growth.  The arrays are column-major like MATLAB's,
    which is also the engine's layout (x fastest), so a volume goes to the
    device without a transposing copy."""

    def __init__(self, mask):
        mask = np.asarray(mask, dtype=float)
        nx, ny, nz = mask.shape
        self.Mask = np.asfortranarray(mask).copy(order="F")
        self.Completeness = np.zeros((nx, ny, nz), order="F")
        self.GrainId = np.zeros((nx, ny, nz), order="F")
        self.Rodrigues = np.zeros((3, nx, ny, nz), order="F")
        self.IPF001 = np.zeros((3, nx, ny, nz), order="F")
        self.EulerZXZ = np.zeros((3, nx, ny, nz), order="F")
        self.Dismedian = np.zeros((nx, ny, nz), order="F")
        self.Ninter = np.zeros((nx, ny, nz), order="F")
        self.VisitFlag = np.zeros((nx, ny, nz), order="F")


_NEIGHBOURS = ((1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 1, 1), (0, 0, 1),
               (0, -1, 1), (0, 1, -1), (0, 0, -1), (0, -1, -1), (-1, 1, 0), (-1, -1, 0), (-1, 0, 1), (-1, 0, -1),
               (-1, 0, 0))


def _mround(v: float) -> float:
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


def grow_connected(eng, ds, seed, indexed_comp, U, drop_off, maxD, max_dmedian, p1):
    """grow_indexed_region.m:1-142, the growth GrainMapping_engine_gpu_comp.m:
    207-214 uses when the spot half length is at most 2 pixels: from the
    seed (1-based voxel), the 18 neighbours of the voxel last added that are
    in the mask and not yet listed are scored with forward_comp on the engine
    (the scoring of cuda_grow_indexed_region, one launch per step for the
    neighbours not scored before: a voxel's score depends only on the voxel,
    U and the stack) and queued when C > I and dis_median <= min(DisMedian,
    maxDmedian), or |C - I| <= 0.01 and Nr_intersect > Ninter; the queued
    voxel of largest C / seed_comp (first maximum) joins the region while that
    ratio exceeds 1 - drop_off and the region is smaller than the mask count
    minus 1, and leaves the queue by the last entry moving into its slot.
    The grown voxels' Completeness, Dismedian and Ninter are written into ds
    (the caller's copy-back, :234-236).  p1: the 11 values [RecVolumePixel(:,1)'
    simap VoxSize Center' Dimension].  Returns (center, replace_center,
    indexed (n, 3) 1-based in find() order)."""
    I = ds.Completeness.copy()
    mask, DM, NI = ds.Mask, ds.Dismedian, ds.Ninter
    x, y, z = (int(v) for v in seed)
    I[x - 1, y - 1, z - 1] = indexed_comp
    seed_comp = I[x - 1, y - 1, z - 1]
    J = np.zeros(I.shape, np.int8)
    upd = {}
    param1 = np.concatenate([np.asarray(p1, float)[:11], [seed_comp, drop_off, max_dmedian]])
    n_mask = int((mask > 0).sum())
    csum = seed_comp
    cpos = [seed_comp * x, seed_comp * y, seed_comp * z]
    c0 = [v / csum for v in cpos]
    queue = []
    scores = {}
    reg_size = 0
    nx, ny, nz = I.shape
    while True:
        cand = []
        for dx, dy, dz in _NEIGHBOURS:
            xn, yn, zn = x + dx, y + dy, z + dz
            if 1 <= xn <= nx and 1 <= yn <= ny and 1 <= zn <= nz and J[xn - 1, yn - 1, zn - 1] == 0 and \
                    mask[xn - 1, yn - 1, zn - 1] > 0:
                cand.append((xn, yn, zn))
        new = [c for c in cand if c not in scores]
        if new:
            rows = np.array([[c[0], c[1], c[2], 0.0, 1.0, 0.0] for c in new])
            out = eng.grow_indexed_region(rows, U, param1)
            for c, o in zip(new, out):
                scores[c] = (o[3], o[4], o[5])
        for c in cand:
            ns, ni, med = scores[c]
            C = ni / ns if ns > 0 else math.nan
            Iv, DMv, NIv = I[c[0] - 1, c[1] - 1, c[2] - 1], DM[c[0] - 1, c[1] - 1, c[2] - 1], NI[c[0] - 1, c[1] - 1,
                                                                                                   c[2] - 1]
            lim = max_dmedian if math.isnan(DMv) else (DMv if math.isnan(max_dmedian) else min(DMv, max_dmedian))
            if (C > Iv and med <= lim) or (abs(C - Iv) <= 0.01 and ni > NIv):
                queue.append((c[0], c[1], c[2], C, med, ni))
                J[c[0] - 1, c[1] - 1, c[2] - 1] = 1
        J[x - 1, y - 1, z - 1] = 2
        if not queue:
            break
        ratios = [q[3] / seed_comp for q in queue]
        best = max(ratios)
        k = ratios.index(best)
        reg_size += 1
        if not (best > 1 - drop_off and reg_size < n_mask - 1):
            break
        qx, qy, qz, C, med, ni = queue[k]
        csum = csum + C
        cpos = [cpos[0] + C * qx, cpos[1] + C * qy, cpos[2] + C * qz]
        upd[(qx, qy, qz)] = (C, med, ni)
        x, y, z = qx, qy, qz
        queue[k] = queue[-1]
        queue.pop()
    lin = np.flatnonzero((J > 1).reshape(-1, order="F"))
    idx = np.stack(np.unravel_index(lin, J.shape, order="F"), 1) + 1
    for xi, yi, zi in idx:
        if (xi, yi, zi) in upd:
            C, med, ni = upd[(xi, yi, zi)]
            ds.Completeness[xi - 1, yi - 1, zi - 1] = C
            DM[xi - 1, yi - 1, zi - 1] = med
            NI[xi - 1, yi - 1, zi - 1] = ni
        else:
            ds.Completeness[xi - 1, yi - 1, zi - 1] = I[xi - 1, yi - 1, zi - 1]
    c1 = [v / csum for v in cpos]
    e0, e1, e2 = c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]
    if math.sqrt((e0 * e0 + e1 * e1) + e2 * e2) > max(maxD, maxD / 10000 * idx.shape[0]):
        return np.array([_mround(v) for v in c1]), 1, idx
    return np.array(c0), 0, idx


def _upload(eng, ds):
    for k in ("Completeness", "Mask", "Dismedian", "Ninter"):
        eng.set_volume(k, getattr(ds, k))


def _upload_at(eng, ds, vox):
    """grow_connected changes Completeness, Dismedian and Ninter of ds only
    at the voxels it returns: write those to the device copies."""
    if vox.shape[0] == 0:
        return
    x, y, z = vox[:, 0] - 1, vox[:, 1] - 1, vox[:, 2] - 1
    for k in ("Completeness", "Dismedian", "Ninter"):
        eng.set_volume_at(k, vox, getattr(ds, k)[x, y, z])


def _download(eng, ds, vox):
    """The device growth changes Completeness, Dismedian and Ninter only at
    the voxels it returns (the seed and the accepted ones): copy those."""
    if vox.shape[0] == 0:
        return
    x, y, z = vox[:, 0] - 1, vox[:, 1] - 1, vox[:, 2] - 1
    for k in ("Completeness", "Dismedian", "Ninter"):
        getattr(ds, k)[x, y, z] = eng.get_volume_at(k, vox)


def grain_mapping_engine(eng, pos_list, OR_q, ds: MappingDS, *, trust_comp, min_comp, drop_off, maxD, max_dmedian,
                         nr_seed, first_grain_id, center, vox_size, rvp_first, tomo_dim, simap=False,
                         check_fit_all=False, refine=REFINE_NELDER_MEAD, geo7=None, fit_log=None):
    """GrainMapping_engine_gpu_comp.m:54-331.  OR_q: the orientation set
    (N, 4); ds is updated in place.  With check_fit_all, the voxels with
    Completeness >= trust_comp at entry (else >= min(1.5 min_comp,
    0.8 trust_comp)) are listed (:46-53), and a first indexing of a seed more
    than 2 voxels from all of them uses the fit_all variant (:119-123;
    MATLAB rejects an empty list there, and so does this function).
    After every seed its SeedIndexing row is recorded (:293-299:
    indexing_final(i,1:11), Nr_indexed_voxel, the growth of the GrainId > 0
    count during the seed, the candidate, the fit_all flag).  Past seed 200
    (:318-330): without check_fit_all, when fewer than 5% of the indexed
    seeds among the last 181 added GrainId voxels, check_fit_all turns on
    with the voxels of GrainId > 0 at entry as the list; with check_fit_all,
    when fewer than 5% of the indexed seeds among the last 51 did, the loop
    ends (the remaining rows stay zero).  Returns (indexing_final (P, 18),
    pos_list_new (P, 3), Nr_indexed_voxel (P,), SeedIndexing (P, 15)).

    refine selects the fits at the fit_OR_constrain_with_local_comp_cuda
    (:87-91, :128-130) and fit_grain_COM (:140-170) call sites:
    REFINE_NELDER_MEAD runs the reference's fminsearchbnd on the engine;
    REFINE_ADJOINT (opt-in, needs geo7 = [Lsam2sou Lsam2det dety00 detz00
    tilt_x tilt_y tilt_z] of the engine's geometry) runs index_seed_adjoint
    and fit_grain_com_adjoint instead: Levenberg-Marquardt on the paired
    spots' centroid residual with the hand-written adjoint Jacobian.  The
    fit_all variant keeps its Nelder-Mead in both modes.  fit_log, a list,
    receives one dict per adjoint fit (site, seed, spots, iterations, rms)."""
    if refine not in (REFINE_NELDER_MEAD, REFINE_ADJOINT):
        raise ValueError(f"refine must be {REFINE_NELDER_MEAD!r} or {REFINE_ADJOINT!r}")
    adjoint = refine == REFINE_ADJOINT
    if adjoint:
        if geo7 is None:
            raise ValueError("refine='adjoint' needs geo7, the geometry's 7 fit values")
        geo7 = np.asarray(geo7, float).reshape(7)
    pos_list = np.asarray(pos_list, dtype=float).reshape(-1, 3)
    OR_q = np.asarray(OR_q, dtype=float).reshape(-1, 4)
    P = pos_list.shape[0]
    dims = ds.Completeness.shape
    ind_final = np.zeros((P, 18))
    pos_new = np.zeros((P, 3))
    nr_vox = np.zeros(P, dtype=np.int64)
    rcc = np.zeros(P, dtype=np.int64)
    seed_ix = np.zeros((P, 15))
    cand_id = np.zeros(P)
    fa_track = np.zeros(P)
    gid_entry = ds.GrainId.reshape(-1, order="F") > 0
    # GrainId > 0 count, kept up to date at each assignment (a recount of the
    # whole volume per seed costs milliseconds)
    n_gid = n_gid_prev = int(np.count_nonzero(gid_entry))
    p1 = np.array([rvp_first[0], rvp_first[1], rvp_first[2], 1.0 if simap else 0.0, vox_size, center[0], center[1],
                   center[2], tomo_dim[0], tomo_dim[1], tomo_dim[2]], dtype=float)
    eng.set_orientation_set(OR_q)
    _upload(eng, ds)
    temp = None
    ind_gid = None
    if check_fit_all:
        flat = ds.Completeness.reshape(-1, order="F")
        lin = np.flatnonzero(flat >= trust_comp)
        if lin.size == 0:
            lin = np.flatnonzero(flat >= min(min_comp * 1.5, trust_comp * 0.8))
        ind_gid = np.stack(np.unravel_index(lin, dims, order="F"), 1) + 1

    def p2i(p):
        return pos2ind(p, center, vox_size, rvp_first, dims, tomo_dim, simap)

    for i in range(P):
        gid = i + first_grain_id
        stop = False
        pos_ix = pos_list[i].copy()
        while not stop and rcc[i] <= 3:
            x, y, z = (v - 1 for v in p2i(pos_ix))
            c_here = ds.Completeness[x, y, z]
            if not ((c_here < trust_comp or (rcc[i] > 0 and c_here >= trust_comp)) and ds.VisitFlag[x, y, z] != 1):
                stop = True
                pos_new[i] = pos_ix
                break
            fit_all = False
            if rcc[i] == 0 and check_fit_all:
                if ind_gid.shape[0] == 0:
                    raise ValueError("check_fit_all needs indexed voxels (an empty list fails in MATLAB)")
                d = ind_gid - np.array([x + 1, y + 1, z + 1])
                fit_all = float(np.sqrt((d * d).sum(1)).min()) > 2
            def index_here():
                if adjoint:
                    row_a, cand_a, inf = index_seed_adjoint(eng, geo7, pos_ix, nr_seed, i + 1)
                    if fit_log is not None:
                        fit_log.append(dict(site="fit_OR_constrain", seed=i + 1, **inf))
                    return row_a, cand_a
                r = eng.index_seeds(pos_ix[None, :], nr_seed=nr_seed)
                return r["indexing"][0].copy(), float(np.asarray(r["candidate"]).reshape(-1)[0])

            if fit_all:
                # :123-127: fit_OR_constrain_with_local_comp_cuda_2 with fit_all_flag = 1
                fa = index_seed_fit_all(eng, pos_ix, OR_q, nr_seed)
                row, cand = fa["indexing"].copy(), float(np.asarray(fa["candidate"]).reshape(-1)[0])
            elif rcc[i] == 0:
                row, cand = index_here()
            else:
                # :87-91: the indexing repeated on ori_local_sampling(q, 4) of the last fit
                local = eng.local_sampling(temp[14:18], 4, 0.004)[0]
                eng.set_orientation_set(local)
                try:
                    row, cand = index_here()
                finally:
                    eng.set_orientation_set(OR_q)
            row[0] = i + 1
            if rcc[i] == 0:
                cand_id[i] = cand
                fa_track[i] = 1.0 if fit_all else 0.0
            U = quaternion2U(row[14:18])
            _, _, _, pairs = eng.pair_spots(pos_ix[None, :], U[None, :, :])
            pairs = pairs[0]
            ind_final[i] = row
            if rcc[i] == 0 and pairs.shape[0]:
                # :140-170: grain centre of mass from the paired spots
                if adjoint:
                    fo, lm = fit_grain_com_adjoint(eng, geo7, pairs, row[11:14])
                    if fit_log is not None:
                        fit_log.append(dict(site="fit_grain_COM", seed=i + 1, n_pairs=int(pairs.shape[0]),
                                            lm_iterations=lm.iterations if lm else 0,
                                            rms_px=lm.rms if lm else math.nan))
                else:
                    fit_out, _, _, _ = eng.fit_grain_com([pairs])
                    fo = fit_out[0]
                d = fo[:3] - ind_final[i, 1:4]
                if math.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) / vox_size > maxD:
                    tx, ty, tz = (v - 1 for v in p2i(fo[:3]))
                    if ds.Completeness[tx, ty, tz] < ind_final[i, 6] and ds.Mask[tx, ty, tz] == 1:
                        ind_final[i, 1:4] = fo[:3]
                        ind_final[i, 8:11] = fo[3:6]
                        rod, quat, _ = euler_to_rod_quat(fo[3:6][None, :])
                        ind_final[i, 11:14] = rod[0]
                        ind_final[i, 14:18] = quat[0]
                        UU = euler2u_deg(ind_final[i, 8:11][None, :])
                        ns, ni, med, _ = eng.pair_spots(ind_final[i, 1:4][None, :], UU)
                        ind_final[i, 4] = ns[0]
                        ind_final[i, 5] = ni[0]
                        ind_final[i, 7] = float(med[0])
                        ind_final[i, 6] = ind_final[i, 5] / ind_final[i, 4] if ns[0] else math.nan
                        pos_ix = fo[:3].copy()
                        x, y, z = (v - 1 for v in p2i(pos_ix))
            Ci, dmed = ind_final[i, 6], ind_final[i, 7]
            if Ci < min_comp or Ci < ds.Completeness[x, y, z] or dmed > max_dmedian:
                stop = True
                pos_new[i] = pos_ix
                ds.VisitFlag[x, y, z] = 0.8
                break
            temp = ind_final[i].copy()
            centre = np.array([x + 1, y + 1, z + 1], dtype=float)
            replace = 0
            if temp[6] >= ds.Completeness[x, y, z]:
                ds.Completeness[x, y, z] = temp[6]
                UU = quaternion2U(temp[14:18])
                if pairs.shape[0]:
                    a = pairs[:, 19]
                    n = a.size
                    mu = sum(float(v) for v in a) / n if n else math.nan
                    sd = math.sqrt(sum((float(v) - mu) ** 2 for v in a) / (n - 1)) if n > 1 else 0.0
                    spot_l = math.sqrt((mu + 2 * sd) / math.pi)
                else:
                    spot_l = 20.0
                if spot_l <= 2:
                    # :207-214: connected-voxel growth on the host volumes
                    gc, grep_, gidx = grow_connected(eng, ds, (x + 1, y + 1, z + 1), temp[6], UU, drop_off, maxD,
                                                     max_dmedian, p1)
                    _upload_at(eng, ds, gidx)
                    g = {"center": gc, "replace_center": grep_, "indexed": gidx}
                else:
                    g = eng.grow_region((x + 1, y + 1, z + 1), temp[6], UU, spot_l, drop_off, maxD, max_dmedian, p1)
                    _download(eng, ds, g["indexed"])
                idx = g["indexed"] - 1
                nr_vox[i] = idx.shape[0]
                xi, yi, zi = idx[:, 0], idx[:, 1], idx[:, 2]
                lin_u = np.unique(np.ravel_multi_index((xi, yi, zi), dims))
                was = np.count_nonzero(ds.GrainId[np.unravel_index(lin_u, dims)] > 0)
                ds.GrainId[xi, yi, zi] = gid
                n_gid += (lin_u.size if gid > 0 else 0) - int(was)
                eu = temp[8:11]
                ds.Rodrigues[:, xi, yi, zi] = temp[11:14][:, None]
                ds.IPF001[:, xi, yi, zi] = (eu / math.sqrt((eu[0] * eu[0] + eu[1] * eu[1]) + eu[2] * eu[2]))[:, None]
                ds.EulerZXZ[:, xi, yi, zi] = eu[:, None]
                ds.VisitFlag[xi, yi, zi] = 0.5
                ds.VisitFlag[x, y, z] = 1.0
                centre = g["center"].copy()
                replace = g["replace_center"]
                if replace == 1:
                    cx, cy, cz = (int(v) - 1 for v in centre)
                    if ds.GrainId[cx, cy, cz] != gid:
                        # :238-253: the grain's voxel nearest to the new centre
                        lin = np.flatnonzero(ds.GrainId.reshape(-1, order="F") == gid)
                        allid = np.stack(np.unravel_index(lin, dims, order="F"), 1) + 1
                        dd = np.sqrt(((allid - centre) ** 2).sum(1))
                        centre = allid[int(np.argmin(dd))].astype(float)
                        dc = centre - np.array([x + 1, y + 1, z + 1])
                        replace = 1 if math.sqrt((dc[0] * dc[0] + dc[1] * dc[1]) + dc[2] * dc[2]) > maxD else 0
            if replace == 0:
                stop = True
                pos_new[i] = pos_ix
            else:
                rcc[i] += 1
                pn = ((centre + np.asarray(rvp_first, float) - 1) - np.asarray(tomo_dim, float) / 2) * vox_size + \
                    np.asarray(center, float)
                if simap:
                    pn[0], pn[1] = -pn[0], -pn[1]
                pos_new[i] = pn
                pos_ix = pn.copy()
        # :293-299: the seed's record; column 13 = voxels that gained a grain id
        seed_ix[i] = np.concatenate([ind_final[i, :11], [nr_vox[i], n_gid - n_gid_prev, cand_id[i], fa_track[i]]])
        n_gid_prev = n_gid
        k = i + 1  # MATLAB's seed number
        if not check_fit_all and k > 200:
            w = seed_ix[k - 181:k]  # SeedIndexing(i-180:i, :)
            n_ix = int(np.count_nonzero(w[:, 0] > 0))
            if n_ix and np.count_nonzero(w[:, 12] > 0) / n_ix < 0.05:
                check_fit_all = True
                lin = np.flatnonzero(gid_entry)
                ind_gid = np.stack(np.unravel_index(lin, dims, order="F"), 1) + 1
        if check_fit_all and k > 200:
            w = seed_ix[k - 51:k]  # SeedIndexing(i-50:i, :)
            n_ix = int(np.count_nonzero(w[:, 0] > 0))
            if n_ix and np.count_nonzero(w[:, 12] > 0) / n_ix < 0.05:
                break
    return ind_final, pos_new, nr_vox, seed_ix
