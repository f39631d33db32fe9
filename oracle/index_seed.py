"""Seed indexing of one position -- TEST INFRASTRUCTURE ONLY (checker of the
engine's grs_index_seeds).

GrainIndexing_engine_gpu_comp.m:46-66 with
fit_OR_constrain_with_local_comp_cuda.m:12-165:
  Output = ForwardCalc_OR_space_comp(pos, OR)            (C oracle)
  index_seeds = first min(Nr_seed, N) rows of sortrows(Output,3,'descend')
  q_all: ori_local_sampling(index_seeds(j,5:8), nr, gridsize) of seed j at
         rows (j-1)*len_j+1 .. j*len_j (:19-26; MATLAB grows the array on
         assignment, later seeds overwrite earlier ones, unwritten rows stay
         zero with column 5 = 0)
  Output = scores of q_all (:47-56)
  Output_max(j,:) = Output row of the first maximum of C over rows
         1+nr_per_OR*(j-1) .. nr_per_OR*j when column 9 contains j (:100-111;
         nr_per_OR is the LAST seed's row count), zeros otherwise
  sortrows(Output_max,3,'descend'); fminsearchbnd from rows 1..Nr_fit
         (:115-126, oracle.fit_or); the orientation bookkeeping (:133-148,
         oracle.orient) and the verification counts with that U
         (forward_comp_pair_spots.m counts and median, oracle.forward_comp);
  the fit of highest C (:161, max: NaN ignored, first maximum).
sortrows(...,'descend') is stable and puts NaN first (MATLAB documentation).
"""
from __future__ import annotations

import math

import numpy as np

from . import forward_comp, local_sampling, or_space_comp
from .fit_or import fit_or_local
from .orient import fit_orientation


def sort_desc(C) -> np.ndarray:
    """Row order of sortrows(X, col, 'descend') for the column values C."""
    C = np.asarray(C, dtype=float)
    idx = np.arange(C.size)
    nan = np.isnan(C)
    rest = idx[~nan]
    return np.concatenate([idx[nan], rest[np.argsort(-C[rest], kind="stable")]])


def matlab_max_index(v) -> int:
    """[~, i] = max(v): NaN ignored, first maximum; all NaN -> 1 (0-based 0)."""
    v = np.asarray(v, dtype=float)
    ok = ~np.isnan(v)
    if not ok.any():
        return 0
    m = v[ok].max()
    return int(np.flatnonzero(ok & (v == m))[0])


def qall_owner(lens) -> np.ndarray:
    """Owner (1-based seed number, 0 = unwritten) of every q_all row."""
    K = len(lens)
    R = max([K * lens[0]] + [(j + 1) * lens[j] for j in range(K)])
    owner = np.zeros(R, dtype=np.int64)
    for j in range(K):
        owner[j * lens[j]:(j + 1) * lens[j]] = j + 1
    return owner


def index_seed(geo, stack, pos, q_set, rot, hkl, nr_seed, nr=8, gridsize=0.004, n_fit=1, half_width=0.02,
               max_fun_evals=2000, max_iter=2000, tol_fun=1e-10, tol_x=1e-4, nthreads=8) -> dict:
    q_set = np.asarray(q_set, dtype=float).reshape(-1, 4)
    out1 = or_space_comp(geo, stack, pos, q_set, rot, hkl, nthreads=nthreads)
    order = sort_desc(out1[:, 2])
    K = min(int(nr_seed), q_set.shape[0])
    seeds = order[:K]
    locals_ = [local_sampling(q_set[u], nr, gridsize) for u in seeds]
    lens = [l.shape[0] for l in locals_]
    owner = qall_owner(lens)
    R = owner.size
    q_all = np.zeros((R, 4))
    for r in range(R):
        j = owner[r]
        if j:
            q_all[r] = locals_[j - 1][r - (j - 1) * lens[j - 1]]
    out2 = or_space_comp(geo, stack, pos, q_all, rot, hkl, nthreads=nthreads)
    nper = lens[-1]
    omax_C = np.zeros(K)
    omax_row = np.full(K, -1)
    for j in range(1, K + 1):
        if (owner == j).any():
            lo = nper * (j - 1)
            r = lo + matlab_max_index(out2[lo:nper * j, 2])
            omax_row[j - 1] = r
            omax_C[j - 1] = out2[r, 2]
    fits = []
    for jj in sort_desc(omax_C)[:min(int(n_fit), K)]:
        r = omax_row[jj]
        x0 = q_all[r, 1:4].copy() if r >= 0 else np.zeros(3)
        group = int(owner[r]) if r >= 0 else 0
        x, fval, nf, it = fit_or_local(geo, stack, pos, x0, rot, hkl, half_width, max_fun_evals, max_iter, tol_fun,
                                       tol_x)
        conv = fit_orientation(x)
        if conv is None:
            ns = ni = med = C = math.nan
            o10 = np.full(10, math.nan)
        else:
            o10, U = conv
            ns, ni, med = forward_comp(geo, stack, U, pos, rot, hkl)
            C = ni / ns if ns > 0 else math.nan
        fits.append(dict(x0=x0, x=np.asarray(x), fval=fval, nf=nf, it=it, group=group, ns=ns, ni=ni, C=C, med=med,
                         o10=o10))
    best = matlab_max_index([f["C"] for f in fits])
    f = fits[best]
    row = np.concatenate([[0.0], np.asarray(pos, float), [f["ns"], f["ni"], f["C"], f["med"]], f["o10"]])
    seed_C = out1[seeds, 2]
    return dict(indexing=row, candidate=f["group"], fit=f, fits=fits, seed_idx=seeds, seed_C=seed_C, lens=lens,
                owner=owner)


def _grid_rows(q_rows, groups, nr, gridsize):
    """q_all of forward_calc_gridding_OR.m:7-15 (and of the first stage of
    fit_OR_constrain_with_local_comp_cuda_2.m:22-30): the local sampling of
    row j at rows (j-1)*len_j+1 .. j*len_j with group id groups[j] (MATLAB
    grows the array on assignment; unwritten rows stay zero, group 0)."""
    locals_ = [local_sampling(q, nr, gridsize) for q in q_rows]
    lens = [l.shape[0] for l in locals_]
    owner = qall_owner(lens)
    q_all = np.zeros((owner.size, 4))
    gid = np.zeros(owner.size)
    for r in range(owner.size):
        j = owner[r]
        if j:
            q_all[r] = locals_[j - 1][r - (j - 1) * lens[j - 1]]
            gid[r] = groups[j - 1]
    return q_all, gid, lens


def _group_max(out9, K, nper):
    """Output_max of forward_calc_gridding_OR.m:26-37: Output sorted by the
    group column (stable), then per group id j = 1..K present in it the row of
    the first maximum of C over sorted rows nper*(j-1)+1 .. nper*j."""
    order = np.argsort(out9[:, 8], kind="stable")
    srt = out9[order]
    omax = np.zeros((K, 9))
    for j in range(1, K + 1):
        if (srt[:, 8] == j).any():
            lo = nper * (j - 1)
            blk = srt[lo:nper * j, 2]
            if blk.size == 0:
                raise IndexError("group block beyond the scored rows (MATLAB index error)")
            omax[j - 1] = srt[lo + matlab_max_index(blk)]
    return srt, omax


def index_seed_fit_all(geo, stack, pos, q_set, rot, hkl, nr_seed, nr=8, gridsize=0.004, regrid_nr=5,
                       regrid_gridsize=0.001, iterations=2, n_fit=5, half_width=0.02, max_fun_evals=2000,
                       max_iter=2000, tol_fun=1e-10, tol_x=1e-4, nthreads=8) -> dict:
    """fit_OR_constrain_with_local_comp_cuda_2.m:8-201 with fit_all_flag = 1
    (Nr_fit = 5), reached from GrainMapping_engine_gpu_comp.m:119-127:
      index_seeds = first min(Nr_seed, N) rows of sortrows(Output,3,'descend')
      of the whole orientation set (:100-109 of the engine);
      stage 1: local sampling (8, 0.004) of every seed, group id j (:22-30),
      scored; per-group maxima over blocks of the first seed's... row count of
      the LAST seed (:73-84); sortrows(.,3,'descend');
      twice forward_calc_gridding_OR(rows, 5, 0.001) (:86-90): local sampling
      around every row with its group id, scored, sorted by group id, the
      per-group maxima (block size = the last row's sampling length), sorted
      by C descending;
      fminsearchbnd from the first Nr_fit rows' (q1, q2, q3) (:151-170), the
      orientation bookkeeping, the verification counts; the fit of highest C
      (:193-197, first maximum).
    The reference calls forward_calc_gridding_OR with three outputs
    ([Output, Output_max_temp, nr_per_OR1]) while the function declares two,
    which MATLAB rejects; the third is taken here as the function's own block
    size nr_per_OR1 (the last row's sampling length), which the recomputation
    of Output_max at :136-147 then reproduces.  Fewer than Nr_fit seeds raise
    (MATLAB's index error at :152)."""
    q_set = np.asarray(q_set, dtype=float).reshape(-1, 4)
    out1 = or_space_comp(geo, stack, pos, q_set, rot, hkl, nthreads=nthreads)
    order = sort_desc(out1[:, 2])
    K = min(int(nr_seed), q_set.shape[0])
    seeds = order[:K]
    if n_fit > K:
        raise IndexError(f"Nr_fit = {n_fit} fits need at least {n_fit} seeds (got {K})")
    q_all, gid, lens = _grid_rows(q_set[seeds], np.arange(1, K + 1, dtype=float), nr, gridsize)
    sc = or_space_comp(geo, stack, pos, q_all, rot, hkl, nthreads=nthreads)
    out9 = np.column_stack([sc, q_all, gid])
    # stage 1 maxima: Output is in q_all order (group ids ascending already)
    nper = lens[-1]
    omax = np.zeros((K, 9))
    for j in range(1, K + 1):
        if (gid == j).any():
            lo = nper * (j - 1)
            omax[j - 1] = out9[lo + matlab_max_index(out9[lo:nper * j, 2])]
    omax = omax[sort_desc(omax[:, 2])]
    for _ in range(int(iterations)):
        q2, gid2, lens2 = _grid_rows(omax[:, 4:8], omax[:, 8], regrid_nr, regrid_gridsize)
        sc2 = or_space_comp(geo, stack, pos, q2, rot, hkl, nthreads=nthreads)
        _, om = _group_max(np.column_stack([sc2, q2, gid2]), K, lens2[-1])
        omax = om[sort_desc(om[:, 2])]
    fits = []
    for j in range(int(n_fit)):
        x0 = omax[j, 5:8].copy()
        x, fval, nf, it = fit_or_local(geo, stack, pos, x0, rot, hkl, half_width, max_fun_evals, max_iter, tol_fun,
                                       tol_x)
        conv = fit_orientation(x)
        if conv is None:
            ns = ni = med = C = math.nan
            o10 = np.full(10, math.nan)
        else:
            o10, U = conv
            ns, ni, med = forward_comp(geo, stack, U, pos, rot, hkl)
            C = ni / ns if ns > 0 else math.nan
        fits.append(dict(x0=x0, x=np.asarray(x), fval=fval, nf=nf, it=it, group=int(omax[j, 8]), ns=ns, ni=ni, C=C,
                         med=med, o10=o10))
    best = matlab_max_index([f["C"] for f in fits])
    f = fits[best]
    row = np.concatenate([[0.0], np.asarray(pos, float), [f["ns"], f["ni"], f["C"], f["med"]], f["o10"]])
    return dict(indexing=row, candidate=f["group"], fit=f, fits=fits, seed_idx=seeds, output_max=omax)
