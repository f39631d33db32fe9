"""Region growth around an indexed seed -- TEST INFRASTRUCTURE ONLY (checker
of grs_grow_region).

grow_indexed_region_cuda.m:29-151 with numpy volumes indexed [x-1, y-1, z-1]
and the C oracle's cuda_grow_indexed_region (oracle.grow).  Sums and norms of
3-vectors are sequential.
"""
from __future__ import annotations

import math

import numpy as np

from . import grow


def _mround(x: float) -> float:
    """MATLAB round: halfway away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def grow_region(geo, stack, vols, seed, indexed_comp, U, spot_L, drop_off, maxD, maxDmedian, p1, rot, hkl):
    """vols: dict Completeness, Mask, Dismedian, Ninter (updated in place the
    way the caller copies the outputs back).  Returns (center, replace_center,
    indexed (n, 3) 1-based in find() order, n_checked)."""
    I, Imask, DM, NI = vols["Completeness"], vols["Mask"], vols["Dismedian"], vols["Ninter"]
    x, y, z = [int(v) for v in seed]
    I[x - 1, y - 1, z - 1] = indexed_comp
    J = np.zeros(I.shape, bool)
    J[x - 1, y - 1, z - 1] = True
    seed_comp = indexed_comp
    s_sum = seed_comp
    s_pos = [seed_comp * x, seed_comp * y, seed_comp * z]
    c0 = [v / s_sum for v in s_pos]
    py, pz = geo.pixelysize, geo.pixelzsize
    box_l = math.ceil((spot_L * ((py + pz) / 2.0)) / ((geo.Lsam2det / geo.Lsam2sou) * p1[4]))
    lo, hi = [], []
    for k, pk in enumerate((x, y, z)):
        lo.append(pk - box_l if pk - box_l > 0 else 1)
        hi.append(pk + box_l if pk + box_l < I.shape[k] else I.shape[k])
    cp = [_mround(float(hi[k] - lo[k])) / 2.0 + lo[k] for k in range(3)]
    rows = []
    n_checked = 0
    for zz in range(lo[2], hi[2] + 1):
        for yy in range(lo[1], hi[1] + 1):
            for xx in range(lo[0], hi[0] + 1):
                ex, ey, ez = xx - cp[0], yy - cp[1], zz - cp[2]
                d = math.sqrt((ex * ex + ey * ey) + ez * ez)
                if d <= box_l:
                    n_checked += 1
                    rows.append([xx, yy, zz, I[xx - 1, yy - 1, zz - 1], Imask[xx - 1, yy - 1, zz - 1],
                                 DM[xx - 1, yy - 1, zz - 1]])
    rows = np.array(rows, dtype=float).reshape(-1, 6)
    param1 = np.concatenate([np.asarray(p1, float)[:11], [seed_comp, drop_off, maxDmedian]])
    out = grow(geo, stack, rows, U, rot, hkl, param1, nthreads=8) if rows.shape[0] else np.zeros((0, 7))
    for o in out[out[:, 6] == 1]:
        xn, yn, zn = int(o[0]) - 1, int(o[1]) - 1, int(o[2]) - 1
        J[xn, yn, zn] = True
        I[xn, yn, zn] = o[4] / o[3]
        DM[xn, yn, zn] = o[5]
        NI[xn, yn, zn] = o[4]
    lin = np.flatnonzero(J.reshape(-1, order="F"))
    idx = np.stack(np.unravel_index(lin, J.shape, order="F"), 1) + 1
    if idx.shape[0] > 1:
        for xi, yi, zi in idx:
            cv = I[xi - 1, yi - 1, zi - 1]
            s_sum = s_sum + cv
            s_pos = [s_pos[0] + cv * xi, s_pos[1] + cv * yi, s_pos[2] + cv * zi]
    c1 = [v / s_sum for v in s_pos]
    e0, e1, e2 = c1[0] - c0[0], c1[1] - c0[1], c1[2] - c0[2]
    dist = math.sqrt((e0 * e0 + e1 * e1) + e2 * e2)
    if dist > max(maxD, maxD / 10000 * idx.shape[0]):
        return np.array([_mround(v) for v in c1]), 1, idx, n_checked
    return np.array(c0), 0, idx, n_checked
