"""Spot pairing -- TEST INFRASTRUCTURE ONLY (checker of grs_pair_spots).

forward_comp_pair_spots.m:3-93 for one (U, pos): the selected events of
forward_comp.m (C oracle, with their wavelengths), and for every hit
(dis <= minEucDis/mean([pixelysize pixelzsize]), :70) the nearest spot of the
same projection (sqrt((y - dety)^2 + (z - detz)^2), first minimum, :73-74),
kept when that distance is below min(sqrt(area), 15) (:75).  Rows
[reshape(U,1,9) pos hkl' Energy rot_angle spot_y spot_z area dety detz dis]
(:78-79) with pos(2) - RotAxisOffset (:16), rounded to single precision (the
row is concatenated with the single-precision dis).
"""
from __future__ import annotations

import math

import numpy as np

from . import forward_comp


def pair_spots(geo, stack, U, pos, rot, hkl, frames):
    U = np.asarray(U, dtype=float).reshape(3, 3)
    hkl = np.asarray(hkl, dtype=float).reshape(-1, 3)
    ns, ni, med, ev = forward_comp(geo, stack, U, pos, rot, hkl, events=True)
    thr = geo.minEucDis / ((geo.pixelysize + geo.pixelzsize) / 2.0)
    p = [float(pos[0]), float(pos[1]) - geo.RotAxisOffset, float(pos[2])]
    rows = []
    for om, h, dy, dz, _y22, _z22, dis, lam in ev:
        if not dis <= thr:
            continue
        sp = np.asarray(frames[int(om)], dtype=float).reshape(-1, 3)
        if sp.shape[0] == 0:
            continue
        best, bi = math.nan, -1
        for k in range(sp.shape[0]):
            ey, ez = sp[k, 0] - dy, sp[k, 1] - dz
            d = math.sqrt(ey * ey + ez * ez)
            if d == d and (bi < 0 or d < best):
                best, bi = d, k
        if bi < 0:
            continue
        sa = math.sqrt(sp[bi, 2])
        lim = sa if (sa == sa and sa < 15.0) else 15.0
        if not best < lim:
            continue
        row = (list(U.flatten(order="F")) + p + list(hkl[int(h)]) + [12.398 / lam, float(rot[int(om)])] +
               list(sp[bi]) + [dy, dz, dis])
        rows.append(np.float32(row).astype(np.float64))
    return ns, ni, med, np.array(rows).reshape(-1, 23)
