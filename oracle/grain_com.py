"""Grain centre-of-mass fit -- TEST INFRASTRUCTURE ONLY (checker of
grs_fit_grain_com).

fit_grain_COM.m:5-89 with fit_grain_COM_only = 0: x0 = [position from
columns 11:13 of the first pair row, u2euler_corr(reshape(columns 2:10, 3, 3))]
(u2euler_corr.m:1-52), bounds x0 -/+ 0.1 mm and max(x0-3, 0) ..
min(x0+3, 360) degrees, fminsearchbnd (oracle.fminsearch) of
fit_grain_COM_fun: for every pair the event of its (hkl, omega) simulated
from pos = [x1, x2-RotAxisOffset, x3] and U = euler2u(x(4:6) degrees)
through the forward_comp.m statements, the pixel rounded, the distance to
the measured centroid; ErrMean = the sequential sum over the count.
Rows use the grs_pair_spots layout (23 values, no leading seed column).
norm() of a 3-vector is the square root of the sequential sum of squares
(parity unpinned against MATLAB's dnrm2 at the last bit, like forward_comp's).
"""
from __future__ import annotations

import math

import numpy as np

from . import euler2u_deg, math_eval
from .fminsearch import fminsearchbnd
from .orient import atan2

PI = float.fromhex("0x1.921fb54442d18p+1")


def _m(fn, x):
    return math_eval(fn, np.asarray(x, dtype=float))


def _euler2u_rad(p1, P, p2):
    c1, s1, c2, s2 = _m(1, p1)[0], _m(0, p1)[0], _m(1, p2)[0], _m(0, p2)[0]
    cP, sP = _m(1, P)[0], _m(0, P)[0]
    return np.array([[c1 * c2 - (s1 * s2) * cP, -c1 * s2 - (s1 * c2) * cP, s1 * sP],
                     [s1 * c2 + (c1 * s2) * cP, -s1 * s2 + (c1 * c2) * cP, -c1 * sP],
                     [s2 * sP, c2 * sP, cP]])


def u2euler_corr(U):
    U = np.asarray(U, dtype=float).reshape(3, 3)
    ph = float(_m(2, U[2, 2])[0])
    if ph < 0.0001:
        p2 = float(_m(3, U[1, 0] / U[0, 0])[0]) / 2.0
        return np.array([(p2 * 180.0) / PI, (ph * 180.0) / PI, (p2 * 180.0) / PI])
    a1 = atan2(-U[0, 2], U[1, 2])
    b1 = atan2(U[2, 0], U[2, 1])
    if a1 < 0.0:
        a1 = a1 + 2.0 * PI
        a2 = a1 - PI
    else:
        a2 = a1 + PI
    if b1 < 0.0:
        b1 = b1 + 2.0 * PI
        b2 = b1 - PI
    else:
        b2 = b1 + PI
    cands = [(a1, b1), (a2, b1), (a1, b2), (a2, b2)]
    best, bi = None, 0
    for k, (p1, p2) in enumerate(cands):
        Uk = _euler2u_rad(p1, ph, p2)
        tot = 0.0
        for c in range(3):
            col = 0.0
            for r in range(3):
                col = col + abs(Uk[r, c] - U[r, c])
            tot = tot + col
        if best is None or tot < best:
            best, bi = tot, k
    p1, p2 = cands[bi]
    return np.array([(p1 * 180.0) / PI, (ph * 180.0) / PI, (p2 * 180.0) / PI])


def _mround(x):
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def objective(geo, rows, x):
    S, B, R = np.asarray(geo.S, float), np.asarray(geo.B, float), np.asarray(geo.RotDet, float).reshape(-1)
    U = euler2u_deg(x[3], x[4], x[5])
    SU = np.array([[(S[i, 0] * U[0, j] + S[i, 1] * U[1, j]) + S[i, 2] * U[2, j] for j in range(3)] for i in range(3)])
    SUB = np.array([[(SU[i, 0] * B[0, j] + SU[i, 1] * B[1, j]) + SU[i, 2] * B[2, j] for j in range(3)]
                    for i in range(3)])
    rot, hkl, ye, ze = rows[:, 16], rows[:, 12:15], rows[:, 17], rows[:, 18]
    om = (rot * PI) / 180.0
    c, s = _m(1, om), _m(0, om)
    RAO, Lss, Lsd = geo.RotAxisOffset, geo.Lsam2sou, geo.Lsam2det
    L = Lss + Lsd
    p0, p1, p2 = x[0], x[1] - RAO, x[2]
    SW = []
    for r in range(3):
        if r == 0:
            o = [c * S[0, j] - s * S[1, j] for j in range(3)]
        elif r == 1:
            o = [s * S[0, j] + c * S[1, j] for j in range(3)]
        else:
            o = [np.full(c.shape, S[2, j]) for j in range(3)]
        SW.append((o[0] * p0 + o[1] * p1) + o[2] * p2)
    gp1, gp2, gp3 = Lss + SW[0], (SW[1] - geo.P0y) + RAO, SW[2] - geo.P0z
    ctr1, ctr2 = (gp2 * L) / gp1, (gp3 * L) / gp1
    alpha = _m(3, np.sqrt(gp2 * gp2 + gp3 * gp3) / gp1)
    ng = np.sqrt((gp1 * gp1 + gp2 * gp2) + gp3 * gp3)
    nc = np.sqrt((L * L + ctr1 * ctr1) + ctr2 * ctr2)
    u0, u1, u2 = gp1 / ng, gp2 / ng, gp3 / ng
    cu1, cu2 = ctr1 / nc, ctr2 / nc
    L2 = (Lsd - SW[0]) / _m(1, alpha)
    k0 = Lsd - SW[0]
    tnum = (R[0] * (Lsd - SW[0]) + R[3] * ((geo.dety00 - RAO) - SW[1])) + R[6] * (geo.detz00 - SW[2])
    ov1, ov2 = (SW[1] - geo.dety00) + RAO, SW[2] - geo.detz00
    gw = [(SUB[r, 0] * hkl[:, 0] + SUB[r, 1] * hkl[:, 1]) + SUB[r, 2] * hkl[:, 2] for r in range(3)]
    G0, G1, G2 = c * gw[0] - s * gw[1], s * gw[0] + c * gw[1], gw[2]
    Glen = np.sqrt((G0 * G0 + G1 * G1) + G2 * G2)
    beta = _m(2, (u0 * (G0 / Glen) + u1 * (G1 / Glen)) + u2 * (G2 / Glen))
    theta = beta - PI / 2.0
    sintth = _m(0, 2.0 * theta)
    nv = np.sqrt(G1 * G1 + G2 * G2)
    phix = _m(2, (G1 / nv) * cu1 + (G2 / nv) * cu2)
    diffvec = (L2 * sintth) / _m(0, phix - 2.0 * theta)
    y22 = ctr1 + ((diffvec * G1) / nv)
    z22 = ctr2 + ((diffvec * G2) / nv)
    k1, k2 = y22 - SW[1], z22 - SW[2]
    nk = np.sqrt((k0 * k0 + k1 * k1) + k2 * k2)
    K0, K1, K2 = k0 / nk, k1 / nk, k2 / nk
    t = tnum / ((R[0] * K0 + R[3] * K1) + R[6] * K2)
    w0, w1, w2 = t * K0 + (-k0), t * K1 + ov1, t * K2 + ov2
    Y22 = (R[1] * w0 + R[4] * w1) + R[7] * w2
    Z22 = (R[2] * w0 + R[5] * w1) + R[8] * w2
    ey = _mround((-Y22) / geo.pixelysize + geo.dety0) - ye
    ez = _mround((-Z22) / geo.pixelzsize + geo.detz0) - ze
    dis = np.sqrt(ey * ey + ez * ez)
    acc = 0.0
    for v in dis:
        acc = acc + float(v)
    return acc / dis.size


def fit_grain_com(geo, rows, max_fun_evals=2000, max_iter=500, tol_fun=1e-10, tol_x=1e-4):
    rows = np.asarray(rows, dtype=float).reshape(-1, 23)
    Ucm = rows[0, 0:9]
    x0 = np.concatenate([rows[0, 9:12], u2euler_corr(Ucm.reshape(3, 3, order="F"))])
    lb = np.concatenate([x0[:3] - 0.1, [max(v - 3.0, 0.0) for v in x0[3:]]])
    ub = np.concatenate([x0[:3] + 0.1, [min(v + 3.0, 360.0) for v in x0[3:]]])
    sin = lambda u: float(_m(0, u)[0])
    asin = lambda z: float(_m(7, z)[0])
    return fminsearchbnd(lambda x: objective(geo, rows, x), x0, lb, ub, max_fun_evals, max_iter, tol_fun, tol_x,
                         sin=sin, asin=asin)
