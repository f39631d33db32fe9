"""The reference's geometry fit on the CPU -- TEST INFRASTRUCTURE and the CPU
baseline of the geometry-fit leg of bench.py (never part of the engine).

geo_fitting.m:5-79: the objective of the 7 geometry values x = [Lsam2sou
Lsam2det dety00 detz00 tilt_x tilt_y tilt_z] (tilts in degrees,
get_det_R.m) is the mean Euclidean distance between the rounded simulated
pixel of every matched spot and its measured centroid, the grains (U, centre)
held.  P0y and P0z are both shifted by -RotAxisOffset (:19-20) and the
position's y by -RotAxisOffset (:32), and RotAxisOffset is not added back in
the spot construction (:36-39, :64-67).
L_shift_tilt_fitting.m:23-50 ('FitAllOnce'): fminsearchbnd (MaxFunEvals 2000,
MaxIter 500, TolFun 1e-10) within x0 -/+ [0.5 mm, 0.5 mm, 50 px, 50 px,
0.5 deg x 3] (the 50 px use pixelysize for both offsets, :30-31); while the
result lies within 0.1 of a bound the box is re-centred on it and the fit
repeated (:43-47).  GrainMapping_fit_geo.m:190-210 runs this from the current
geometry (iter_fit = 1) and keeps the best.

The grain-dependent terms (U, Gw, Omega*S*pos) are computed once with
numpy; every evaluation runs the geometry-dependent part in C with OpenMP
(oracle_geo_fitting, grs_geofit.c).  GeoObjective.sim_numpy restates that
part in numpy as an independent check.  Grain orientations are Rodrigues
vectors (U = quaternion2U(rod2quat(r))).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import lib
from .fminsearch import fminsearchbnd

_dp = ctypes.POINTER(ctypes.c_double)


def _U_from_rod(r):
    s = 1.0 / np.sqrt(1.0 + (r * r).sum(-1))
    q0, q1, q2, q3 = s, r[:, 0] * s, r[:, 1] * s, r[:, 2] * s
    return np.stack([
        np.stack([2 * (q0 * q0 + q1 * q1) - 1, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)], -1),
        np.stack([2 * (q1 * q2 + q0 * q3), 2 * (q0 * q0 + q2 * q2) - 1, 2 * (q2 * q3 - q0 * q1)], -1),
        np.stack([2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 2 * (q0 * q0 + q3 * q3) - 1], -1),
    ], 1)


def _rot_det(t_deg):
    """get_det_R.m: RotZ*RotY*RotX, angles in degrees."""
    a = np.asarray(t_deg, float) * (math.pi / 180.0)
    cx, sx, cy, sy, cz, sz = math.cos(a[0]), math.sin(a[0]), math.cos(a[1]), math.sin(a[1]), math.cos(a[2]), \
        math.sin(a[2])
    RX = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    RY = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    RZ = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return RZ @ RY @ RX


class GeoObjective:
    """geo_fitting.m for a fixed spot set and fixed grains."""

    def __init__(self, geo, hkl, omega_deg, meas, grain, gr6, nthreads=8):
        S, B = np.asarray(geo.S, float), np.asarray(geo.B, float)
        gr6 = np.asarray(gr6, float)
        gi = np.asarray(grain, int)
        U = _U_from_rod(gr6[:, :3])[gi]
        om = np.asarray(omega_deg, float) * (math.pi / 180.0)
        c, s = np.cos(om), np.sin(om)
        self.Gw = np.einsum("ij,njk,kl,nl->ni", S, U, B, np.asarray(hkl, float))
        RAO = geo.RotAxisOffset
        pos = gr6[gi, 3:].copy()
        pos[:, 1] -= RAO
        p = pos @ S.T
        self.SW = np.stack([c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1], p[:, 2]], 1)
        self.Gt = np.stack([c * self.Gw[:, 0] - s * self.Gw[:, 1], s * self.Gw[:, 0] + c * self.Gw[:, 1],
                            self.Gw[:, 2]], 1)
        self.P0y, self.P0z = geo.P0y - RAO, geo.P0z - RAO
        self.py, self.pz, self.dety0, self.detz0 = geo.pixelysize, geo.pixelzsize, geo.dety0, geo.detz0
        self.meas = np.ascontiguousarray(meas, dtype=float)
        self.n = self.meas.shape[0]
        self.evals = 0
        self.nthreads = int(nthreads)
        self.SW = np.ascontiguousarray(self.SW)
        self.Gt = np.ascontiguousarray(self.Gt)

    def _c(self, x, sim):
        x = np.ascontiguousarray(x, dtype=float)
        return lib().oracle_geo_fitting(self.SW.ctypes.data_as(_dp), self.Gt.ctypes.data_as(_dp),
                                        self.meas.ctypes.data_as(_dp), self.n, x.ctypes.data_as(_dp), self.P0y,
                                        self.P0z, self.py, self.pz, self.dety0, self.detz0,
                                        sim.ctypes.data_as(_dp) if sim is not None else None, self.nthreads)

    def sim(self, x):
        """Unrounded simulated pixel coordinates (n, 2) at geometry x."""
        out = np.zeros((self.n, 2))
        self._c(x, out)
        return out

    def sim_numpy(self, x):
        """Unrounded simulated pixel coordinates (n, 2) at geometry x."""
        Lss, Lsd, dy00, dz00 = (float(v) for v in x[:4])
        R = _rot_det(x[4:7])
        SW, Gt = self.SW, self.Gt
        L = Lss + Lsd
        gp0 = Lss + SW[:, 0]
        gp1 = SW[:, 1] - self.P0y
        gp2 = SW[:, 2] - self.P0z
        c1, c2 = gp1 * L / gp0, gp2 * L / gp0
        alpha = np.arctan(np.sqrt(gp1 * gp1 + gp2 * gp2) / gp0)
        ng = np.sqrt(gp0 * gp0 + gp1 * gp1 + gp2 * gp2)
        glen = np.sqrt((Gt * Gt).sum(1))
        beta = np.arccos((gp0 * Gt[:, 0] + gp1 * Gt[:, 1] + gp2 * Gt[:, 2]) / ng / glen)
        theta = beta - math.pi / 2
        nv = np.sqrt(Gt[:, 1] ** 2 + Gt[:, 2] ** 2)
        nc = np.sqrt(L * L + c1 * c1 + c2 * c2)
        phix = np.arccos((Gt[:, 1] * c1 + Gt[:, 2] * c2) / nv / nc)
        phiy = phix - 2 * theta
        L2 = (Lsd - SW[:, 0]) / np.cos(alpha)
        diffvec = L2 * np.sin(2 * theta) / np.sin(phiy)
        y22 = c1 + diffvec * Gt[:, 1] / nv
        z22 = c2 + diffvec * Gt[:, 2] / nv
        k = np.stack([Lsd - SW[:, 0], y22 - SW[:, 1], z22 - SW[:, 2]], 1)
        K = k / np.sqrt((k * k).sum(1))[:, None]
        t = (R[0, 0] * (Lsd - SW[:, 0]) + R[1, 0] * (dy00 - SW[:, 1]) + R[2, 0] * (dz00 - SW[:, 2])) / \
            (R[0, 0] * K[:, 0] + R[1, 0] * K[:, 1] + R[2, 0] * K[:, 2])
        w = t[:, None] * K + np.stack([SW[:, 0] - Lsd, SW[:, 1] - dy00, SW[:, 2] - dz00], 1)
        Y22 = R[0, 1] * w[:, 0] + R[1, 1] * w[:, 1] + R[2, 1] * w[:, 2]
        Z22 = R[0, 2] * w[:, 0] + R[1, 2] * w[:, 1] + R[2, 2] * w[:, 2]
        return np.stack([-Y22 / self.py + self.dety0, -Z22 / self.pz + self.detz0], 1)

    def __call__(self, x):
        """ErrMean of geo_fitting.m: mean distance of the rounded pixel (half away from zero)."""
        self.evals += 1
        return float(self._c(x, None))

    def rms(self, x):
        r = self.sim(x) - self.meas
        return float(np.sqrt(np.mean((r * r).sum(1))))


def fit_all_once(obj, x0, pixelysize, max_rounds=10):
    """L_shift_tilt_fitting.m:23-50; at most max_rounds re-centred fits (the
    reference has no cap).  Returns (x, fval, evaluations, rounds)."""
    x0 = np.asarray(x0, float).copy()
    half = np.array([0.5, 0.5, pixelysize * 50, pixelysize * 50, 0.5, 0.5, 0.5])
    total, rounds = 0, 0
    while True:
        lb, ub = x0 - half, x0 + half
        x, fval, nf, _ = fminsearchbnd(obj, x0, lb, ub, max_fun_evals=2000, max_iter=500, tol_fun=1e-10)
        total += nf
        rounds += 1
        if (np.all(np.abs(x - lb) > 0.1) and np.all(np.abs(x - ub) > 0.1)) or rounds >= max_rounds:
            return x, fval, total, rounds
        x0 = np.asarray(x, float).copy()
