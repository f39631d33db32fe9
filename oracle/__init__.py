"""CPU oracle for the completeness scorer -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker; the engine (grainrecon_amd) never
imports, links or calls it.

Two independent restatements of GrainRecon's forward_comp.m:
  * liboracle.so (grs_oracle.c): statement-by-statement FP64 C restatement,
    OpenMP over orientations / voxels; also used as the CPU baseline;
  * np_restatement.py: a vectorised numpy restatement written separately,
    used to cross-check the C oracle's formula order and indexing.

Pins (tests/test_oracle.py):
  * closed form: an on-axis voxel with an untilted detector must put every
    spot where the mirror reflection k_out = k_in - 2 (k_in . g) g hits the
    plane x = Lsam2det (forward_comp.m:40-47 reduces to that construction);
  * tilted detector: the spot must be the ray/plane intersection of the
    diffracted ray with the plane of normal RotDet(:,1) (forward_comp.m:49-55);
  * numpy restatement agreement on random inputs (counts, medians, pixels);
  * invariants: 0 <= Nr_intersect <= Nr_simu <= Nomega*H, Friedel exclusivity,
    determinism, cubic-symmetry invariance of counts;
  * known answer: a mask rendered from (voxel, U) scores C = 1 at (voxel, U);
  * the FP64 math (grs_math.h) against mpmath to < 1 ulp.
The MATLAB built-ins themselves (libm last bits, BLAS summation order,
median of an even count) cannot be run here; those choices are documented in
grs_oracle.c and counted as unpinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64 = ctypes.c_int64


class OracleGeo(ctypes.Structure):
    _fields_ = [
        ("S", ctypes.c_double * 9), ("B", ctypes.c_double * 9), ("RotDet", ctypes.c_double * 9),
        ("thetamax", ctypes.c_double), ("lambda_min", ctypes.c_double), ("lambda_max", ctypes.c_double),
        ("Lsam2sou", ctypes.c_double), ("Lsam2det", ctypes.c_double), ("minEucDis", ctypes.c_double),
        ("dety00", ctypes.c_double), ("detz00", ctypes.c_double), ("P0y", ctypes.c_double),
        ("P0z", ctypes.c_double), ("RotAxisOffset", ctypes.c_double),
        ("pixelysize", ctypes.c_double), ("pixelzsize", ctypes.c_double), ("dety0", ctypes.c_double),
        ("detz0", ctypes.c_double), ("detysize", ctypes.c_double), ("detzsize", ctypes.c_double),
        ("BeamStopY", ctypes.c_double * 2), ("BeamStopZ", ctypes.c_double * 2),
    ]


class OracleProj(ctypes.Structure):
    _fields_ = [("proj", _fp), ("nz", _i64), ("ny", _i64), ("nrot", _i64)]


_lib = None


def build() -> str:
    subprocess.run(["sh", os.path.join(HERE, "build_oracle.sh")], check=True, capture_output=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_forward_comp.restype = _i64
        L.oracle_forward_comp.argtypes = [ctypes.POINTER(OracleGeo), ctypes.POINTER(OracleProj), _dp, _dp, _dp, _i64,
                                          _dp, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _fp, _dp, _i64]
        L.oracle_or_space_comp.restype = None
        L.oracle_or_space_comp.argtypes = [ctypes.POINTER(OracleGeo), ctypes.POINTER(OracleProj), _dp, _dp, _i64,
                                           _dp, _i64, _dp, _i64, _dp, ctypes.c_int]
        L.oracle_grow.restype = None
        L.oracle_grow.argtypes = [ctypes.POINTER(OracleGeo), ctypes.POINTER(OracleProj), _dp, _i64, _dp, _dp, _i64,
                                  _dp, _i64, _dp, _dp, ctypes.c_int]
        L.oracle_compete.restype = None
        L.oracle_compete.argtypes = [ctypes.POINTER(OracleGeo), ctypes.POINTER(OracleProj), _dp, _i64, _dp, _dp,
                                     _dp, _i64, _dp, _i64, _dp, _dp, ctypes.c_int]
        L.oracle_margin_audit.restype = None
        L.oracle_margin_audit.argtypes = [ctypes.POINTER(OracleGeo), ctypes.POINTER(OracleProj), _dp, _dp, _i64,
                                          _dp, _i64, _dp, _i64, _dp, ctypes.c_int]
        L.oracle_geo_fitting.restype = ctypes.c_double
        L.oracle_geo_fitting.argtypes = [_dp, _dp, _dp, _i64, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]
        L.oracle_math_eval.restype = None
        L.oracle_math_eval.argtypes = [ctypes.c_int, _dp, _dp, _i64]
        L.oracle_set_libm.restype = None
        L.oracle_set_libm.argtypes = [ctypes.c_int]
        L.oracle_euler2u_deg.restype = None
        L.oracle_euler2u_deg.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.oracle_colon.restype = _i64
        L.oracle_colon.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _i64]
        L.oracle_local_sampling.restype = _i64
        L.oracle_local_sampling.argtypes = [_dp, ctypes.c_int32, ctypes.c_double, _dp, _i64]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.oracle_find_neighbor_ors.restype = _i64
        L.oracle_find_neighbor_ors.argtypes = [_i32p, _i64, _i64, _i64, _i32p, _i64, ctypes.c_int32, ctypes.c_int32,
                                               _i32p, ctypes.POINTER(_i64)]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def geo_from(geo) -> OracleGeo:
    """OracleGeo from a grainrecon_amd.geometry.Geometry (or any object with the same fields)."""
    g = OracleGeo()
    for name in ("S", "B", "RotDet"):
        getattr(g, name)[:] = np.asarray(getattr(geo, name), dtype=float).reshape(9).tolist()
    for f in ("thetamax", "lambda_min", "lambda_max", "Lsam2sou", "Lsam2det", "minEucDis", "dety00", "detz00",
              "P0y", "P0z", "RotAxisOffset", "pixelysize", "pixelzsize", "dety0", "detz0", "detysize",
              "detzsize"):
        setattr(g, f, float(getattr(geo, f)))
    g.BeamStopY[:] = [float(x) for x in geo.BeamStopY]
    g.BeamStopZ[:] = [float(x) for x in geo.BeamStopZ]
    return g


class Stack:
    """Keeps a MATLAB-order float32 copy of a (nz, ny, nrot) distance stack alive."""

    def __init__(self, proj: np.ndarray):
        proj = np.asarray(proj, dtype=np.float32)
        self.shape = proj.shape
        self.flat = np.ascontiguousarray(proj.reshape(-1, order="F"))
        self.c = OracleProj(self.flat.ctypes.data_as(_fp), proj.shape[0], proj.shape[1], proj.shape[2])


def forward_comp(geo, stack: Stack, U, pos, rot, hkl, events: bool = False):
    """forward_comp.m for one (U, pos).  Returns (Nr_simu, Nr_intersect,
    dis_median[, events (n, 8): omega idx, hkl idx, dety, detz, dety22, detz22, dis, lambdahkl])."""
    g = geo_from(geo)
    U = _d(np.asarray(U).reshape(9))
    pos = _d(np.asarray(pos).reshape(3))
    rot = _d(np.asarray(rot).reshape(-1))
    hkl = _d(np.asarray(hkl).reshape(-1, 3))
    ns, ni = _i64(), _i64()
    med = ctypes.c_float()
    cap = rot.size * hkl.shape[0] if events else 0
    ev = np.zeros((max(cap, 1), 8))
    n = lib().oracle_forward_comp(ctypes.byref(g), ctypes.byref(stack.c), U.ctypes.data_as(_dp),
                                  pos.ctypes.data_as(_dp), rot.ctypes.data_as(_dp), rot.size,
                                  hkl.ctypes.data_as(_dp), hkl.shape[0], ctypes.byref(ns), ctypes.byref(ni),
                                  ctypes.byref(med), ev.ctypes.data_as(_dp) if events else None, cap)
    if events:
        return ns.value, ni.value, med.value, ev[:n].copy()
    return ns.value, ni.value, med.value


def or_space_comp(geo, stack: Stack, pos, q, rot, hkl, nthreads: int = 8) -> np.ndarray:
    """ForwardCalc_OR_space_comp.m: (N, 4) [Nr_simu Nr_intersect C dis_median]."""
    g = geo_from(geo)
    pos = _d(np.asarray(pos).reshape(3))
    q = _d(np.asarray(q).reshape(-1, 4))
    rot = _d(np.asarray(rot).reshape(-1))
    hkl = _d(np.asarray(hkl).reshape(-1, 3))
    out = np.zeros((q.shape[0], 4))
    lib().oracle_or_space_comp(ctypes.byref(g), ctypes.byref(stack.c), pos.ctypes.data_as(_dp),
                               q.ctypes.data_as(_dp), q.shape[0], rot.ctypes.data_as(_dp), rot.size,
                               hkl.ctypes.data_as(_dp), hkl.shape[0], out.ctypes.data_as(_dp), int(nthreads))
    return out


def grow(geo, stack: Stack, vox6, U, rot, hkl, param1, nthreads: int = 8) -> np.ndarray:
    g = geo_from(geo)
    vox = _d(np.asarray(vox6).reshape(-1, 6))
    U = _d(np.asarray(U).reshape(9))
    rot = _d(np.asarray(rot).reshape(-1))
    hkl = _d(np.asarray(hkl).reshape(-1, 3))
    p1 = _d(np.asarray(param1).reshape(14))
    out = np.zeros((vox.shape[0], 7))
    lib().oracle_grow(ctypes.byref(g), ctypes.byref(stack.c), vox.ctypes.data_as(_dp), vox.shape[0],
                      U.ctypes.data_as(_dp), rot.ctypes.data_as(_dp), rot.size, hkl.ctypes.data_as(_dp),
                      hkl.shape[0], p1.ctypes.data_as(_dp), out.ctypes.data_as(_dp), int(nthreads))
    return out


def compete(geo, stack: Stack, vox5, id_all, euler, rot, hkl, param1, nthreads: int = 8) -> np.ndarray:
    g = geo_from(geo)
    vox = _d(np.asarray(vox5).reshape(-1, 5))
    ids = _d(np.asarray(id_all).reshape(-1))
    eu = _d(np.asarray(euler).reshape(-1, 3))
    rot = _d(np.asarray(rot).reshape(-1))
    hkl = _d(np.asarray(hkl).reshape(-1, 3))
    p1 = _d(np.asarray(param1).reshape(-1)[:11])
    out = np.zeros((vox.shape[0], 8))
    lib().oracle_compete(ctypes.byref(g), ctypes.byref(stack.c), vox.ctypes.data_as(_dp), vox.shape[0],
                         ids.ctypes.data_as(_dp), eu.ctypes.data_as(_dp), rot.ctypes.data_as(_dp), rot.size,
                         hkl.ctypes.data_as(_dp), hkl.shape[0], p1.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                         int(nthreads))
    return out


AUDIT_FIELDS = ("events", "events_differing", "orientations_differing", "max_dbeta_rad", "max_dlambda_rel",
                "max_dpixel_px", "min_margin_beta_low_rad", "min_margin_beta_high_rad", "min_margin_lambda_min_rel",
                "min_margin_lambda_max_rel", "min_margin_rounding_px", "events_in_window",
                "min_ratio_beta_low", "min_ratio_beta_high", "min_ratio_lambda_min", "min_ratio_lambda_max",
                "min_ratio_rounding")


def margin_audit(geo, stack: Stack, pos, q, rot, hkl, nthreads: int = 8) -> dict:
    """Every event of one voxel x N orientation rows computed with the
    project's FP64 math and with the C library's: the decisions and counts
    that differ, the largest differences of beta, lambda and the raw pixel
    coordinates, and the smallest margin of every continuous decision
    (oracle_margin_audit in grs_oracle.c)."""
    g = geo_from(geo)
    pos = _d(np.asarray(pos).reshape(3))
    q = _d(np.asarray(q).reshape(-1, 4))
    rot = _d(np.asarray(rot).reshape(-1))
    hkl = _d(np.asarray(hkl).reshape(-1, 3))
    out = np.zeros(len(AUDIT_FIELDS))
    lib().oracle_margin_audit(ctypes.byref(g), ctypes.byref(stack.c), pos.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                              q.shape[0], rot.ctypes.data_as(_dp), rot.size, hkl.ctypes.data_as(_dp), hkl.shape[0],
                              out.ctypes.data_as(_dp), int(nthreads))
    return dict(zip(AUDIT_FIELDS, out.tolist()))


def math_eval(fn: int, x) -> np.ndarray:
    """The engine's FP64 math evaluated on the host: 0 sin, 1 cos, 2 acos, 3 atan, 4 round, 7 asin."""
    x = _d(np.asarray(x).reshape(-1))
    y = np.zeros_like(x)
    lib().oracle_math_eval(int(fn), x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), x.size)
    return y


def set_libm(on: bool) -> None:
    """Switch the oracle's transcendentals to the C library (sensitivity studies only)."""
    lib().oracle_set_libm(1 if on else 0)


def euler2u_deg(p1, P, p2) -> np.ndarray:
    U = np.zeros(9)
    lib().oracle_euler2u_deg(float(p1), float(P), float(p2), U.ctypes.data_as(_dp))
    return U.reshape(3, 3)


def colon(a: float, d: float, b: float) -> np.ndarray:
    """MATLAB a:d:b (oracle_colon)."""
    v = np.zeros(4096)
    n = lib().oracle_colon(float(a), float(d), float(b), v.ctypes.data_as(_dp), v.size)
    if n < 0:
        raise ValueError("colon: too many elements or non-finite argument")
    return v[:n].copy()


def local_sampling(centre, nr: int, gridsize: float) -> np.ndarray:
    """ori_local_sampling.m:12-34 -> OR_local.q (sorted unique rows)."""
    c = np.asarray(centre, dtype=float).reshape(-1)
    if c.size == 3:
        c = np.array([np.sqrt(1.0 - ((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2])), *c])
    c = _d(c)
    cap = 1 + (int(nr) + 3) ** 3
    out = np.zeros((cap, 4))
    n = lib().oracle_local_sampling(c.ctypes.data_as(_dp), int(nr), float(gridsize), out.ctypes.data_as(_dp), cap)
    if n < 0:
        raise ValueError("local_sampling failed")
    return out[:n].copy()


def find_neighbor_ors(grain_id: np.ndarray, vox, search_radius: int, nr_max: int = 30):
    """find_neighbor_ORs.m:3-96.  grain_id: (nx, ny, nz) array indexed like
    DS_in.GrainId(x, y, z); vox: (M, 3) 1-based [x y z].  Returns
    (id_neigb_all (K,), id_neigb_ind (M, 2) 1-based [first last])."""
    g = np.asarray(grain_id)
    nx, ny, nz = g.shape
    gf = np.ascontiguousarray(g.reshape(-1, order="F"), dtype=np.int32)
    v = np.ascontiguousarray(np.asarray(vox).reshape(-1, 3), dtype=np.int32)
    M = v.shape[0]
    ids = np.zeros(max(M * nr_max, 1), np.int32)
    ind = np.zeros((M, 2), np.int64)
    i32 = ctypes.POINTER(ctypes.c_int32)
    n = lib().oracle_find_neighbor_ors(gf.ctypes.data_as(i32), nx, ny, nz, v.ctypes.data_as(i32), M,
                                       int(search_radius), int(nr_max), ids.ctypes.data_as(i32),
                                       ind.ctypes.data_as(ctypes.POINTER(_i64)))
    if n < 0:
        raise ValueError("the grain-id volume holds no positive id")
    return ids[:n].copy(), ind
