"""Host-side mirror of GrainRecon's native scoring operators on the gfx950 engine.

Two layers:

* ``Engine``: a device context with the geometry, reflections, rotation angles
  and the distance stack resident in HBM (FP64 native entries).
* ``cuda_forward_comp``, ``cuda_grow_indexed_region``,
  ``cuda_forward_comp_compete``: the three MEX operators with their exact
  argument lists and single-precision outputs (call sites
  ForwardCalc_OR_space_comp_cuda.m:32, grow_indexed_region_cuda.m:105-107,
  grow_unindexed_compete_comp_cuda.m:39-41), so the MATLAB wrappers' logic can
  be driven unchanged from Python.

Everything runs on the GPU through libgrscore.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .geometry import Geometry

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)


def _d(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def stack_to_matlab_order(proj: np.ndarray) -> np.ndarray:
    """A distance stack indexed like proj_bin_bw(detz, dety, i), i.e. a numpy
    array of shape (nz, ny, nrot), flattened in MATLAB column-major order."""
    proj = np.asarray(proj, dtype=np.float32)
    if proj.ndim != 3:
        raise ValueError("distance stack must be 3-D (detz, dety, omega)")
    return proj.reshape(-1, order="F")


def fit_orientations(x):
    """fit_OR_constrain_with_local_comp_cuda.m:133-148 on the host for fitted
    vector parts x (M, 3): returns (out10 (M, 10) = [Euler ZXZ deg, Rodrigues,
    quaternion], U (M, 3, 3), ok (M,))."""
    x = _d(np.asarray(x).reshape(-1, 3))
    M = x.shape[0]
    out = np.zeros((M, 10))
    U = np.zeros((M, 9))
    ok = np.zeros(M, np.int32)
    code = L.lib().grs_fit_orientations(_ptr(x, _dp), M, _ptr(out, _dp), _ptr(U, _dp), _ptr(ok, _ip))
    if code != 0:
        raise ValueError("grs_fit_orientations: invalid argument")
    return out, U.reshape(M, 3, 3), ok


def euler_to_rod_quat(euler_deg):
    """rod = angle2rod(euler*pi/180,'ZXZ'), quat = rod2quat(rod) for (M, 3)
    degrees: returns ((M, 3) rod, (M, 4) quat, ok (M,))."""
    e = _d(np.asarray(euler_deg).reshape(-1, 3))
    out = np.zeros((e.shape[0], 7))
    ok = np.zeros(e.shape[0], np.int32)
    if L.lib().grs_euler_to_rod_quat(_ptr(e, _dp), e.shape[0], _ptr(out, _dp), _ptr(ok, _ip)) != 0:
        raise ValueError("grs_euler_to_rod_quat: invalid argument")
    return out[:, :3].copy(), out[:, 3:].copy(), ok


def euler2u_deg(euler_deg):
    """euler2u.m of (M, 3) ZXZ angles in degrees with the engine's FP64
    functions: (M, 3, 3)."""
    e = _d(np.asarray(euler_deg).reshape(-1, 3))
    U = np.zeros((e.shape[0], 9))
    if L.lib().grs_euler2u_deg(_ptr(e, _dp), e.shape[0], _ptr(U, _dp)) != 0:
        raise ValueError("grs_euler2u_deg: invalid argument")
    return U.reshape(-1, 3, 3)


def u2euler(U_colmajor):
    """u2euler_corr.m for matrices given column-major (M, 9): (M, 3) degrees."""
    U = _d(np.asarray(U_colmajor).reshape(-1, 9))
    out = np.zeros((U.shape[0], 3))
    if L.lib().grs_u2euler(_ptr(U, _dp), U.shape[0], _ptr(out, _dp)) != 0:
        raise ValueError("grs_u2euler: invalid argument")
    return out


class Engine:
    """FP64 scoring context on one GPU (grs_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = ctypes.c_void_p()
        L.check(self._lib.grs_create(int(device), ctypes.byref(h)))
        self._h = h
        self.H = 0
        self.nrot = 0
        self._vshape = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.grs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        L.check(code, self._h)

    # -- state ------------------------------------------------------------
    def set_geometry(self, geo: Geometry, hkl: np.ndarray, rot_angles: np.ndarray) -> None:
        hkl = _d(np.asarray(hkl).reshape(-1, 3))
        rot = _d(np.asarray(rot_angles).reshape(-1))
        g = geo.to_c()
        self._chk(self._lib.grs_set_geometry(self._h, ctypes.byref(g), _ptr(hkl, _dp), hkl.shape[0],
                                             _ptr(rot, _dp), rot.shape[0]))
        self.H, self.nrot = hkl.shape[0], rot.shape[0]

    def set_distance_stack(self, proj: np.ndarray) -> None:
        """proj: (nz, ny, nrot) array indexed like proj_bin_bw(detz, dety, i)."""
        proj = np.asarray(proj, dtype=np.float32)
        nz, ny, nr = proj.shape
        flat = np.ascontiguousarray(proj.reshape(-1, order="F"))
        self._chk(self._lib.grs_set_distance_stack(self._h, flat.ctypes.data, nz, ny, nr, 0))

    def set_distance_stack_device(self, dptr: int, nz: int, ny: int, nrot: int) -> None:
        """Copies a device-resident MATLAB-order float stack (e.g. a torch tensor's data_ptr())."""
        self._chk(self._lib.grs_set_distance_stack(self._h, ctypes.c_void_p(int(dptr)), nz, ny, nrot, 1))

    # -- operators ----------------------------------------------------------
    def forward_comp(self, pos, q, method: int = L.METHOD_FAST):
        """One voxel (pos, mm) x N orientations [q0 q1 q2 q3] ->
        (Nr_simu, Nr_intersect, C, dis_median)."""
        pos = _d(np.asarray(pos).reshape(3))
        q = _d(np.asarray(q).reshape(-1, 4))
        N = q.shape[0]
        ns = np.zeros(N, np.int32)
        ni = np.zeros(N, np.int32)
        C = np.zeros(N, np.float64)
        med = np.zeros(N, np.float32)
        self._chk(self._lib.grs_forward_comp(self._h, _ptr(pos, _dp), _ptr(q, _dp), N, _ptr(ns, _ip),
                                             _ptr(ni, _ip), _ptr(C, _dp), _ptr(med, _fp), int(method)))
        return ns, ni, C, med

    def forward_comp_device(self, pos, d_q: int, N: int, d_ns: int, d_ni: int, d_C: int, d_med: int,
                            method: int = L.METHOD_FAST) -> None:
        pos = _d(np.asarray(pos).reshape(3))
        self._chk(self._lib.grs_forward_comp_device(self._h, _ptr(pos, _dp), ctypes.c_void_p(d_q), int(N),
                                                    ctypes.c_void_p(d_ns), ctypes.c_void_p(d_ni),
                                                    ctypes.c_void_p(d_C), ctypes.c_void_p(d_med), int(method)))

    def forward_comp_device_noalloc(self, pos, d_q: int, N: int, d_ns: int, d_ni: int, d_C: int,
                                    d_med: int) -> None:
        """grs_forward_comp_device_noalloc: the fast method with no device
        allocation (working memory in the library's static per-GPU scratch)."""
        pos = _d(np.asarray(pos).reshape(3))
        self._chk(self._lib.grs_forward_comp_device_noalloc(self._h, _ptr(pos, _dp), ctypes.c_void_p(d_q), int(N),
                                                            ctypes.c_void_p(d_ns), ctypes.c_void_p(d_ni),
                                                            ctypes.c_void_p(d_C), ctypes.c_void_p(d_med)))

    def grow_indexed_region(self, vox6, U, param1, method: int = L.METHOD_FAST) -> np.ndarray:
        vox = _d(np.asarray(vox6).reshape(-1, 6))
        U = _d(np.asarray(U).reshape(9))
        p1 = _d(np.asarray(param1).reshape(14))
        out = np.zeros((vox.shape[0], 7))
        self._chk(self._lib.grs_grow_indexed_region(self._h, _ptr(vox, _dp), vox.shape[0], _ptr(U, _dp),
                                                    _ptr(p1, _dp), _ptr(out, _dp), int(method)))
        return out

    def forward_comp_compete(self, vox5, id_all, euler, param1, method: int = L.METHOD_FAST) -> np.ndarray:
        vox = _d(np.asarray(vox5).reshape(-1, 5))
        ids = _d(np.asarray(id_all).reshape(-1))
        eu = _d(np.asarray(euler).reshape(-1, 3))
        p1 = _d(np.asarray(param1).reshape(-1)[:11])
        out = np.zeros((vox.shape[0], 8))
        self._chk(self._lib.grs_forward_comp_compete(self._h, _ptr(vox, _dp), vox.shape[0], _ptr(ids, _dp),
                                                     ids.shape[0], _ptr(eu, _dp), eu.shape[0], _ptr(p1, _dp),
                                                     _ptr(out, _dp), int(method)))
        return out

    def compete_csr(self, pos, row_ptr, cand_q, nr: int, gridsize: float):
        """Batched compete with local sampling over CSR rows (grs_compete_csr).
        Returns (best index in the row's expanded list, best quaternion,
        Nr_simu, Nr_intersect, C, dis_median) per row."""
        pos = _d(np.asarray(pos).reshape(-1, 3))
        rp = np.ascontiguousarray(np.asarray(row_ptr).reshape(-1), dtype=np.int64)
        cq = _d(np.asarray(cand_q).reshape(-1, 4))
        M = pos.shape[0]
        if rp.size != M + 1:
            raise ValueError("row_ptr must hold M+1 offsets")
        best = np.zeros(M, np.int64)
        bq = np.zeros((M, 4))
        ns = np.zeros(M, np.int32)
        ni = np.zeros(M, np.int32)
        C = np.zeros(M)
        med = np.zeros(M, np.float32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        self._chk(self._lib.grs_compete_csr(self._h, _ptr(pos, _dp), M, rp.ctypes.data_as(i64p), _ptr(cq, _dp),
                                            cq.shape[0], int(nr), float(gridsize), best.ctypes.data_as(i64p),
                                            _ptr(bq, _dp), _ptr(ns, _ip), _ptr(ni, _ip), _ptr(C, _dp),
                                            _ptr(med, _fp)))
        return best, bq, ns, ni, C, med

    # -- orientation fit ----------------------------------------------------------
    def fit_or_local(self, pos, x0, half_width: float = 0.02, max_fun_evals: int = 2000, max_iter: int = 2000,
                     tol_fun: float = 1e-10, tol_x: float = 1e-4):
        """Batched fminsearchbnd of 1 - C over (q1, q2, q3) per voxel
        (fit_OR_constrain_with_local_comp_cuda.m:116-126).  pos, x0: (M, 3).
        Returns (x (M, 3), fval (M,), n_fun_evals (M,), n_iter (M,))."""
        pos = _d(np.asarray(pos).reshape(-1, 3))
        x0 = _d(np.asarray(x0).reshape(-1, 3))
        M = pos.shape[0]
        if x0.shape[0] != M:
            raise ValueError("pos and x0 must have the same number of rows")
        x = np.zeros((M, 3))
        fv = np.zeros(M)
        nf = np.zeros(M, np.int32)
        it = np.zeros(M, np.int32)
        self._chk(self._lib.grs_fit_or_local(self._h, _ptr(pos, _dp), _ptr(x0, _dp), M, float(half_width),
                                             int(max_fun_evals), int(max_iter), float(tol_fun), float(tol_x),
                                             _ptr(x, _dp), _ptr(fv, _dp), _ptr(nf, _ip), _ptr(it, _ip)))
        return x, fv, nf, it

    # -- seed indexing ------------------------------------------------------------
    def set_orientation_set(self, q) -> None:
        """Resident orientation set OR.q (N, 4) of the seed indexing."""
        q = _d(np.asarray(q).reshape(-1, 4))
        self._chk(self._lib.grs_set_orientation_set(self._h, _ptr(q, _dp), q.shape[0]))

    def index_seeds(self, pos, nr_seed: int, nr: int = 8, gridsize: float = 0.004, n_fit: int = 1,
                    half_width: float = 0.02, max_fun_evals: int = 2000, max_iter: int = 2000,
                    tol_fun: float = 1e-10, tol_x: float = 1e-4) -> dict:
        """GrainIndexing_engine_gpu_comp.m:46-66 with
        fit_OR_constrain_with_local_comp_cuda.m for every seed position
        (pos (P, 3) mm) against the resident orientation set.  Returns a dict:
        indexing (P, 18) [seed number, pos, Nr_simu, Nr_intersect, C,
        dis_median, Euler ZXZ deg, Rodrigues, quaternion], candidate (P,),
        x0 / x (P, 3), fval, n_fun_evals, n_iter (P,), seed_idx / seed_C
        (P, nr_seed)."""
        pos = _d(np.asarray(pos).reshape(-1, 3))
        P = pos.shape[0]
        ind = np.zeros((P, 18))
        cand = np.zeros(P, np.int32)
        info = np.zeros((P, 9))
        sidx = np.zeros((P, int(nr_seed)), np.int32)
        sC = np.zeros((P, int(nr_seed)))
        self._chk(self._lib.grs_index_seeds(self._h, _ptr(pos, _dp), P, int(nr_seed), int(nr), float(gridsize),
                                            int(n_fit), float(half_width), int(max_fun_evals), int(max_iter),
                                            float(tol_fun), float(tol_x), _ptr(ind, _dp), _ptr(cand, _ip),
                                            _ptr(info, _dp), _ptr(sidx, _ip), _ptr(sC, _dp)))
        return dict(indexing=ind, candidate=cand, x0=info[:, 0:3], x=info[:, 3:6], fval=info[:, 6],
                    n_fun_evals=info[:, 7].astype(np.int64), n_iter=info[:, 8].astype(np.int64), seed_idx=sidx,
                    seed_C=sC)

    # -- spot pairing ---------------------------------------------------------------
    def set_frame_spots(self, frames) -> None:
        """Experimental spots per projection: a list (one entry per rotation
        angle) of (n_i, 3) arrays [centroid y, centroid z, area]
        (Spots{i}(:,9:11))."""
        rows = [np.asarray(f, dtype=float).reshape(-1, 3) for f in frames]
        fp = np.zeros(len(rows) + 1, np.int64)
        fp[1:] = np.cumsum([r.shape[0] for r in rows])
        sp = _d(np.concatenate(rows) if fp[-1] else np.zeros((1, 3)))
        self._chk(self._lib.grs_set_frame_spots(self._h, fp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                len(rows), _ptr(sp, _dp), int(fp[-1])))

    def pair_spots(self, pos, U):
        """forward_comp_pair_spots.m for (pos (M, 3), U (M, 3, 3)).  Returns
        (ns, ni, med, [HittedSpots rows (n_m, 23) per pair])."""
        pos = _d(np.asarray(pos).reshape(-1, 3))
        U = _d(np.asarray(U).reshape(-1, 9))
        M = pos.shape[0]
        ns = np.zeros(M, np.int32)
        ni = np.zeros(M, np.int32)
        med = np.zeros(M, np.float32)
        rp = np.zeros(M + 1, np.int64)
        i64p = ctypes.POINTER(ctypes.c_int64)
        self._chk(self._lib.grs_pair_spots(self._h, _ptr(pos, _dp), _ptr(U, _dp), M, _ptr(ns, _ip), _ptr(ni, _ip),
                                           _ptr(med, _fp), rp.ctypes.data_as(i64p), None, 0))
        rows = np.zeros((max(int(rp[-1]), 1), 23))
        self._chk(self._lib.grs_pair_spots(self._h, _ptr(pos, _dp), _ptr(U, _dp), M, _ptr(ns, _ip), _ptr(ni, _ip),
                                           _ptr(med, _fp), rp.ctypes.data_as(i64p), _ptr(rows, _dp), rows.shape[0]))
        return ns, ni, med, [rows[rp[m]:rp[m + 1]].copy() for m in range(M)]

    # -- grain centre-of-mass fit -----------------------------------------------------
    def fit_grain_com(self, pair_rows, max_fun_evals: int = 2000, max_iter: int = 500, tol_fun: float = 1e-10,
                      tol_x: float = 1e-4):
        """fit_grain_COM.m for a list of per-grain HittedSpots row arrays
        ((n_m, 23), grs_pair_spots layout).  Returns (FitOutput (M, 6),
        fval (M,), n_fun_evals, n_iter)."""
        rows = [np.asarray(r, dtype=float).reshape(-1, 23) for r in pair_rows]
        M = len(rows)
        rp = np.zeros(M + 1, np.int64)
        rp[1:] = np.cumsum([r.shape[0] for r in rows])
        flat = _d(np.concatenate(rows) if M else np.zeros((1, 23)))
        x = np.zeros((M, 6))
        fv = np.zeros(M)
        nf = np.zeros(M, np.int32)
        it = np.zeros(M, np.int32)
        self._chk(self._lib.grs_fit_grain_com(self._h, _ptr(flat, _dp), rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              M, int(max_fun_evals), int(max_iter), float(tol_fun), float(tol_x),
                                              _ptr(x, _dp), _ptr(fv, _dp), _ptr(nf, _ip), _ptr(it, _ip)))
        return x, fv, nf, it

    # -- G-vector matching score ---------------------------------------------------
    def forward_gt(self, pos, q, fam_sq, d_possible, semantics: int = L.GT_MATCH) -> np.ndarray:
        """G-vector matching score of every orientation row q (N, 4) at pos
        against the spots of grs_set_frame_spots, (N, 4):
        GT_MATCH (calc_Gt_match, ForwardCalc_OR_space_Gt.m:117-205) gives
        [Nr_match, Gt_matched_cost, Nr_match/NrSpotExpectAll, mean angle];
        GT_MATCH_FUN (calc_Gt_match_fun.m) gives [Nr_match, median angle,
        Nr_match/NrSpotExpectAll, mean angle]."""
        pos = _d(np.asarray(pos).reshape(3))
        q = _d(np.asarray(q).reshape(-1, 4))
        fs = _d(np.asarray(fam_sq, dtype=float).reshape(-1))
        dp = _d(np.asarray(d_possible, dtype=float).reshape(-1))
        out = np.zeros((q.shape[0], 4))
        self._chk(self._lib.grs_forward_gt(self._h, _ptr(pos, _dp), _ptr(q, _dp), q.shape[0], _ptr(fs, _dp),
                                           _ptr(dp, _dp), int(min(fs.size, dp.size)), int(semantics),
                                           _ptr(out, _dp)))
        return out

    # -- region growth ------------------------------------------------------------
    VOLUMES = {"Completeness": 0, "Mask": 1, "Dismedian": 2, "Ninter": 3}

    def set_volume(self, name: str, v) -> None:
        """Resident volume V(x, y, z) (numpy array indexed [x-1, y-1, z-1])."""
        v = np.asarray(v, dtype=float)
        flat = np.ascontiguousarray(v.transpose(2, 1, 0)).reshape(-1)
        self._chk(self._lib.grs_set_volume(self._h, self.VOLUMES[name], _ptr(flat, _dp), v.shape[0], v.shape[1],
                                           v.shape[2]))
        self._vshape = v.shape

    def _volume_shape(self):
        if self._vshape is None:
            raise L.GrsError(L.GRS_ERR_STATE, "no volume set")
        return self._vshape

    def get_volume(self, name: str) -> np.ndarray:
        nx, ny, nz = self._volume_shape()
        flat = np.zeros(nx * ny * nz)
        self._chk(self._lib.grs_get_volume(self._h, self.VOLUMES[name], _ptr(flat, _dp), flat.size))
        return flat.reshape(nz, ny, nx).transpose(2, 1, 0)  # column-major view of the fresh buffer

    def _linear(self, vox) -> np.ndarray:
        """Linear indices (x fastest, 0-based) of (n, 3) 1-based [x y z] rows;
        a row with a component outside its axis maps to -1, which the C ABI
        rejects (it could otherwise land inside the volume)."""
        nx, ny, nz = self._volume_shape()
        v = np.asarray(vox, dtype=np.int64).reshape(-1, 3) - 1
        lin = np.ascontiguousarray((v[:, 0] + nx * (v[:, 1] + ny * v[:, 2])).astype(np.int64))
        lin[np.any((v < 0) | (v >= np.array([nx, ny, nz])), axis=1)] = -1
        return lin

    def set_volume_at(self, name: str, vox, values) -> None:
        """V(x, y, z) = values at the (n, 3) 1-based [x y z] rows of vox (the
        rows distinct; an index outside the volume or repeated raises
        GrsError)."""
        lin = self._linear(vox)
        val = _d(np.asarray(values, dtype=float).reshape(-1))
        if val.size != lin.size:
            raise ValueError("one value per voxel row")
        self._chk(self._lib.grs_set_volume_at(self._h, self.VOLUMES[name],
                                              lin.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lin.size,
                                              _ptr(val, _dp)))

    def get_volume_at(self, name: str, vox) -> np.ndarray:
        """V(x, y, z) at the (n, 3) 1-based [x y z] rows of vox, gathered on
        the device (an index outside the volume raises GrsError)."""
        lin = self._linear(vox)
        out = np.zeros(lin.size)
        self._chk(self._lib.grs_get_volume_at(self._h, self.VOLUMES[name],
                                              lin.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lin.size,
                                              _ptr(out, _dp)))
        return out

    def grow_region(self, seed, indexed_comp: float, U, spot_L: float, drop_off: float, maxD: float,
                    maxDmedian: float, p1) -> dict:
        """grow_indexed_region_cuda.m on the resident volumes (seed 1-based
        [x y z]).  Returns center, replace_center, indexed (n, 3) 1-based in
        find() order, n_checked."""
        sd = np.asarray(seed, dtype=np.int32).reshape(3)
        U = _d(np.asarray(U).reshape(9))
        p1 = _d(np.asarray(p1).reshape(-1)[:11])
        center = np.zeros(3)
        rc = ctypes.c_int32()
        n = ctypes.c_int64()
        nchk = ctypes.c_int64()
        cap = int(np.prod(self._volume_shape()))
        idx = np.zeros((cap, 3), np.int32)
        self._chk(self._lib.grs_grow_region(self._h, _ptr(sd, _ip), float(indexed_comp), _ptr(U, _dp), float(spot_L),
                                            float(drop_off), float(maxD), float(maxDmedian), _ptr(p1, _dp),
                                            _ptr(center, _dp), ctypes.byref(rc), ctypes.byref(n), _ptr(idx, _ip),
                                            cap, ctypes.byref(nchk)))
        return dict(center=center, replace_center=int(rc.value), indexed=idx[:n.value].copy(),
                    n_checked=int(nchk.value))

    # -- candidate generation ---------------------------------------------------
    def local_sampling(self, centres, nr: int, gridsize: float = 0.004) -> list:
        """ori_local_sampling.m:12-34 for each centre row [q0 q1 q2 q3] (a
        3-vector gets q0 = sqrt(1-(q1^2+q2^2+q3^2)), :12-14).  Returns one
        (n, 4) array of sorted unique quaternion rows (OR_local.q) per centre."""
        c = np.asarray(centres, dtype=float)
        c = c.reshape(1, -1) if c.ndim == 1 else c
        if c.shape[1] == 3:
            q0 = np.sqrt(1.0 - ((c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1]) + c[:, 2] * c[:, 2]))
            c = np.column_stack([q0, c])
        c = _d(c.reshape(-1, 4))
        K = c.shape[0]
        cap = 1 + (int(nr) + 3) ** 3
        out = np.zeros((K, cap, 4))
        n = np.zeros(K, np.int64)
        self._chk(self._lib.grs_local_sampling(self._h, _ptr(c, _dp), K, int(nr), float(gridsize), _ptr(out, _dp), cap,
                                               n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return [out[k, :n[k]].copy() for k in range(K)]

    def find_neighbor_ors(self, grain_id: np.ndarray, vox, search_radius: int = 20, nr_max: int = 30):
        """find_neighbor_ORs.m:3-96.  grain_id: (nx, ny, nz) indexed like
        DS_in.GrainId(x, y, z); vox: (M, >=3) 1-based [x y z ...].  Returns
        (id_neigb_all (K,), id_neigb_ind (M, 2) 1-based [first last])."""
        g = np.asarray(grain_id)
        nx, ny, nz = g.shape
        gf = np.ascontiguousarray(g.reshape(-1, order="F"), dtype=np.int32)
        v = np.ascontiguousarray(np.asarray(vox)[:, :3], dtype=np.int32)
        M = v.shape[0]
        ids = np.zeros(max(M * nr_max, 1), np.int32)
        ind = np.zeros((M, 2), np.int64)
        n = ctypes.c_int64()
        i32 = ctypes.POINTER(ctypes.c_int32)
        i64 = ctypes.POINTER(ctypes.c_int64)
        self._chk(self._lib.grs_find_neighbor_ors(self._h, gf.ctypes.data_as(i32), nx, ny, nz, v.ctypes.data_as(i32), M,
                                                  int(search_radius), int(nr_max), ids.ctypes.data_as(i32),
                                                  ind.ctypes.data_as(i64), ctypes.byref(n)))
        return ids[:n.value].copy(), ind

    # -- input preparation ------------------------------------------------------
    def render_spots(self, pos, grain, U, clear: bool = True) -> None:
        """Forward-simulates spot masks: voxel i at pos[i] (mm) with orientation U[grain[i]]."""
        pos = _d(np.asarray(pos).reshape(-1, 3))
        grain = np.ascontiguousarray(np.asarray(grain).reshape(-1), dtype=np.int32)
        U = _d(np.asarray(U).reshape(-1, 9))
        self._chk(self._lib.grs_render_spots(self._h, _ptr(pos, _dp), _ptr(grain, _ip), pos.shape[0],
                                             _ptr(U, _dp), U.shape[0], 1 if clear else 0))

    def set_binary_stack(self, mask: np.ndarray) -> None:
        """mask: (nz, ny, nrot) boolean/uint8 indexed like proj_bin(detz, dety, i)."""
        m = np.asarray(mask).astype(np.uint8)
        nz, ny, nr = m.shape
        flat = np.ascontiguousarray(m.reshape(-1, order="F"))
        self._chk(self._lib.grs_set_binary_stack(self._h, flat.ctypes.data, nz, ny, nr, 0))

    def distance_from_mask(self) -> None:
        self._chk(self._lib.grs_distance_from_mask(self._h))

    def get_mask(self, shape) -> np.ndarray:
        nz, ny, nr = shape
        out = np.zeros(nz * ny * nr, np.uint8)
        self._chk(self._lib.grs_get_mask(self._h, out.ctypes.data, out.size))
        return out.reshape((nz, ny, nr), order="F")

    def get_distance_stack(self, shape) -> np.ndarray:
        nz, ny, nr = shape
        out = np.zeros(nz * ny * nr, np.float32)
        self._chk(self._lib.grs_get_distance_stack(self._h, _ptr(out, _fp), out.size))
        return out.reshape((nz, ny, nr), order="F")

    # -- device buffers (HBM-resident inputs/outputs) ---------------------------
    def device_alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self._chk(self._lib.grs_device_alloc(self._h, int(nbytes), ctypes.byref(p)))
        return int(p.value)

    def device_free(self, dptr: int) -> None:
        self._chk(self._lib.grs_device_free(self._h, ctypes.c_void_p(dptr)))

    def to_device(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        d = self.device_alloc(arr.nbytes)
        self._chk(self._lib.grs_copy_to_device(self._h, ctypes.c_void_p(d), arr.ctypes.data, arr.nbytes))
        return d

    def device_copy(self, d_dst: int, d_src: int, nbytes: int) -> None:
        self._chk(self._lib.grs_device_copy(self._h, ctypes.c_void_p(d_dst), ctypes.c_void_p(d_src), int(nbytes)))

    def to_host(self, dptr: int, shape, dtype) -> np.ndarray:
        out = np.zeros(shape, dtype)
        self._chk(self._lib.grs_copy_to_host(self._h, out.ctypes.data, ctypes.c_void_p(dptr), out.nbytes))
        return out

    # -- measurement ----------------------------------------------------------
    def set_timing(self, on: bool) -> None:
        self._chk(self._lib.grs_set_timing(self._h, 1 if on else 0))

    def kernel_time(self, reset: bool = False):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self._lib.grs_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n), 1 if reset else 0))
        return ms.value, n.value

    def set_autotune(self, on: bool = True) -> None:
        """Launch autotuning of the scoring kernel (waves per block, hit-test source)."""
        self._chk(self._lib.grs_set_autotune(self._h, 1 if on else 0))

    MASK_NAMES = {L.MASK_BITS: "bits", L.MASK_KEYS: "keys", L.MASK_TILED: "bits_tiled32"}

    def tuning(self, kind: int):
        """Chosen configuration of kind 0 (with median), 1 (counts-only CSR) or
        2 (counts-only grid, seed indexing): dict with waves_per_block,
        mask_layout ("bits", "keys", "bits_tiled32"), xcd_map and ms_per_eval
        of every candidate; None before the first large enough call."""
        w, m, x, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = (ctypes.c_double * L.TUNE_MAX)()
        st = self._lib.grs_get_tuning(self._h, int(kind), ctypes.byref(w), ctypes.byref(m), ctypes.byref(x),
                                      ctypes.byref(n), ms)
        if st != 0:
            return None
        return {"waves_per_block": w.value, "mask_layout": self.MASK_NAMES[m.value], "xcd_map": bool(x.value),
                "ms_per_eval": list(ms)[:n.value]}

    def set_launch_config(self, kind: int, waves_per_block: int, mask_layout: int = L.MASK_BITS,
                          xcd_map: bool = False) -> None:
        """Pins kind's launch configuration (grs_set_launch_config);
        waves_per_block = 0 forgets it."""
        self._chk(self._lib.grs_set_launch_config(self._h, int(kind), int(waves_per_block), int(mask_layout),
                                                  int(bool(xcd_map))))

    def set_graphs(self, on: bool = True) -> None:
        """HIP-graph capture and replay of the batched compete and seed-indexing
        pipelines."""
        self._chk(self._lib.grs_set_graphs(self._h, 1 if on else 0))

    def graph_stats(self):
        """(captures, replays) of the batched-compete graphs so far."""
        cap, rep = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._lib.grs_graph_stats(self._h, ctypes.byref(cap), ctypes.byref(rep)))
        return cap.value, rep.value

    def set_audit(self, on: bool = True) -> None:
        """Arms (and clears) the audit of the fast kernel's resolution stage."""
        self._chk(self._lib.grs_set_audit(self._h, 1 if on else 0))

    def audit(self) -> dict:
        """Audit counters since set_audit(True): events examined by the
        resolution stage, resolved selected / not selected, sent to the exact
        pipeline, disagreements, events that entered the spot construction,
        and the largest coordinate difference (px)."""
        counts = (ctypes.c_int64 * 6)()
        dev = ctypes.c_double()
        self._chk(self._lib.grs_get_audit(self._h, counts, ctypes.byref(dev)))
        return {"examined": counts[0], "resolved_selected": counts[1], "resolved_rejected": counts[2],
                "exact": counts[3], "mismatched": counts[4], "in_window": counts[5], "max_dev_px": dev.value}

    def math_device(self, fn: int, x) -> np.ndarray:
        x = _d(np.asarray(x).reshape(-1))
        y = np.zeros_like(x)
        self._chk(self._lib.grs_math_device(self._h, int(fn), _ptr(x, _dp), _ptr(y, _dp), x.shape[0]))
        return y

    def synchronize(self) -> None:
        self._chk(self._lib.grs_synchronize(self._h))


# ---- the MEX operators ------------------------------------------------------

def _legacy_check(code):
    L.check(code, None)


def cuda_forward_comp(pos, q, S, B, hkl, proj, rot_angles, RotDet, param) -> np.ndarray:
    """cuda_forward_comp(pos, q, S, B, hkl, proj, rot, RotDet, param): single
    row vectors as built by ForwardCalc_OR_space_comp_cuda.m:17-27; returns the
    1 x 4N single vector [Nr_simu Nr_intersect C dis_median] per orientation."""
    pos, q, S, B, hkl, proj, rot, R, param = (_f(x) for x in (pos, q, S, B, hkl, proj, rot_angles, RotDet, param))
    out = np.zeros(q.size, np.float32)
    _legacy_check(L.lib().grs_cuda_forward_comp(
        _ptr(pos, _fp), _ptr(q, _fp), q.size, _ptr(S, _fp), _ptr(B, _fp), _ptr(hkl, _fp), hkl.size,
        _ptr(proj, _fp), proj.size, _ptr(rot, _fp), rot.size, _ptr(R, _fp), _ptr(param, _fp), param.size,
        _ptr(out, _fp)))
    return out


def cuda_grow_indexed_region(vox, U, S, B, hkl, proj, rot_angles, RotDet, param, param1) -> np.ndarray:
    """cuda_grow_indexed_region (grow_indexed_region_cuda.m:83-107): returns the
    1 x 7M single vector [x y z Nr_simu Nr_intersect dis_median growth_flag]."""
    vox, U, S, B, hkl, proj, rot, R, param, p1 = (_f(x) for x in (vox, U, S, B, hkl, proj, rot_angles, RotDet,
                                                                     param, param1))
    out = np.zeros(vox.size // 6 * 7, np.float32)
    _legacy_check(L.lib().grs_cuda_grow_indexed_region(
        _ptr(vox, _fp), vox.size, _ptr(U, _fp), _ptr(S, _fp), _ptr(B, _fp), _ptr(hkl, _fp), hkl.size,
        _ptr(proj, _fp), proj.size, _ptr(rot, _fp), rot.size, _ptr(R, _fp), _ptr(param, _fp), param.size,
        _ptr(p1, _fp), p1.size, _ptr(out, _fp)))
    return out


def cuda_forward_comp_compete(vox, id_all, Euler, S, B, hkl, proj, rot_angles, RotDet, param,
                              param1) -> np.ndarray:
    """cuda_forward_comp_compete (grow_unindexed_compete_comp_cuda.m:17-41):
    returns the 1 x 8M single vector [i phi1 PHI phi2 GrainId C Nr_intersect dis_median]."""
    vox, ids, eu, S, B, hkl, proj, rot, R, param, p1 = (
        _f(x) for x in (vox, id_all, Euler, S, B, hkl, proj, rot_angles, RotDet, param, param1))
    out = np.zeros(vox.size // 5 * 8, np.float32)
    _legacy_check(L.lib().grs_cuda_forward_comp_compete(
        _ptr(vox, _fp), vox.size, _ptr(ids, _fp), ids.size, _ptr(eu, _fp), eu.size, _ptr(S, _fp), _ptr(B, _fp),
        _ptr(hkl, _fp), hkl.size, _ptr(proj, _fp), proj.size, _ptr(rot, _fp), rot.size, _ptr(R, _fp),
        _ptr(param, _fp), param.size, _ptr(p1, _fp), p1.size, _ptr(out, _fp)))
    return out
