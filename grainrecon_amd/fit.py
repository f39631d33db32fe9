"""Differentiable geometry and grain refinement on the engine's adjoint.

Replaces the derivative-free fminsearchbnd fits of GrainRecon
(L_shift_tilt_fitting.m:195-222 over geo_fitting.m, fit_grain_COM.m:5-29)
with the unrounded spot-centroid residual and its hand-written adjoint
(grs_spot_residual, include/grscore.h).

Unknowns: 7 geometry values [Lsam2sou Lsam2det dety00 detz00 tilt_x tilt_y
tilt_z] and per grain [r1 r2 r3 x y z] (Rodrigues vector, centre of mass).
The normal equations have a block-arrow structure -- every spot touches the
7 shared unknowns and the 6 of its own grain -- so a Levenberg-Marquardt
step eliminates the 6x6 grain blocks (Schur complement) and solves a 7x7
system; one step costs one Jacobian launch plus O(G) small solves.
"""
from __future__ import annotations

import ctypes
import dataclasses
import time

import numpy as np

from . import _lib as L
from .engine import Engine, _d, _ptr

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class SpotFit:
    """A matched-spot set resident on the engine (grs_set_spots)."""

    def __init__(self, engine: Engine, hkl, omega, meas, grain, n_grains: int):
        self.eng = engine
        order = np.argsort(np.asarray(grain), kind="stable")
        self.order = order
        self.hkl = _d(np.asarray(hkl, float).reshape(-1, 3)[order])
        self.omega = _d(np.asarray(omega, float).reshape(-1)[order])
        self.meas = _d(np.asarray(meas, float).reshape(-1, 2)[order])
        self.grain = np.ascontiguousarray(np.asarray(grain).reshape(-1)[order], dtype=np.int32)
        self.G = int(n_grains)
        self.n = self.grain.size
        self.goff = np.searchsorted(self.grain, np.arange(self.G + 1))
        engine._chk(engine._lib.grs_set_spots(engine._h, self.n, _ptr(self.hkl, _dp), _ptr(self.omega, _dp),
                                              _ptr(self.meas, _dp), _ptr(self.grain, _ip), self.G))

    def _call(self, g7, gr6, upstream=None, grad=False, jac=False):
        g7 = _d(np.asarray(g7, float).reshape(7))
        gr6 = _d(np.asarray(gr6, float).reshape(self.G, 6))
        r = np.zeros((self.n, 2))
        up = None if upstream is None else _d(np.asarray(upstream, float).reshape(self.n, 2))
        d7 = np.zeros(7) if grad else None
        d6 = np.zeros((self.G, 6)) if grad else None
        J = np.zeros((self.n, 2, 13)) if jac else None
        nul = ctypes.cast(None, _dp)
        self.eng._chk(self.eng._lib.grs_spot_residual(
            self.eng._h, _ptr(g7, _dp), _ptr(gr6, _dp), _ptr(r, _dp), nul if up is None else _ptr(up, _dp),
            nul if d7 is None else _ptr(d7, _dp), nul if d6 is None else _ptr(d6, _dp),
            nul if J is None else _ptr(J, _dp)))
        return r, d7, d6, J

    def residual(self, g7, gr6) -> np.ndarray:
        return self._call(g7, gr6)[0]

    def value_and_grad(self, g7, gr6, upstream=None):
        """Residual and J^T u (u = upstream, or the residual: grad of 0.5|r|^2)."""
        r, d7, d6, _ = self._call(g7, gr6, upstream=upstream, grad=True)
        return r, d7, d6

    def jacobian(self, g7, gr6):
        r, _, _, J = self._call(g7, gr6, jac=True)
        return r, J

    def normal(self, g7, gr6, delta: float = 0.0) -> dict:
        """Normal equations of the (weighted) residual reduced on the device
        (grs_spot_normal): Agg (7, 7), bg (7,), Akk (G, 6, 6), Agk (G, 7, 6),
        bk (G, 6), cost (sum rho) and sumsq (sum |r|^2)."""
        g7 = _d(np.asarray(g7, float).reshape(7))
        gr6 = _d(np.asarray(gr6, float).reshape(self.G, 6))
        Agg, bg, sums = np.zeros((7, 7)), np.zeros(7), np.zeros(2)
        Akk, Agk, bk = np.zeros((self.G, 6, 6)), np.zeros((self.G, 7, 6)), np.zeros((self.G, 6))
        self.eng._chk(self.eng._lib.grs_spot_normal(
            self.eng._h, _ptr(g7, _dp), _ptr(gr6, _dp), float(delta), _ptr(Agg, _dp), _ptr(bg, _dp),
            _ptr(Akk, _dp), _ptr(Agk, _dp), _ptr(bk, _dp), _ptr(sums, _dp)))
        return dict(Agg=Agg, bg=bg, Akk=Akk, Agk=Agk, bk=bk, cost=float(sums[0]), sumsq=float(sums[1]))


def rms_px(r: np.ndarray) -> float:
    """Root mean square Euclidean distance (pixels) of residual pairs."""
    return float(np.sqrt(np.mean(np.sum(r * r, axis=1))))


def mean_rounded_px(r_unrounded_sim: np.ndarray, meas: np.ndarray) -> float:
    """The reference objective (geo_fitting.m:70-78): mean Euclidean distance
    between the rounded simulated pixel and the measured centroid."""
    sim = r_unrounded_sim
    rounded = np.sign(sim) * np.floor(np.abs(sim) + 0.5)
    return float(np.mean(np.sqrt(np.sum((rounded - meas) ** 2, axis=1))))


@dataclasses.dataclass
class FitResult:
    g7: np.ndarray
    gr6: np.ndarray
    rms: float
    iterations: int
    seconds: float
    history: list


def host_normal(fit, g7, gr6, delta: float = 0.0) -> dict:
    """The normal equations of grs_spot_normal computed on the host from the
    Jacobian rows (fit.jacobian): for fit objects without a device reduction
    and as its check."""
    r, J = fit.jacobian(g7, gr6)
    G, goff = fit.G, fit.goff
    s2 = np.sum(r * r, axis=1)
    if delta > 0.0:
        h = np.sqrt(s2 + delta * delta)
        cost = float(np.sum(h - delta))
        sw = 1.0 / np.sqrt(h)
        r = r * sw[:, None]
        J = J * sw[:, None, None]
    else:
        cost = 0.5 * float(np.sum(s2))
    Jg = J[:, :, :7].reshape(-1, 7)
    Jk = J[:, :, 7:]
    Agg = Jg.T @ Jg
    bg = Jg.T @ r.reshape(-1)
    Akk = np.zeros((G, 6, 6))
    Agk = np.zeros((G, 7, 6))
    bk = np.zeros((G, 6))
    nz = goff[:-1] < goff[1:]
    if fit.n:
        Akk[nz] = np.add.reduceat(np.einsum("nai,naj->nij", Jk, Jk), goff[:-1][nz], axis=0)
        Agk[nz] = np.add.reduceat(np.einsum("nai,naj->nij", J[:, :, :7], Jk), goff[:-1][nz], axis=0)
        bk[nz] = np.add.reduceat(np.einsum("nai,na->ni", Jk, r), goff[:-1][nz], axis=0)
    return dict(Agg=Agg, bg=bg, Akk=Akk, Agk=Agk, bk=bk, cost=cost, sumsq=float(np.sum(s2)))


def levenberg_marquardt(fit: SpotFit, g7, gr6, max_iter: int = 50, tol: float = 1e-10, lam: float = 1e-3,
                        fix_geometry: bool = False, fix_grains: bool = False,
                        fix_position: bool = False, delta: float = 0.0) -> FitResult:
    """Joint refinement of geometry and grains (block-arrow LM); either block
    can be held fixed (fix_grains=True is the reference's geometry-only fit,
    L_shift_tilt_fitting.m:23-50).  fix_position holds the grain centres and
    fits the orientations only (the free parameters of
    fit_OR_constrain_with_local_comp_cuda.m:117-126).  delta > 0 minimises the
    Charbonnier sum of sqrt(|r|^2 + delta^2) - delta (the mean Euclidean
    distance the reference's objectives average, smoothed at delta pixels;
    robust to mispaired spots) by iteratively reweighted Gauss-Newton steps;
    delta = 0 is least squares.  The normal equations come from the device
    (SpotFit.normal, grs_spot_normal) when the fit object has them."""
    t0 = time.perf_counter()
    g7 = np.array(g7, float)
    gr6 = np.array(gr6, float).reshape(fit.G, 6)
    normal = getattr(fit, "normal", None)

    def ne_at(a7, a6):
        return normal(a7, a6, delta) if normal is not None else host_normal(fit, a7, a6, delta)

    def rms(ne):
        return float(np.sqrt(ne["sumsq"] / max(fit.n, 1)))

    ne = ne_at(g7, gr6)
    hist = [rms(ne)]
    it = 0
    eye6 = np.eye(6)[None]
    for it in range(1, max_iter + 1):
        Agg, bg = ne["Agg"], ne["bg"]
        Akk, Agk, bk = ne["Akk"].copy(), ne["Agk"].copy(), ne["bk"].copy()
        if fix_position:
            # decouple the centre block: its step solves I dk = 0
            Akk[:, 3:, :] = 0.0
            Akk[:, :, 3:] = 0.0
            Akk[:, 3:, 3:] = np.eye(3)
            Agk[:, :, 3:] = 0.0
            bk[:, 3:] = 0.0
        while True:
            Dg = np.diag(np.diag(Agg)) * lam + np.eye(7) * 1e-30
            Ak = Akk + lam * np.einsum("gii->gi", Akk)[:, :, None] * eye6 + 1e-30 * eye6
            if fix_geometry:
                dg = np.zeros(7)
            elif fix_grains:
                dg = np.linalg.solve(Agg + Dg, -bg)
            else:
                Ak_inv = np.linalg.inv(Ak)
                Sg = Agg + Dg - np.einsum("gij,gjk,glk->il", Agk, Ak_inv, Agk)
                rhs = -(bg - np.einsum("gij,gjk,gk->i", Agk, Ak_inv, bk))
                dg = np.linalg.solve(Sg, rhs)
            if fix_grains:
                dk = np.zeros_like(gr6)
            else:
                dk = -np.linalg.solve(Ak, (bk + np.einsum("gji,j->gi", Agk, dg))[:, :, None])[:, :, 0]
            g7n, gr6n = g7 + dg, gr6 + dk
            nen = ne_at(g7n, gr6n)
            if np.isfinite(nen["cost"]) and nen["cost"] <= ne["cost"]:
                lam = max(lam / 10.0, 1e-12)
                break
            lam *= 10.0
            if lam > 1e12:
                break
        if not (np.isfinite(nen["cost"]) and nen["cost"] <= ne["cost"]):
            break
        rel = (ne["cost"] - nen["cost"]) / max(ne["cost"], 1e-300)
        g7, gr6, ne = g7n, gr6n, nen
        hist.append(rms(ne))
        if rel < tol:
            break
    return FitResult(g7=g7, gr6=gr6, rms=rms(ne), iterations=it, seconds=time.perf_counter() - t0, history=hist)


def torch_residual():
    """A torch.autograd.Function over the engine: forward = residual, backward
    = the hand-written adjoint (vector-Jacobian product)."""
    import torch

    class SpotResidual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, g7, gr6, fit):
            ctx.fit = fit
            ctx.save_for_backward(g7, gr6)
            r = fit.residual(g7.detach().cpu().double().numpy(), gr6.detach().cpu().double().numpy())
            return torch.from_numpy(r).to(g7.device)

        @staticmethod
        def backward(ctx, grad_r):
            g7, gr6 = ctx.saved_tensors
            _, d7, d6 = ctx.fit.value_and_grad(g7.detach().cpu().double().numpy(), gr6.detach().cpu().double().numpy(),
                                               upstream=grad_r.detach().cpu().double().numpy())
            return torch.from_numpy(d7).to(g7.device), torch.from_numpy(d6).to(gr6.device), None

    return SpotResidual


def simulate_matched_spots(eng: Engine, g7, com, U, rng=None, noise_px: float = 0.0):
    """Spot list of each grain (voxel at its centre of mass, orientation U)
    with the engine's geometry: (hkl, omega, measured centroid, grain).
    Noise is Gaussian, per axis, in pixels."""
    G = com.shape[0]
    pos = _d(com)
    grain = np.arange(G, dtype=np.int32)
    Uf = _d(np.asarray(U).reshape(G, 9))
    cap = G * eng.nrot * eng.H
    vox = np.zeros(cap, np.int32)
    om = np.zeros(cap, np.int32)
    h = np.zeros(cap, np.int32)
    yz = np.zeros((cap, 2))
    n = ctypes.c_int64()
    eng._chk(eng._lib.grs_simulate_spots(eng._h, _ptr(pos, _dp), _ptr(grain, _ip), G, _ptr(Uf, _dp), G, cap,
                                         _ptr(vox, _ip), _ptr(om, _ip), _ptr(h, _ip), _ptr(yz, _dp),
                                         ctypes.byref(n)))
    k = n.value
    meas = yz[:k].copy()
    if noise_px and rng is not None:
        meas += rng.normal(scale=noise_px, size=meas.shape)
    return dict(grain=vox[:k], omega_idx=om[:k], hkl_idx=h[:k], meas=meas, clean=yz[:k].copy())
