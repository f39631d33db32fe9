"""Bounded Nelder-Mead with fminsearchbnd semantics -- CPU BASELINE ONLY.

The reference fits geometry and grains with fminsearchbnd
(L_shift_tilt_fitting.m:195-222, fit_grain_COM.m:25-28): MATLAB's
fminsearch (the Nelder-Mead variant of J. C. Lagarias, J. A. Reeds, M. H.
Wright, P. E. Wright, "Convergence properties of the Nelder-Mead simplex
method in low dimensions", SIAM J. Optim. 9 (1998)) over variables mapped
through a sine transform for two-sided bounds (fminsearchbnd.m:38-41).  This
is this project's restatement, used only to time the derivative-free baseline
against the adjoint-based fit.

Defaults follow fminsearch: reflection 1, expansion 2, contraction 1/2,
shrink 1/2, initial simplex x0 with each coordinate scaled by 1.05 (0.00025
for zeros), stop when the simplex diameter <= TolX and the spread of f <=
TolFun, or after MaxFunEvals / MaxIter.
"""
from __future__ import annotations

import math

import numpy as np


def _to_bounded(u, lb, ub):
    # fminsearchbnd xtransform for two-sided bounds
    x = (np.sin(u) + 1.0) / 2.0 * (ub - lb) + lb
    return np.clip(x, lb, ub)


def _from_bounded(x, lb, ub):
    z = 2.0 * (x - lb) / (ub - lb) - 1.0
    return 2.0 * math.pi + np.arcsin(np.clip(z, -1.0, 1.0))


def fminsearchbnd(fun, x0, lb, ub, max_fun_evals=2000, max_iter=500, tol_fun=1e-10, tol_x=1e-4):
    """Returns (x, fval, n_fun_evals, n_iter)."""
    x0 = np.asarray(x0, float)
    lb = np.asarray(lb, float)
    ub = np.asarray(ub, float)
    n = x0.size
    f = lambda u: fun(_to_bounded(u, lb, ub))  # noqa: E731
    u0 = _from_bounded(x0, lb, ub)
    sim = np.empty((n + 1, n))
    sim[0] = u0
    for k in range(n):
        y = u0.copy()
        y[k] = y[k] * 1.05 if y[k] != 0.0 else 0.00025
        sim[k + 1] = y
    fv = np.array([f(v) for v in sim])
    nf = n + 1
    it = 0
    while nf < max_fun_evals and it < max_iter:
        order = np.argsort(fv, kind="stable")
        sim, fv = sim[order], fv[order]
        if np.max(np.abs(fv[0] - fv[1:])) <= tol_fun and np.max(np.abs(sim[1:] - sim[0])) <= tol_x:
            break
        it += 1
        xbar = sim[:-1].mean(axis=0)
        xr = 2.0 * xbar - sim[-1]
        fr = f(xr)
        nf += 1
        if fr < fv[0]:
            xe = 3.0 * xbar - 2.0 * sim[-1]
            fe = f(xe)
            nf += 1
            if fe < fr:
                sim[-1], fv[-1] = xe, fe
            else:
                sim[-1], fv[-1] = xr, fr
        elif fr < fv[-2]:
            sim[-1], fv[-1] = xr, fr
        else:
            shrink = False
            if fr < fv[-1]:
                xc = 1.5 * xbar - 0.5 * sim[-1]
                fc = f(xc)
                nf += 1
                if fc <= fr:
                    sim[-1], fv[-1] = xc, fc
                else:
                    shrink = True
            else:
                xcc = 0.5 * xbar + 0.5 * sim[-1]
                fcc = f(xcc)
                nf += 1
                if fcc < fv[-1]:
                    sim[-1], fv[-1] = xcc, fcc
                else:
                    shrink = True
            if shrink:
                for k in range(1, n + 1):
                    sim[k] = sim[0] + 0.5 * (sim[k] - sim[0])
                    fv[k] = f(sim[k])
                    nf += 1
    k = int(np.argmin(fv))
    return _to_bounded(sim[k], lb, ub), float(fv[k]), nf, it
