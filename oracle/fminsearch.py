"""Bounded Nelder-Mead with fminsearchbnd semantics -- TEST INFRASTRUCTURE
(and the CPU baseline of the geometry fit) ONLY.

The reference fits geometry, grains and orientations with fminsearchbnd
(L_shift_tilt_fitting.m:195-222, fit_grain_COM.m:25-28,
fit_OR_constrain_with_local_comp_cuda.m:117-126): MATLAB's fminsearch (the
Nelder-Mead method as analysed by J. C. Lagarias, J. A. Reeds, M. H. Wright,
P. E. Wright, SIAM J. Optim. 9 (1998) 112-147) over variables mapped through
a sine for two-sided bounds (fminsearchbnd.m:152-165 for the start point,
:263-268 for the transform).

fminsearch's published behaviour, restated step by step:
  * simplex: x0 and x0 with coordinate j scaled by 1.05 (0.00025 when 0),
    sorted by f (stable; NaN last); iteration count 1, n+1 evaluations;
  * loop while evaluations < MaxFunEvals and iterations < MaxIter; stop when
    max|f1 - fj| <= max(TolFun, 10 eps(f1)) and
    max|v_j - v_1| (all coordinates) <= max(TolX, 10 eps(max(v_1)));
  * xbar = (sum of the n best vertices) / n; reflection 2 xbar - v_end;
    expansion 3 xbar - 2 v_end when the reflection beats the best (kept if
    it beats the reflection); the reflection is kept if it beats the second
    worst; otherwise an outside contraction 1.5 xbar - 0.5 v_end (kept if
    <= the reflection) when the reflection beats the worst, else an inside
    contraction 0.5 xbar + 0.5 v_end (kept if it beats the worst); a failed
    contraction shrinks every vertex towards the best one by 1/2;
  * after every iteration the simplex is sorted again;
  * max() ignores NaN as MATLAB's does.
Sums run in index order.  sin and asin are injectable so that the checker
of the engine's batched fit can use the engine's own FP64 functions.
"""
from __future__ import annotations

import math

import numpy as np

_EPS = 2.220446049250313e-16


def _eps(x: float) -> float:
    """MATLAB eps(x): the spacing of doubles at |x|."""
    if x != x:
        return float("nan")
    x = abs(x)
    if x == math.inf:
        return float("nan")
    if x < 2.2250738585072014e-308:
        return 5e-324
    m, e = math.frexp(x)
    return math.ldexp(1.0, e - 53)


def _nanmax(vals):
    """MATLAB max: NaN entries are ignored unless every entry is NaN."""
    vals = [v for v in vals if v == v]
    return max(vals) if vals else float("nan")


def _sort(fv, sim):
    order = sorted(range(len(fv)), key=lambda i: (fv[i] != fv[i], fv[i] if fv[i] == fv[i] else 0.0))
    return [fv[i] for i in order], [sim[i] for i in order]


def initial_simplex(u0):
    """fminsearch's start: u0 and, for every coordinate j, u0 with u0_j
    scaled by 1 + 0.05 (usual_delta), or set to 0.00025 (zero_term_delta)
    when it is 0."""
    u0 = [float(v) for v in u0]
    sim = [list(u0)]
    for j in range(len(u0)):
        y = list(u0)
        y[j] = (1.0 + 0.05) * y[j] if y[j] != 0.0 else 0.00025
        sim.append(y)
    return sim


def bnd_start(x0, lb, ub, asin=math.asin):
    """fminsearchbnd.m:152-165 (two-sided bounds): the transformed start point."""
    u0 = []
    for x, lo, hi in zip(x0, lb, ub):
        if x <= lo:
            u0.append(-math.pi / 2)
        elif x >= hi:
            u0.append(math.pi / 2)
        else:
            z = 2.0 * (x - lo) / (hi - lo) - 1.0
            u0.append(2.0 * math.pi + asin(max(-1.0, min(1.0, z))))
    return u0


def bnd_xform(u, lb, ub, sin=math.sin):
    """fminsearchbnd.m:263-268: u -> x for two-sided bounds."""
    out = []
    for k in range(len(u)):
        t = (sin(u[k]) + 1.0) / 2.0
        t = t * (ub[k] - lb[k]) + lb[k]
        out.append(max(lb[k], min(ub[k], t)))
    return np.array(out)


def fminsearchbnd(fun, x0, lb, ub, max_fun_evals=2000, max_iter=500, tol_fun=1e-10, tol_x=1e-4,
                  sin=math.sin, asin=math.asin):
    """Returns (x, fval, n_fun_evals, n_iter); every bound must be finite."""
    x0 = [float(v) for v in np.asarray(x0, float).reshape(-1)]
    lb = [float(v) for v in np.asarray(lb, float).reshape(-1)]
    ub = [float(v) for v in np.asarray(ub, float).reshape(-1)]

    def f(u):
        return float(fun(bnd_xform(u, lb, ub, sin)))

    u, fval, nf, it, _, _ = nelder_mead(f, initial_simplex(bnd_start(x0, lb, ub, asin)), max_fun_evals, max_iter,
                                        tol_fun, tol_x)
    return bnd_xform(u, lb, ub, sin), fval, nf, it


def nelder_mead(f, sim, max_fun_evals=2000, max_iter=500, tol_fun=1e-10, tol_x=1e-4):
    """fminsearch's iteration from the simplex `sim` (n+1 vertices, evaluated
    in order).  Returns (best vertex, its value, evaluations, iterations,
    final simplex, final values)."""
    sim = [[float(c) for c in v] for v in sim]
    n = len(sim) - 1
    fv = [f(v) for v in sim]
    fv, sim = _sort(fv, sim)
    nf, it = n + 1, 1
    while nf < max_fun_evals and it < max_iter:
        dfv = _nanmax([abs(fv[0] - fv[j]) for j in range(1, n + 1)])
        dx = _nanmax([abs(sim[j][k] - sim[0][k]) for j in range(1, n + 1) for k in range(n)])
        if dfv <= _nanmax([tol_fun, 10.0 * _eps(fv[0])]) and dx <= _nanmax([tol_x, 10.0 * _eps(_nanmax(sim[0]))]):
            break
        xbar = []
        for k in range(n):
            acc = 0.0
            for j in range(n):
                acc += sim[j][k]
            xbar.append(acc / n)
        vend = sim[n]
        xr = [(1.0 + 1.0) * xbar[k] - 1.0 * vend[k] for k in range(n)]
        fxr = f(xr)
        nf += 1
        shrink = False
        if fxr < fv[0]:
            xe = [(1.0 + 2.0) * xbar[k] - 2.0 * vend[k] for k in range(n)]
            fxe = f(xe)
            nf += 1
            if fxe < fxr:
                sim[n], fv[n] = xe, fxe
            else:
                sim[n], fv[n] = xr, fxr
        elif fxr < fv[n - 1]:
            sim[n], fv[n] = xr, fxr
        elif fxr < fv[n]:
            xc = [(1.0 + 0.5) * xbar[k] - 0.5 * vend[k] for k in range(n)]
            fxc = f(xc)
            nf += 1
            if fxc <= fxr:
                sim[n], fv[n] = xc, fxc
            else:
                shrink = True
        else:
            xcc = [(1.0 - 0.5) * xbar[k] + 0.5 * vend[k] for k in range(n)]
            fxcc = f(xcc)
            nf += 1
            if fxcc < fv[n]:
                sim[n], fv[n] = xcc, fxcc
            else:
                shrink = True
        if shrink:
            for j in range(1, n + 1):
                sim[j] = [sim[0][k] + 0.5 * (sim[j][k] - sim[0][k]) for k in range(n)]
                fv[j] = f(sim[j])
            nf += n
        fv, sim = _sort(fv, sim)
        it += 1
    return np.array(sim[0]), fv[0], nf, it, sim, fv
