"""Orientation fit of one voxel -- TEST INFRASTRUCTURE ONLY (checker of the
engine's batched grs_fit_or_local).

fit_OR_constrain_with_local_comp_cuda.m:116-126: fminsearchbnd of
fit_OR_fun_local.m (1 - Nr_intersect/Nr_simu for U from q = [sqrt(1-|x|^2) x],
:9-29, :98) over x = (q1, q2, q3) within max(x0-0.02,-1) .. min(x0+0.02,1),
MaxFunEvals 2000, MaxIter 2000, TolFun 1e-10, TolX default 1e-4.  Counts come
from the C oracle (ForwardCalc_OR_space_comp.m semantics: q0 recomputed),
sin/asin from the engine's FP64 math as compiled into the oracle library.
"""
from __future__ import annotations

import math

import numpy as np

from . import math_eval, or_space_comp
from .fminsearch import fminsearchbnd


def _sin(u: float) -> float:
    return float(math_eval(0, np.array([u]))[0])


def _asin(z: float) -> float:
    return float(math_eval(7, np.array([z]))[0])


def objective(geo, stack, pos, x, rot, hkl) -> float:
    out = or_space_comp(geo, stack, pos, np.array([[0.0, x[0], x[1], x[2]]]), rot, hkl, nthreads=1)
    ns, ni = out[0, 0], out[0, 1]
    return 1.0 - ni / ns if ns > 0 else math.nan


def fit_or_local(geo, stack, pos, x0, rot, hkl, half_width=0.02, max_fun_evals=2000, max_iter=2000,
                 tol_fun=1e-10, tol_x=1e-4):
    """Returns (x, fval, n_fun_evals, n_iter)."""
    x0 = [float(v) for v in x0]
    lb = [max(v - half_width, -1.0) for v in x0]
    ub = [min(v + half_width, 1.0) for v in x0]
    return fminsearchbnd(lambda x: objective(geo, stack, pos, x, rot, hkl), x0, lb, ub, max_fun_evals, max_iter,
                         tol_fun, tol_x, sin=_sin, asin=_asin)
