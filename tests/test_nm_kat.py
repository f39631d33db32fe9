"""Known-answer pins of the Nelder-Mead restatements: the oracle's
(oracle/fminsearch.py) and the engine's state machine (grs_nm.h, driven on
the host through tests/nm_harness.hip).

The hand-derived sequence below starts from a dyadic simplex, so every trial
point is exact in binary, and visits one branch of fminsearch's iteration
per step (Lagarias, Reeds, Wright, Wright, SIAM J. Optim. 9 (1998) 112-147,
with fminsearch's coefficients rho = 1, chi = 2, psi = 1/2, sigma = 1/2):

  it 1  reflection accepted      f(r) in [f_best, f_second)
  it 2  expansion accepted       f(e) < f(r) < f_best
  it 3  outside contraction      f(r) in [f_second, f_worst), f(c) <= f(r)
  it 4  inside contraction       f(r) >= f_worst, f(cc) < f_worst
  it 5  shrink after a failed inside contraction
  it 6  expansion rejected       f(r) < f_best <= f(e)
  it 7  shrink after a failed outside contraction (f(c) > f(r))

The objective is a table of values on exactly the points this sequence
visits: a wrong coefficient, branch test, centroid, vertex order or shrink
formula evaluates a point outside the table and fails, as does any
difference in the evaluation count.  Derivation (simplex sorted best first):

  start  A=(0,0):0  B=(4,0):1  C=(0,4):2
  it 1   xbar=(2,0)        r=(4,-4):0.5                      -> A, r, B
  it 2   xbar=(2,-2)       r=(0,-4):-1   e=(-2,-6):-2        -> e, A, (4,-4)
  it 3   xbar=(-1,-3)      r=(-6,-2):0.25  c=(-3.5,-2.5):0.125
  it 4   xbar=(-1,-3)      r=(1.5,-3.5):3  cc=(-2.25,-2.75):0.0625
  it 5   xbar=(-1,-3)      r=(0.25,-3.25):5  cc=(-1.625,-2.875):7, shrink to
         (-2,-6): (-1,-3):-1.5, (-2.125,-4.375):-3
  it 6   xbar=(-2.0625,-5.1875)  r=(-3.125,-7.375):-4  e=(-4.1875,-9.5625):-3.5
  it 7   xbar=(-2.625,-5.875)    r=(-3.25,-5.75):-2.5  c=(-2.9375,-5.8125):-2.25,
         shrink to (-3.125,-7.375): (-2.625,-5.875):-3.75, (-2.5625,-6.6875):-3.9
  stop   MaxIter = 8, after 3 + 1 + 2 + 2 + 2 + 4 + 2 + 4 = 20 evaluations
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from oracle.fminsearch import bnd_start, bnd_xform, fminsearchbnd, initial_simplex, nelder_mead

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "grainrecon_amd", "csrc")

TABLE = {
    (0.0, 0.0): 0.0, (4.0, 0.0): 1.0, (0.0, 4.0): 2.0,
    (4.0, -4.0): 0.5,
    (0.0, -4.0): -1.0, (-2.0, -6.0): -2.0,
    (-6.0, -2.0): 0.25, (-3.5, -2.5): 0.125,
    (1.5, -3.5): 3.0, (-2.25, -2.75): 0.0625,
    (0.25, -3.25): 5.0, (-1.625, -2.875): 7.0, (-1.0, -3.0): -1.5, (-2.125, -4.375): -3.0,
    (-3.125, -7.375): -4.0, (-4.1875, -9.5625): -3.5,
    (-3.25, -5.75): -2.5, (-2.9375, -5.8125): -2.25, (-2.625, -5.875): -3.75, (-2.5625, -6.6875): -3.9,
}
SEQUENCE = [
    (0.0, 0.0), (4.0, 0.0), (0.0, 4.0),                 # starting simplex
    (4.0, -4.0),                                        # it 1 reflection, accepted
    (0.0, -4.0), (-2.0, -6.0),                          # it 2 reflection, expansion accepted
    (-6.0, -2.0), (-3.5, -2.5),                         # it 3 reflection, outside contraction accepted
    (1.5, -3.5), (-2.25, -2.75),                        # it 4 reflection, inside contraction accepted
    (0.25, -3.25), (-1.625, -2.875), (-1.0, -3.0), (-2.125, -4.375),   # it 5 inside contraction fails, shrink
    (-3.125, -7.375), (-4.1875, -9.5625),               # it 6 expansion rejected
    (-3.25, -5.75), (-2.9375, -5.8125), (-2.625, -5.875), (-2.5625, -6.6875),  # it 7 outside fails, shrink
]
FINAL = [((-3.125, -7.375), -4.0), ((-2.5625, -6.6875), -3.9), ((-2.625, -5.875), -3.75)]
NF, IT = 20, 8
START = [[0.0, 0.0], [4.0, 0.0], [0.0, 4.0]]


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("nm") / "libnmh.so")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
                    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-Wno-unused-command-line-argument", "-I", CSRC, "-o", out, os.path.join(HERE, "nm_harness.hip")],
                   check=True, capture_output=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.OBJ = ctypes.CFUNCTYPE(ctypes.c_double, dp, ctypes.c_int)
    L.nm_run_simplex.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 L.OBJ, dp, dp, ip, ip]
    L.nm_run_follow_ups.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    L.OBJ, dp, dp, ip, ip, dp, ctypes.c_int, ip, ip]
    L.nm_start.argtypes = [ctypes.c_int, dp, dp, dp, dp]
    L.nm_fminsearchbnd.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, L.OBJ, dp, dp, ip, ip]
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _table_objective(visited):
    def f(u):
        key = tuple(float(c) for c in u)
        visited.append(key)
        return TABLE[key]  # a point off the derived sequence raises KeyError
    return f


def test_oracle_follows_the_derived_sequence():
    visited = []
    best, fval, nf, it, sim, fv = nelder_mead(_table_objective(visited), START, max_fun_evals=2000, max_iter=8,
                                              tol_fun=0.0, tol_x=0.0)
    assert visited == SEQUENCE
    assert (nf, it) == (NF, IT)
    assert [(tuple(v), f) for v, f in zip(sim, fv)] == FINAL
    assert tuple(best) == FINAL[0][0] and fval == FINAL[0][1]


def test_engine_state_machine_follows_the_derived_sequence(harness):
    visited = []
    f = _table_objective(visited)

    @harness.OBJ
    def cb(u, n):
        try:
            return f([u[k] for k in range(n)])
        except KeyError:
            return float("nan")  # the assertion on `visited` reports the stray point

    sim0 = np.array(START, dtype=float)
    sim = np.zeros(6)
    fv = np.zeros(3)
    nf, it = ctypes.c_int(), ctypes.c_int()
    assert harness.nm_run_simplex(2, _dp(sim0), 2000, 8, 0.0, 0.0, cb, _dp(sim), _dp(fv), ctypes.byref(nf),
                                  ctypes.byref(it)) == 0
    assert visited == SEQUENCE
    assert (nf.value, it.value) == (NF, IT)
    assert [(tuple(sim[2 * j:2 * j + 2]), fv[j]) for j in range(3)] == FINAL


def test_stopping_test_and_budget():
    """Stops at the first iteration whose simplex spread is within both
    tolerances; a budget of 6 evaluations ends after iteration 2 (6
    evaluations: the loop tests nf < MaxFunEvals before each iteration)."""
    visited = []
    _, _, nf, it, _, _ = nelder_mead(_table_objective(visited), START, max_fun_evals=6, max_iter=100, tol_fun=0.0,
                                     tol_x=0.0)
    assert (nf, it) == (6, 3) and visited == SEQUENCE[:6]
    # f spread 2 <= TolFun and vertex spread 4 <= TolX: no iteration at all
    visited = []
    _, _, nf, it, _, _ = nelder_mead(_table_objective(visited), START, tol_fun=2.0, tol_x=4.0)
    assert (nf, it) == (3, 1) and visited == SEQUENCE[:3]
    # one tolerance short: iterate
    visited = []
    _, _, nf, it, _, _ = nelder_mead(_table_objective(visited), START, max_iter=2, tol_fun=2.0, tol_x=3.99)
    assert (nf, it) == (4, 2)


def _run_follow_ups(harness, f, sim0, max_it, tol_f, tol_x, cap=4096):
    n = sim0.shape[1]

    @harness.OBJ
    def cb(u, nn):
        return f([u[k] for k in range(nn)])

    sim = np.zeros((n + 1) * n)
    fv = np.zeros(n + 1)
    fed = np.zeros(cap * n)
    nf, it, n_fed = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    hits = (ctypes.c_int * 3)()
    assert harness.nm_run_follow_ups(n, _dp(sim0), 2000, max_it, tol_f, tol_x, cb, _dp(sim), _dp(fv),
                                     ctypes.byref(nf), ctypes.byref(it), _dp(fed), cap, ctypes.byref(n_fed),
                                     hits) == 0
    assert n_fed.value <= cap
    pts = [tuple(fed[n * j:n * j + n]) for j in range(n_fed.value)]
    return sim.reshape(n + 1, n), fv, nf.value, it.value, pts, list(hits)


def test_follow_ups_feed_the_derived_sequence(harness):
    """The engine's fits evaluate each reflection's three follow-ups
    (grs_nm::follow_ups) in the same launch and feed the expansion or
    contraction the method requests from them.  On the derived sequence
    (off-table points are NaN, which only a wrongly fed follow-up would
    reach) the points fed, the counts and the final simplex are the known
    answer, and every follow-up kind is used: the expansions of it 2 and 6,
    the outside contractions of it 3 and 7, the inside ones of it 4 and 5."""
    f = lambda u: TABLE.get(tuple(float(c) for c in u), math.nan)  # noqa: E731
    sim, fv, nf, it, fed, hits = _run_follow_ups(harness, f, np.array(START, dtype=float), 8, 0.0, 0.0)
    assert fed == SEQUENCE
    assert (nf, it) == (NF, IT)
    assert [(tuple(sim[j]), fv[j]) for j in range(3)] == FINAL
    assert hits == [2, 2, 2]


def test_follow_ups_equal_the_plain_drive_bit_for_bit(harness):
    """On a three-dimensional Rosenbrock run to convergence, driving the
    state machine with follow-ups gives the plain drive's points, values and
    counts bit for bit."""
    def rosen(u):
        return 100.0 * (u[1] - u[0] * u[0]) ** 2 + (1.0 - u[0]) ** 2 + 0.5 * (u[2] - 0.25) ** 2

    sim0 = np.array([[-1.2, 1.0, 0.9], [-1.26, 1.0, 0.9], [-1.2, 1.05, 0.9], [-1.2, 1.0, 0.945]])
    visited = []

    @harness.OBJ
    def cb(u, n):
        p = [u[k] for k in range(n)]
        visited.append(tuple(p))
        return rosen(p)

    sim = np.zeros(12)
    fv = np.zeros(4)
    nf, it = ctypes.c_int(), ctypes.c_int()
    assert harness.nm_run_simplex(3, _dp(sim0), 2000, 500, 1e-12, 1e-9, cb, _dp(sim), _dp(fv), ctypes.byref(nf),
                                  ctypes.byref(it)) == 0
    sim2, fv2, nf2, it2, fed, hits = _run_follow_ups(harness, rosen, sim0, 500, 1e-12, 1e-9)
    assert fed == visited
    assert np.array_equal(sim2.reshape(-1), sim) and np.array_equal(fv2, fv) and (nf2, it2) == (nf.value, it.value)
    assert min(hits) > 0 and nf.value > 200


def test_nan_vertex_sorts_last(harness):
    """A NaN value sorts after every number (stable sort, NaN last); the
    reflection then takes the NaN vertex's place."""
    vals = {(0.0, 0.0): float("nan"), (4.0, 0.0): 1.0, (0.0, 4.0): 2.0, (4.0, 4.0): 1.5}
    seen = []

    def f(u):
        seen.append(tuple(u))
        return vals[tuple(u)]

    _, _, nf, it, sim, fv = nelder_mead(f, START, max_iter=2, tol_fun=0.0, tol_x=0.0)
    # sorted start B(1), C(2), A(NaN); r = 2*(2,2) - (0,0) = (4,4): 1.5 in [1, 2) -> accepted in A's place
    assert seen == [(0.0, 0.0), (4.0, 0.0), (0.0, 4.0), (4.0, 4.0)]
    assert [tuple(v) for v in sim] == [(4.0, 0.0), (4.0, 4.0), (0.0, 4.0)] and fv == [1.0, 1.5, 2.0]
    assert (nf, it) == (4, 2)

    seen_e = []

    @harness.OBJ
    def cb(u, n):
        k = (u[0], u[1])
        seen_e.append(k)
        return vals.get(k, float("inf"))

    sim0 = np.array(START, dtype=float)
    s = np.zeros(6)
    v = np.zeros(3)
    a, b = ctypes.c_int(), ctypes.c_int()
    harness.nm_run_simplex(2, _dp(sim0), 2000, 2, 0.0, 0.0, cb, _dp(s), _dp(v), ctypes.byref(a), ctypes.byref(b))
    assert seen_e == seen and list(v) == fv and (a.value, b.value) == (nf, it)


def test_bounded_start_and_transform_closed_forms(harness):
    """fminsearchbnd.m:152-165, 263-268: x0 at 3/4 of [lb, ub] maps to
    u0 = 2 pi + asin(1/2) = 13 pi / 6, a start on or beyond a bound to -+pi/2;
    the first simplex scales u0 coordinate-wise by 1.05; the transform maps
    u0 back to x0 and sin(u) = +-1 to the bounds."""
    lb, ub = [-1.0, 2.0, 0.0], [3.0, 6.0, 1.0]
    x0 = [2.0, 2.0, 5.0]          # 3/4 of the way, on lb, beyond ub
    u0 = bnd_start(x0, lb, ub)
    assert abs(u0[0] - 13 * math.pi / 6) < 1e-15 and u0[1] == -math.pi / 2 and u0[2] == math.pi / 2
    sim = initial_simplex(u0)
    assert len(sim) == 4 and sim[0] == u0
    for j in range(3):
        expect = list(u0)
        expect[j] = u0[j] * 1.05
        assert np.allclose(sim[j + 1], expect, rtol=1e-15, atol=0)
    assert initial_simplex([0.0, 1.0])[1] == [0.00025, 1.0]
    x = bnd_xform(u0, lb, ub)
    assert abs(x[0] - 2.0) < 1e-15 and x[1] == 2.0 and x[2] == 1.0
    # the engine builds the same start with its own asin
    xe, lbe, ube = (np.array(a, dtype=float) for a in (x0, lb, ub))
    out = np.zeros(12)
    harness.nm_start(3, _dp(xe), _dp(lbe), _dp(ube), _dp(out))
    eng = out.reshape(4, 3)
    assert abs(eng[0, 0] - 13 * math.pi / 6) < 1e-15 and eng[0, 1] == -math.pi / 2 and eng[0, 2] == math.pi / 2
    asin_e = lambda z: float(oracle.math_eval(7, [z])[0])  # noqa: E731
    assert np.array_equal(eng, np.array(initial_simplex(bnd_start(x0, lb, ub, asin=asin_e))))


def test_engine_and_oracle_fminsearchbnd_agree_bit_for_bit(harness):
    """Both restatements with the engine's FP64 sin/asin on a bounded
    Rosenbrock: the same point, value, evaluation and iteration counts."""
    sin_e = lambda u: float(oracle.math_eval(0, [u])[0])  # noqa: E731
    asin_e = lambda z: float(oracle.math_eval(7, [z])[0])  # noqa: E731

    def rosen(x):
        return 100.0 * (x[1] - x[0] * x[0]) ** 2 + (1.0 - x[0]) ** 2 + 0.5 * (x[2] - 0.25) ** 2

    x0, lb, ub = [-1.2, 1.0, 0.9], [-2.0, -2.0, -1.0], [2.0, 2.0, 1.0]
    x, fv, nf, it = fminsearchbnd(rosen, x0, lb, ub, max_fun_evals=2000, max_iter=500, tol_fun=1e-12, tol_x=1e-9,
                                  sin=sin_e, asin=asin_e)

    @harness.OBJ
    def cb(u, n):
        return rosen([u[k] for k in range(n)])

    xe = np.zeros(3)
    fe = ctypes.c_double()
    a, b = ctypes.c_int(), ctypes.c_int()
    harness.nm_fminsearchbnd(3, _dp(np.array(x0)), _dp(np.array(lb)), _dp(np.array(ub)), 2000, 500, 1e-12, 1e-9, cb,
                             _dp(xe), ctypes.byref(fe), ctypes.byref(a), ctypes.byref(b))
    assert np.array_equal(xe, x) and fe.value == fv and (a.value, b.value) == (nf, it)
    assert np.allclose(x, [1.0, 1.0, 0.25], atol=1e-4)
