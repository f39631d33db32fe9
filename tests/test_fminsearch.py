"""The bounded Nelder-Mead baseline (fminsearchbnd semantics) on problems with known minima."""
import numpy as np

from oracle.fminsearch import fminsearchbnd


def test_rosenbrock_interior_minimum():
    f = lambda x: 100.0 * (x[1] - x[0] ** 2) ** 2 + (1.0 - x[0]) ** 2  # noqa: E731
    x, fv, nf, it = fminsearchbnd(f, [-1.2, 1.0], [-5, -5], [5, 5], tol_x=1e-8)
    assert np.allclose(x, [1.0, 1.0], atol=1e-4)
    assert fv < 1e-8 and nf <= 2000 and it <= 500


def test_minimum_on_the_bound_is_reached_not_crossed():
    # unconstrained minimum at (2, -3); box [-1, 1]^2 -> constrained minimum at (1, -1)
    f = lambda x: (x[0] - 2.0) ** 2 + (x[1] + 3.0) ** 2  # noqa: E731
    x, fv, _, _ = fminsearchbnd(f, [0.0, 0.0], [-1, -1], [1, 1], tol_x=1e-10)
    assert np.all(x <= 1.0) and np.all(x >= -1.0)
    assert np.allclose(x, [1.0, -1.0], atol=1e-6)
    assert abs(fv - 5.0) < 1e-9


def test_budget_is_respected():
    calls = []

    def f(x):
        calls.append(1)
        return float(np.sum(np.cos(3 * x) + x * x))

    x, fv, nf, it = fminsearchbnd(f, np.ones(7) * 0.3, -np.ones(7), np.ones(7), max_fun_evals=200, tol_fun=0, tol_x=0)
    assert len(calls) == nf
    assert nf <= 200 + 7   # one shrink can overrun the budget by at most n evaluations, as in fminsearch
