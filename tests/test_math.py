"""Pins the engine's FP64 elementary functions (grs_math.h) on the host.

The same header is compiled into the gfx950 kernels; tests/test_engine_gpu.py
checks that the device returns the same bits.  Here the host build (through
the oracle library) is checked against mpmath at 40 digits.
"""
import math
import os
import re

import mpmath as mp
import numpy as np
import pytest

import oracle

mp.mp.dps = 40
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def max_ulp(fn_id, mpfn, xs):
    ys = oracle.math_eval(fn_id, xs)
    worst = 0.0
    for x, y in zip(xs, ys):
        ex = mpfn(mp.mpf(float(x)))
        exf = float(ex)
        u = math.ulp(exf) if exf != 0.0 else math.ulp(0.0)
        err = abs(mp.mpf(float(y)) - ex) / u
        worst = max(worst, float(err))
    return worst


def test_header_coefficients_match_generator():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", os.path.join(ROOT, "grainrecon_amd", "csrc",
                                                                      "gen_math_coeffs.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    block = gen.c_block()
    header = open(os.path.join(ROOT, "grainrecon_amd", "csrc", "grs_math.h")).read()
    assert block in header


@pytest.mark.parametrize("fn_id,mpfn,lo,hi", [
    (0, mp.sin, -7.0, 7.0),
    (1, mp.cos, -7.0, 7.0),
    (0, mp.sin, -0.8, 0.8),
    (1, mp.cos, -0.8, 0.8),
    (2, mp.acos, -1.0, 1.0),
    (3, mp.atan, -8.0, 8.0),
    (7, mp.asin, -1.0, 1.0),
    (7, mp.asin, -0.5, 0.5),
])
def test_under_one_ulp(fn_id, mpfn, lo, hi):
    rng = np.random.default_rng(fn_id * 7 + int(hi))
    xs = rng.uniform(lo, hi, 4000)
    assert max_ulp(fn_id, mpfn, xs) < 1.0


def test_acos_near_endpoints_and_branch_points():
    xs = []
    for c in (-1.0, -0.5, 0.0, 0.5, 1.0):
        for k in range(1, 60):
            for s in (-1, 1):
                x = c + s * 2.0 ** (-k)
                if -1.0 <= x <= 1.0:
                    xs.append(x)
    xs += [-1.0, -0.5, 0.0, -0.0, 0.5, 1.0, math.nextafter(0.5, 1), math.nextafter(-0.5, -1)]
    assert max_ulp(2, mp.acos, np.array(xs)) < 1.0
    y = oracle.math_eval(2, np.array([1.0, -1.0, 0.0, -0.0, 1.0000000000000002, -1.0000000000000002, np.nan]))
    assert y[0] == 0.0 and y[1] == math.pi and y[2] == math.pi / 2 and y[3] == math.pi / 2
    assert np.isnan(y[4]) and np.isnan(y[5]) and np.isnan(y[6])


def test_sin_cos_multiples_of_pi_over_2_and_angles_in_degrees():
    # omega = rot*pi/180 for the rotation grids used by the engine
    rot = np.concatenate([np.arange(0, 361, 1.0), np.arange(0, 361, 2.0), np.linspace(0, 360, 37)])
    om = rot * np.pi / 180
    assert max_ulp(0, mp.sin, om) < 1.0
    assert max_ulp(1, mp.cos, om) < 1.0
    y = oracle.math_eval(0, np.array([0.0, -0.0, np.inf, np.nan, 1e7]))
    assert y[0] == 0.0 and math.copysign(1, y[1]) == -1.0
    assert np.isnan(y[2]) and np.isnan(y[3]) and np.isnan(y[4])


def test_atan_large_and_reduction_boundaries():
    xs = np.array([0.4375, 0.6875, 1.1875, 2.4375, 1e3, 1e300, 1e-300, 0.0, -0.0, 3.0, -3.0])
    xs = np.concatenate([xs, np.nextafter(xs, np.inf), np.nextafter(xs, -np.inf)])
    assert max_ulp(3, mp.atan, xs) < 1.0
    assert oracle.math_eval(3, np.array([np.inf]))[0] == math.pi / 2


def test_round_half_away_from_zero():
    xs = np.array([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 0.49999999999999994, -0.49999999999999994, 2.0 ** 52 + 1,
                   -(2.0 ** 52) - 1, 1e300, 3.7, -3.7, 0.0])
    want = np.array([1.0, 2.0, 3.0, -1.0, -2.0, -3.0, 0.0, -0.0, 2.0 ** 52 + 1, -(2.0 ** 52) - 1, 1e300, 4.0,
                     -4.0, 0.0])
    assert np.array_equal(oracle.math_eval(4, xs), want)


def test_asin_near_endpoints_and_branch_points():
    xs = []
    for c in (-1.0, -0.5, 0.0, 0.5, 1.0):
        for k in range(1, 60):
            for s in (-1, 1):
                x = c + s * 2.0 ** (-k)
                if -1.0 <= x <= 1.0:
                    xs.append(x)
    xs += [-1.0, -0.5, 0.0, 0.5, 1.0, 1e-300, -1e-300]
    assert max_ulp(7, mp.asin, np.array(xs)) < 1.0
    y = oracle.math_eval(7, np.array([1.0, -1.0, -0.0, 1.0000000000000002, np.nan]))
    assert y[0] == math.pi / 2 and y[1] == -math.pi / 2
    assert y[2] == 0.0 and math.copysign(1.0, y[2]) < 0
    assert np.isnan(y[3]) and np.isnan(y[4])


def _edges():
    xs = []
    for c in (-1.0, -0.5, 0.0, 0.5, 1.0):
        for k in range(1, 60):
            for s in (-1, 1):
                x = c + s * 2.0 ** (-k)
                if -1.0 <= x <= 1.0:
                    xs.append(x)
    return xs + [-1.0, -0.5, 0.0, 0.5, 1.0]


def test_filter_acos_error_bound():
    """grs_filt_acos (the omega solve's arc ends, widened there by 1e-6 rad):
    absolute error below 5e-12 rad over [-1, 1], at the branch points and
    the ends; NaN outside."""
    rng = np.random.default_rng(11)
    xs = np.concatenate([np.linspace(-1.0, 1.0, 200001), rng.uniform(-1, 1, 20000), _edges()])
    y = oracle.math_eval(8, xs)
    err = max(abs(float(mp.mpf(float(a)) - mp.acos(mp.mpf(float(x))))) for x, a in zip(xs[::7], y[::7]))
    assert err < 5e-12
    assert np.max(np.abs(y - np.arccos(xs))) < 5e-12
    assert np.isnan(oracle.math_eval(8, np.array([1.0000000000000002, -1.5, np.nan]))).all()


def test_filter_atan2_error_bound():
    """grs_filt_atan2 at (sin x, cos x) returns x within 3e-12 rad over the
    whole circle, including the axes, the diagonals and the reduction point
    tan(pi/8) of every octant."""
    rng = np.random.default_rng(12)
    special = [k * math.pi / 8 for k in range(-8, 9)]
    xs = np.concatenate([np.linspace(-math.pi, math.pi, 200001), rng.uniform(-math.pi, math.pi, 20000),
                         special, [s + d for s in special for d in (-1e-9, 1e-9, -1e-15, 1e-15)]])
    xs = xs[(xs > -math.pi) & (xs <= math.pi)]
    y = oracle.math_eval(9, xs)
    d = np.abs((y - xs + math.pi) % (2 * math.pi) - math.pi)
    assert d.max() < 3e-12
    assert (y > -math.pi - 1e-15).all() and (y <= math.pi + 1e-15).all()
