"""Brute-force golden vectors of forward_comp.m (tests/golden/forward_comp_tiny.json,
written by tests/golden/make_golden.py from two independent restatements that
agree on every count, event and median)."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from grainrecon_amd.geometry import Geometry
from oracle.np_restatement import forward_comp_np

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "forward_comp_tiny.json")) as _f:
    GOLDEN = json.load(_f)["cases"]
IDS = [c["name"] for c in GOLDEN]


def _geo(c):
    g = dict(c["geometry"])
    for f in ("S", "B", "RotDet"):
        g[f] = np.array(g[f]).reshape(3, 3)
    g["BeamStopY"], g["BeamStopZ"] = tuple(g["BeamStopY"]), tuple(g["BeamStopZ"])
    return Geometry(**g)


def _proj(c):
    return np.array(c["proj_fortran_order"], np.float32).reshape(c["proj_shape"], order="F")


def _expected(c):
    return np.array([[[np.nan if x is None else x for x in r] for r in rp] for rp in c["expected"]])


def _U(q):
    ss = (q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]
    q0 = math.sqrt(1.0 - ss)
    q1, q2, q3 = q[1], q[2], q[3]
    return np.array([
        [2.0 * (q0 * q0 + q1 * q1) - 1.0, 2.0 * (q1 * q2 - q0 * q3), 2.0 * (q1 * q3 + q0 * q2)],
        [2.0 * (q1 * q2 + q0 * q3), 2.0 * (q0 * q0 + q2 * q2) - 1.0, 2.0 * (q2 * q3 - q0 * q1)],
        [2.0 * (q1 * q3 - q0 * q2), 2.0 * (q2 * q3 + q0 * q1), 2.0 * (q0 * q0 + q3 * q3) - 1.0],
    ])


def test_fixtures_cover_the_edge_cases():
    for c in GOLDEN:
        e = _expected(c)
        assert np.isnan(e[2, :34, 2]).all() and (e[2, :34, 3] == 2000).all()   # Nr_simu = 0
        assert not e[:, 34:, :].any()                                          # |q_vec| > 1 rows
        assert (e[0, 0, 2] == 1.0) and (e[1, 1, 2] == 1.0)                     # rendered grains
        ns = e[:2, :34, 0]
        assert (ns % 2 == 0).any() and (ns % 2 == 1).any()                     # both median branches
        assert ((e[:2, :34, 1] > 0) & (e[:2, :34, 1] < ns)).any()              # partial hits


@pytest.mark.parametrize("case", GOLDEN, ids=IDS)
def test_c_oracle_reproduces_golden(case):
    geo, stack = _geo(case), oracle.Stack(_proj(case))
    e = _expected(case)
    for p, pos in enumerate(case["positions"]):
        out = oracle.or_space_comp(geo, stack, pos, case["q"], case["rot"], case["hkl"], nthreads=2)
        np.testing.assert_array_equal(out, e[p])


@pytest.mark.parametrize("case", GOLDEN, ids=IDS)
def test_numpy_restatement_reproduces_golden(case):
    geo, proj = _geo(case), _proj(case)
    e = _expected(case)
    for p, pos in enumerate(case["positions"][:2]):
        for k, q in enumerate(case["q"][:34]):
            ns, ni, med, info = forward_comp_np(geo, proj, _U(q), pos, case["rot"], case["hkl"])
            assert (ns, ni) == (e[p, k, 0], e[p, k, 1])
            assert np.float32(med) == np.float32(e[p, k, 3])
    for g in range(2):
        _, _, _, info = forward_comp_np(geo, proj, _U(case["q"][g]), case["positions"][g], case["rot"], case["hkl"])
        np.testing.assert_array_equal(info["sel"][:, :4], np.array(case["truth_events"][g]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN, ids=IDS)
@pytest.mark.parametrize("method", [0, 1], ids=["fast", "plain"])
def test_engine_reproduces_golden(engine, case, method):
    engine.set_geometry(_geo(case), np.array(case["hkl"]), np.array(case["rot"]))
    engine.set_distance_stack(_proj(case))
    e = _expected(case)
    q = np.array(case["q"])
    for p, pos in enumerate(case["positions"]):
        ns, ni, C, med = engine.forward_comp(pos, q, method=method)
        np.testing.assert_array_equal(ns, e[p, :, 0])
        np.testing.assert_array_equal(ni, e[p, :, 1])
        np.testing.assert_array_equal(C, e[p, :, 2])
        np.testing.assert_array_equal(med, e[p, :, 3].astype(np.float32))
