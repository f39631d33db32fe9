"""Orientation fit (fit_OR_constrain_with_local_comp_cuda.m:116-126): the
batched device fit against the sequential oracle fit, step for step."""
import numpy as np
import pytest

import oracle
import scenes
from oracle.fit_or import fit_or_local, objective


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene(seed=33, ngrains=3, spread_mm=0.05, nrot=37)


def _starts(s, rng, k):
    pos, x0 = [], []
    for i in range(k):
        g = i % 3
        pos.append(s.voxels[g][int(rng.integers(0, 27))])
        x0.append(s.q[g, 1:] + rng.uniform(-0.006, 0.006, 3))
    return np.array(pos), np.array(x0)


def test_oracle_fit_improves_completeness(scene):
    s = scene
    rng = np.random.default_rng(1)
    pos, x0 = _starts(s, rng, 1)
    f0 = objective(s.geo, s.stack, pos[0], x0[0], s.rot, s.hkl)
    x, fv, nf, it = fit_or_local(s.geo, s.stack, pos[0], x0[0], s.rot, s.hkl)
    assert fv <= f0 and nf <= 2000 + 3
    assert np.all(np.abs(x - x0[0]) <= 0.02)
    assert fv == objective(s.geo, s.stack, pos[0], x, s.rot, s.hkl)


def test_oracle_fit_respects_budget(scene):
    s = scene
    pos, x0 = _starts(s, np.random.default_rng(2), 1)
    _, _, nf, it = fit_or_local(s.geo, s.stack, pos[0], x0[0], s.rot, s.hkl, max_fun_evals=12)
    assert nf <= 12 + 3


@pytest.mark.gpu
def test_batched_fit_equals_sequential_oracle(engine, scene):
    s = scene
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(3)
    pos, x0 = _starts(s, rng, 6)
    x0[5] = [0.9, 0.5, 0.3]          # |q_vec| > 1 everywhere near it: NaN objective
    x, fv, nf, it = engine.fit_or_local(pos, x0)
    for m in range(pos.shape[0]):
        rx, rf, rnf, rit = fit_or_local(s.geo, s.stack, pos[m], x0[m], s.rot, s.hkl)
        np.testing.assert_array_equal(x[m], rx)
        assert (fv[m] == rf) or (np.isnan(fv[m]) and np.isnan(rf))
        assert nf[m] == rnf and it[m] == rit
    assert np.sum(fv[:5] < 0.5) >= 3 and np.isnan(fv[5]) and nf[5] >= 2000
