"""Grain centre-of-mass fit (fit_grain_COM.m): u2euler_corr and the
objective pinned on the CPU, the batched GPU fit (grs_fit_grain_com) against
the sequential oracle fit."""
import numpy as np
import pytest

import oracle
import scenes
from oracle import grain_com as GC
from oracle.pair_spots import pair_spots


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)


def _frames_at_pixels(s, g):
    _, _, _, ev = oracle.forward_comp(s.geo, s.stack, s.U[g], s.centers[g], s.rot, s.hkl, events=True)
    frames = [[] for _ in range(s.rot.size)]
    for om, _h, dy, dz, *_ in ev:
        frames[int(om)].append([dy, dz, 30.0])
    return [np.array(f).reshape(-1, 3) for f in frames]


def _rows(s, g, frames, shift=(0.0, 0.0, 0.0)):
    _, _, _, rows = pair_spots(s.geo, s.stack, s.U[g], s.centers[g] + np.asarray(shift), s.rot, s.hkl, frames)
    return rows


def test_u2euler_corr_round_trip():
    from grainrecon_amd.engine import u2euler

    rng = np.random.default_rng(5)
    for _ in range(100):
        e = np.array([rng.uniform(0, 360), rng.uniform(0.1, 179.9), rng.uniform(0, 360)])
        U = oracle.euler2u_deg(*e)
        a = GC.u2euler_corr(U)
        np.testing.assert_allclose(oracle.euler2u_deg(*a), U, atol=1e-12)
        assert np.all((a >= 0) & (a < 360))
        np.testing.assert_array_equal(u2euler(U.flatten(order="F")[None, :])[0], a)


def test_objective_zero_at_the_pairs_own_grain(scene):
    s = scene
    frames = _frames_at_pixels(s, 0)
    rows = _rows(s, 0, frames)
    assert rows.shape[0] > 20
    x_true = np.concatenate([s.centers[0], GC.u2euler_corr(rows[0, 0:9].reshape(3, 3, order="F"))])
    f0 = GC.objective(s.geo, rows, x_true)
    assert f0 < 0.05
    assert GC.objective(s.geo, rows, x_true + np.array([0.0, 0.03, 0.0, 0, 0, 0])) > f0 + 1.0
    assert GC.objective(s.geo, rows, x_true + np.array([0, 0, 0, 0.3, 0.0, 0.0])) > f0 + 0.5


@pytest.mark.gpu
def test_fit_grain_com_equals_oracle(engine, scene):
    s = scene
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    sets = []
    for g, shift in ((0, (0.002, -0.001, 0.0)), (1, (0.0, 0.0, 0.0)), (2, (-0.003, 0.0, 0.002))):
        rows = _rows(s, g, _frames_at_pixels(s, g), shift)
        if rows.shape[0]:
            sets.append(rows)
    assert len(sets) >= 2
    # a grain with more pairs than a wave has lanes (its rows three times, so
    # the in-order sum runs over several 64-pair chunks)
    sets.append(np.concatenate([sets[0]] * max(3, 130 // sets[0].shape[0] + 1)))
    assert sets[-1].shape[0] > 128
    print(f"fit_grain_com: {len(sets)} grains, {[r.shape[0] for r in sets]} pairs", flush=True)
    for lim in ((2000, 500), (40, 500), (2000, 7)):  # defaults, the evaluation limit, the iteration limit
        x, fv, nf, it = engine.fit_grain_com(sets, max_fun_evals=lim[0], max_iter=lim[1])
        for m, rows in enumerate(sets):
            rx, rf, rnf, rit = GC.fit_grain_com(s.geo, rows, max_fun_evals=lim[0], max_iter=lim[1])
            np.testing.assert_array_equal(x[m], rx)
            assert fv[m] == rf and nf[m] == rnf and it[m] == rit, (lim, m)
        if lim[0] == 40:
            assert np.all(nf >= 40) and np.all(nf <= 40 + 6)
        if lim[1] == 7:
            assert np.all(it == 7)
    from grainrecon_amd._lib import GrsError

    with pytest.raises(GrsError):
        engine.fit_grain_com(sets, max_iter=100001)
