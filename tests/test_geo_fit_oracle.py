"""The CPU baseline of the geometry fit (oracle/geo_fit.py: geo_fitting.m and
the FitAllOnce loop of L_shift_tilt_fitting.m) against the torch FP64
restatement of the spot construction (oracle/residual_ref.py) and against
known answers."""
import numpy as np
import torch

import oracle
import scenes
from grainrecon_amd.geometry import labdct_geometry
from oracle import residual_ref
from oracle.geo_fit import GeoObjective, fit_all_once

TILTS = (0.2, -0.15, 0.1)


def _problem():
    s = scenes.make_scene(seed=71, ngrains=3, spread_mm=0.05, nrot=37)
    geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=256, pixel_mm=6.72e-3 * 4, tilts=TILTS, dety00=0.02,
                          detz00=-0.03)
    hkl, om, grain = [], [], []
    for g in range(3):
        _, _, _, ev = oracle.forward_comp(geo, s.stack, s.U[g], s.centers[g], s.rot, s.hkl, events=True)
        hkl.append(s.hkl[ev[:, 1].astype(int)])
        om.append(s.rot[ev[:, 0].astype(int)])
        grain.append(np.full(ev.shape[0], g))
    gr6 = np.array([np.concatenate([s.q[g, 1:] / s.q[g, 0], s.centers[g]]) for g in range(3)])
    g7 = np.array([geo.Lsam2sou, geo.Lsam2det, geo.dety00, geo.detz00, *TILTS])
    return geo, np.vstack(hkl), np.concatenate(om), np.concatenate(grain), gr6, g7


def test_simulated_pixels_equal_the_torch_restatement():
    geo, hkl, om, grain, gr6, g7 = _problem()
    rng = np.random.default_rng(3)
    meas = rng.uniform(0, 256, (hkl.shape[0], 2))
    obj = GeoObjective(geo, hkl, om, meas, grain, gr6)
    assert obj.n > 100
    for x in (g7, g7 + np.array([-0.2, 0.3, 0.05, -0.05, 0.1, -0.1, 0.2])):
        r = residual_ref.residual(geo, dict(hkl=hkl, omega=om, meas=meas, grain=grain),
                                  torch.tensor(x, dtype=torch.float64), torch.tensor(gr6, dtype=torch.float64))
        sim = r.numpy() + meas
        assert np.abs(obj.sim(x) - sim).max() < 1e-8
        assert np.abs(obj.sim_numpy(x) - sim).max() < 1e-8
        rounded = np.sign(sim) * np.floor(np.abs(sim) + 0.5)
        assert abs(obj(x) - np.mean(np.sqrt(((rounded - meas) ** 2).sum(1)))) < 1e-12


def test_objective_is_zero_at_the_truth_and_the_fit_lowers_it():
    """Measured centroids at the rounded simulated pixels of the true geometry:
    geo_fitting is exactly 0 there, positive at a perturbed start, and the
    FitAllOnce loop (re-centred while a bound is within 0.1) lowers it."""
    geo, hkl, om, grain, gr6, g7 = _problem()
    probe = GeoObjective(geo, hkl, om, np.zeros((hkl.shape[0], 2)), grain, gr6)
    sim = probe.sim(g7)
    meas = np.sign(sim) * np.floor(np.abs(sim) + 0.5)
    obj = GeoObjective(geo, hkl, om, meas, grain, gr6)
    assert obj(g7) == 0.0
    px = geo.pixelysize
    x0 = g7 + np.array([-0.05, 0.05, 3 * px, -3 * px, 0.05, -0.05, 0.05])
    f0 = obj(x0)
    x, f, nf, rounds = fit_all_once(obj, x0, px, max_rounds=3)
    assert f0 > 1.0 and f < 0.5 * f0 and nf > 100 and 1 <= rounds <= 3
    assert obj.evals >= nf
