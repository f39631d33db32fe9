"""The opt-in adjoint refinement of the mapping loop (grainrecon_amd.mapping,
refine='adjoint'): Levenberg-Marquardt on the paired spots' centroid residual
at the fit_OR_constrain_with_local_comp_cuda.m:117-126 and fit_grain_COM.m:
5-29 call sites in place of fminsearchbnd.

Known answers: with the spot centroids simulated exactly (unrounded) at the
true orientation and centre, the true parameters are the residual's zero up
to the single-precision rounding of the centroids in the pair rows
(forward_comp_pair_spots.m concatenates them with a single; half an ulp of a
256 px coordinate is 7.6e-6 px), so the fits must return them to that level
from the start the reference would use; the Nelder-Mead of the rounded
objective is run beside them for comparison.  The LM's centre-holding mode is pinned on the CPU on a linear
least-squares problem whose solution is known in closed form."""
import math

import numpy as np
import pytest

from grainrecon_amd.fit import levenberg_marquardt


class _LinearFit:
    """r = J x - b for 1 grain with constant J (n, 2, 13): the interface
    levenberg_marquardt uses (G, goff, n, jacobian)."""

    def __init__(self, rng, n=40):
        self.G, self.n = 1, n
        self.goff = np.array([0, n])
        self.J = rng.normal(size=(n, 2, 13))
        self.b = rng.normal(size=(n, 2))

    def jacobian(self, g7, gr6):
        x = np.concatenate([g7, gr6[0]])
        return np.einsum("naj,j->na", self.J, x) - self.b, self.J


def test_lm_fix_position_solves_the_orientation_block_only():
    rng = np.random.default_rng(4)
    f = _LinearFit(rng)
    g7 = rng.normal(size=7)
    gr6 = rng.normal(size=(1, 6))
    res = levenberg_marquardt(f, g7, gr6, fix_geometry=True, fix_position=True, max_iter=200, tol=1e-15)
    np.testing.assert_array_equal(res.g7, g7)
    np.testing.assert_array_equal(res.gr6[0, 3:], gr6[0, 3:])
    # closed form: least squares in the 3 orientation columns with the rest held
    A = f.J[:, :, 7:10].reshape(-1, 3)
    rhs = (f.b - np.einsum("naj,j->na", f.J[:, :, list(range(7)) + [10, 11, 12]],
                           np.concatenate([g7, gr6[0, 3:]]))).reshape(-1)
    x = np.linalg.lstsq(A, rhs, rcond=None)[0]
    np.testing.assert_allclose(res.gr6[0, :3], x, rtol=0, atol=1e-9)
    # and the full grain block without fix_position
    res6 = levenberg_marquardt(f, g7, gr6, fix_geometry=True, max_iter=200, tol=1e-15)
    A6 = f.J[:, :, 7:].reshape(-1, 6)
    rhs6 = (f.b - np.einsum("naj,j->na", f.J[:, :, :7], g7)).reshape(-1)
    np.testing.assert_allclose(res6.gr6[0], np.linalg.lstsq(A6, rhs6, rcond=None)[0], rtol=0, atol=1e-9)


def test_lm_charbonnier_reaches_the_reweighted_stationary_point_and_resists_outliers():
    """delta > 0: at the result the weighted gradient sum w J^T r vanishes (the
    IRLS fixed point of sum sqrt(|r|^2 + delta^2)); with 3 of 40 spots moved
    30 px off, the Charbonnier fit stays near the inliers' least-squares
    solution while plain least squares is pulled away."""
    from grainrecon_amd.fit import host_normal

    rng = np.random.default_rng(9)
    f = _LinearFit(rng)
    g7 = rng.normal(size=7)
    x_true = rng.normal(size=6)
    f.b = np.einsum("naj,j->na", f.J, np.concatenate([g7, x_true])) + rng.normal(scale=0.1, size=(f.n, 2))
    clean = f.b.copy()
    f.b[[3, 17, 29]] += 30.0
    gr6 = np.zeros((1, 6))
    rob = levenberg_marquardt(f, g7, gr6, fix_geometry=True, delta=0.5, max_iter=500, tol=1e-16)
    ne = host_normal(f, g7, rob.gr6, 0.5)
    assert np.abs(ne["bk"]).max() < 1e-8 * max(1.0, np.abs(ne["Akk"]).max())
    ls = levenberg_marquardt(f, g7, gr6, fix_geometry=True, max_iter=200, tol=1e-16)
    f_clean = _LinearFit(np.random.default_rng(9))
    f_clean.J, f_clean.b = f.J, clean
    ref = levenberg_marquardt(f_clean, g7, gr6, fix_geometry=True, max_iter=200, tol=1e-16)
    err_rob = np.abs(rob.gr6 - ref.gr6).max()
    err_ls = np.abs(ls.gr6 - ref.gr6).max()
    assert err_rob < 0.2 * err_ls


def test_refine_argument_is_checked_before_any_engine_call():
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine

    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=2.0, max_dmedian=20.0, nr_seed=6, first_grain_id=1,
              center=[0, 0, 0], vox_size=0.005, rvp_first=[1, 1, 1], tomo_dim=(4, 4, 4))
    ds = MappingDS(np.ones((4, 4, 4)))
    with pytest.raises(ValueError, match="refine"):
        grain_mapping_engine(None, np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]), ds, refine="gradient", **kw)
    with pytest.raises(ValueError, match="geo7"):
        grain_mapping_engine(None, np.zeros((1, 3)), np.array([[1.0, 0, 0, 0]]), ds, refine="adjoint", **kw)


def _misorientation_deg(qa, qb):
    d = min(1.0, abs(float(np.dot(qa, qb))))
    return math.degrees(2.0 * math.acos(d))


def _exact_frames(engine, s, grains, area):
    from grainrecon_amd import fit as F

    com = np.array([s.centers[g] for g in grains])
    sp = F.simulate_matched_spots(engine, None, com, np.stack([s.U[g] for g in grains]))
    frames = []
    for i in range(s.rot.size):
        yz = sp["meas"][sp["omega_idx"] == i]
        frames.append(np.column_stack([yz, np.full(yz.shape[0], area)]))
    return frames


def _geo7(geo):
    return np.array([geo.Lsam2sou, geo.Lsam2det, geo.dety00, geo.detz00, 0.0, 0.0, 0.0])


@pytest.mark.gpu
def test_adjoint_orientation_fit_recovers_the_true_orientation(engine):
    """At grain 0's centre, the start of fit_OR_constrain (the best local
    sample) is a few tenths of a degree off; the LM of the Rodrigues vector
    on the paired exact centroids returns the true orientation.  The
    Nelder-Mead of 1 - completeness from the same start is run beside it."""
    from grainrecon_amd.mapping import index_seed_adjoint
    from test_index import _index_problem

    s, qs, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_exact_frames(engine, s, [0], 400.0))
    engine.set_orientation_set(qs)
    pos = s.centers[0]
    print("index_seed_adjoint at grain 0's centre", flush=True)
    row, cand, info = index_seed_adjoint(engine, _geo7(s.geo), pos, 6, 1)
    nm = engine.index_seeds(pos[None, :], nr_seed=6)
    x0 = nm["x0"][0]
    q_start = np.concatenate([[math.sqrt(1.0 - x0 @ x0)], x0])
    err_start = _misorientation_deg(q_start, s.q[0])
    err_adj = _misorientation_deg(row[14:18], s.q[0])
    err_nm = _misorientation_deg(nm["indexing"][0, 14:18], s.q[0])
    print(f"start {err_start:.3e} deg, adjoint {err_adj:.3e} deg, Nelder-Mead {err_nm:.3e} deg, "
          f"{info['n_pairs']} pairs, {info['lm_iterations']} LM iterations, rms {info['rms_px']:.2e} px", flush=True)
    assert err_start > 0.001
    assert info["n_pairs"] >= 20 and info["rms_px"] < 2e-5
    assert err_adj < 1e-4
    assert err_adj <= err_nm
    assert row[6] > 0.95 and row[6] >= nm["indexing"][0, 6]
    np.testing.assert_array_equal(row[1:4], pos)
    assert cand == float(nm["candidate"][0])


@pytest.mark.gpu
def test_adjoint_centre_fit_recovers_the_true_centre(engine):
    """fit_grain_COM from spot pairs made 2 voxels away from grain 1's
    centre with its true orientation: the LM over [Rodrigues vector, centre]
    returns the true centre and orientation; the Nelder-Mead of the mean
    rounded pixel distance stops within the rounding's dead zone."""
    from grainrecon_amd.engine import euler_to_rod_quat
    from grainrecon_amd.geometry import quaternion2U
    from grainrecon_amd.mapping import fit_grain_com_adjoint
    from test_index import _index_problem

    s, _, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_exact_frames(engine, s, [1], 400.0))
    start = s.centers[1] + np.array([0.01, -0.006, 0.004])
    U = quaternion2U(s.q[1])
    _, _, _, pr = engine.pair_spots(start[None, :], U[None, :, :])
    pairs = pr[0]
    print(f"fit_grain_com_adjoint on {pairs.shape[0]} pairs", flush=True)
    fo, lm = fit_grain_com_adjoint(engine, _geo7(s.geo), pairs, s.q[1][1:] / s.q[1][0])
    fo_nm = engine.fit_grain_com([pairs])[0][0]
    _, q_adj, _ = euler_to_rod_quat(fo[3:][None, :])
    err_pos = float(np.abs(fo[:3] - s.centers[1]).max())
    err_pos_nm = float(np.abs(fo_nm[:3] - s.centers[1]).max())
    print(f"centre error: adjoint {err_pos:.3e} mm, Nelder-Mead {err_pos_nm:.3e} mm; LM {lm.iterations} "
          f"iterations, rms {lm.rms:.2e} px", flush=True)
    assert pairs.shape[0] >= 20
    assert err_pos < 1e-6 and lm.rms < 2e-5
    assert _misorientation_deg(q_adj[0], s.q[1]) < 1e-4
    assert err_pos <= err_pos_nm


@pytest.mark.gpu
def test_mapping_loop_adjoint_mode_maps_like_nelder_mead(engine):
    """The mapping loop of the end-to-end test in both modes on exact
    centroids of all three grains: the same seeds index the same grains, the
    adjoint rows' orientations are the true ones, and the grown GrainId maps
    agree on nearly every voxel."""
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine
    from test_index import _index_problem

    s, qs, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_exact_frames(engine, s, [0, 1, 2], 30.0))
    n, vox = 14, 0.005
    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=2.0, max_dmedian=20.0, nr_seed=6, first_grain_id=1,
              center=s.centers[0], vox_size=vox, rvp_first=[1, 1, 1], tomo_dim=(n, n, n))
    pos = np.array([s.centers[0], s.centers[0] + [0.005, 0.0, 0.0]])
    out = {}
    for mode in ("nelder-mead", "adjoint"):
        ds = MappingDS(np.ones((n, n, n)))
        log = []
        print(f"grain_mapping_engine refine={mode}: 2 seeds on a 14^3 volume", flush=True)
        o, _, nvox, _ = grain_mapping_engine(engine, pos, qs, ds, refine=mode, geo7=_geo7(s.geo), fit_log=log, **kw)
        out[mode] = (o, nvox, ds, log)
    o_nm, n_nm, d_nm, log_nm = out["nelder-mead"]
    o_ad, n_ad, d_ad, log_ad = out["adjoint"]
    assert log_nm == []
    assert [e["site"] for e in log_ad][:2] == ["fit_OR_constrain", "fit_grain_COM"]
    assert o_ad[0, 6] > 0.95 and n_ad[0] > 1 and d_ad.GrainId[6, 6, 6] == 1
    err_ad = _misorientation_deg(o_ad[0, 14:18], s.q[0])
    err_nm = _misorientation_deg(o_nm[0, 14:18], s.q[0])
    agree = np.mean((d_nm.GrainId > 0) == (d_ad.GrainId > 0))
    print(f"first seed at {o_ad[0, 1:4]} (centre {s.centers[0]}): orientation error adjoint {err_ad:.3e} deg, "
          f"Nelder-Mead {err_nm:.3e} deg; voxels: Nelder-Mead {n_nm.tolist()}, adjoint {n_ad.tolist()}, "
          f"occupancy agreement {agree:.4f}", flush=True)
    # the loop re-indexes at the grown region's centre, which need not be the
    # grain's centre voxel: the orientation is then fitted at that position
    assert err_ad < 0.05 and err_ad <= err_nm + 0.01
    assert agree >= 0.9
