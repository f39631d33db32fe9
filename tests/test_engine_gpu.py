"""GPU parity: the gfx950 engine against the CPU oracle, bit for bit.

Integer counts must be identical; completeness (a ratio of the same
integers) and the single-precision median must be bitwise identical.
"""
import math

import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd import engine as E
from grainrecon_amd import _lib

pytestmark = pytest.mark.gpu

METHODS = [_lib.METHOD_FAST, _lib.METHOD_PLAIN]


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def test_math_host_equals_device(engine):
    rng = np.random.default_rng(0)
    cases = {
        0: np.concatenate([rng.uniform(-7, 7, 20000), np.linspace(0, 2 * np.pi, 721)]),
        1: np.concatenate([rng.uniform(-7, 7, 20000), np.linspace(0, 2 * np.pi, 721)]),
        2: np.concatenate([rng.uniform(-1, 1, 20000), 1 - rng.random(2000) ** 8, -1 + rng.random(2000) ** 8]),
        3: rng.uniform(-10, 10, 20000),
        4: rng.uniform(-5000, 5000, 20000).round(1) + rng.choice([0.0, 0.5, -0.5], 20000),
        7: np.concatenate([rng.uniform(-1, 1, 20000), 1 - rng.random(2000) ** 8, -1 + rng.random(2000) ** 8]),
        8: np.concatenate([rng.uniform(-1, 1, 20000), 1 - rng.random(2000) ** 8, -1 + rng.random(2000) ** 8]),
        9: np.concatenate([rng.uniform(-np.pi, np.pi, 20000), np.arange(-8, 9) * (np.pi / 8)]),
    }
    for fn, x in cases.items():
        assert np.array_equal(_bits(engine.math_device(fn, x)), _bits(oracle.math_eval(fn, x))), fn
    x = np.abs(rng.normal(size=50000)) * 10.0 ** rng.integers(-20, 20, 50000)
    assert np.array_equal(_bits(engine.math_device(5, x)), _bits(np.sqrt(x)))
    assert np.array_equal(_bits(engine.math_device(6, x)), _bits(1.0 / x))


def test_omega_solve_math_error_bounds(engine):
    """The omega solve's own acos and atan2 (the filter-grade polynomials with
    refined hardware reciprocals) keep the filter-grade bounds, 5e-12 and
    3e-12 rad, and NaN outside [-1, 1]; the refined reciprocal and reciprocal
    square root that the conservative filters use are within 5e-15 relative
    (their margins are 1e-12 or wider)."""
    rng = np.random.default_rng(21)
    xa = np.concatenate([np.linspace(-1.0, 1.0, 200001), rng.uniform(-1, 1, 20000), 1 - rng.random(2000) ** 8,
                         -1 + rng.random(2000) ** 8, [-1.0, -0.5, 0.0, 0.5, 1.0]])
    xa = np.concatenate([xa, np.nextafter(0.5, 1.0) * np.array([1.0, -1.0]), np.nextafter(0.5, 0.0) * np.array([1.0, -1.0])])
    ya = engine.math_device(10, xa)
    assert np.max(np.abs(ya - np.arccos(xa))) < 5e-12
    assert np.isnan(engine.math_device(10, np.array([1.0000000000000002, -1.5, np.nan]))).all()
    special = np.arange(-8, 9) * (np.pi / 8)
    xt = np.concatenate([np.linspace(-np.pi, np.pi, 200001), rng.uniform(-np.pi, np.pi, 20000), special,
                         (special[:, None] + np.array([-1e-9, 1e-9, -1e-15, 1e-15])).ravel()])
    xt = xt[(xt > -np.pi) & (xt <= np.pi)]
    yt = engine.math_device(11, xt)
    assert np.max(np.abs((yt - xt + np.pi) % (2 * np.pi) - np.pi)) < 3e-12
    x = (1.0 + rng.random(100000)) * 10.0 ** rng.integers(-30, 30, 100000) * rng.choice([-1.0, 1.0], 100000)
    assert np.max(np.abs(engine.math_device(12, x) * x - 1.0)) < 5e-15
    xp = np.abs(x)
    r = engine.math_device(13, xp)
    assert np.max(np.abs(r * np.sqrt(xp) - 1.0)) < 5e-15


def _variant_scene(name):
    kw = dict(seed=31, ngrains=3)
    if name == "tilted":
        kw.update(tilts=(0.3, -0.2, 0.25), dety00=0.04, detz00=-0.03)
    if name == "beamstop":
        kw.update(beamstop=((110.0, 150.0), (100.0, 160.0)))
    if name == "off_centre":
        # detector well off the beam and tilted: the corner rays and slopes
        # of the candidate bounds are far from symmetric
        kw.update(tilts=(2.0, -1.5, 3.0), dety00=1.5, detz00=-1.0)
    s = scenes.make_scene(**kw)
    if name == "offsets":
        s.geo.P0y, s.geo.P0z, s.geo.RotAxisOffset = 0.02, -0.015, 0.01
    if name == "beam_offsets":
        # beam offsets far above the 2-pixel margin of the detector corners: the
        # ray K_out of forward_comp.m:49 is then visibly off the 2 theta cone
        # around the beam direction
        s.geo.P0y, s.geo.P0z, s.geo.RotAxisOffset = 0.3, -0.3, 0.05
    if name == "irregular_rot":
        # rotation angles not on a uniform grid: no omega solve, every event
        # goes through the phase-A filter
        s.rot = s.rot + np.random.default_rng(9).uniform(-1.5, 1.5, s.rot.size)
    if name == "sample_tilt":
        # a tilted sample frame with offsets: the beam direction at a voxel then
        # depends on every component of S*pos
        a, b = np.radians(2.0), np.radians(-1.5)
        Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
        Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
        s.geo.S = Rx @ Ry
        s.geo.P0y, s.geo.P0z, s.geo.RotAxisOffset = -0.03, 0.02, 0.015
    return s


def _candidate_q(s, rng, n_random=300):
    """Truth orientations, local perturbations around them (partial C) and random ones."""
    rows = [s.q]
    for g in range(s.q.shape[0]):
        d = rng.normal(scale=0.004, size=(60, 3))
        qv = s.q[g, 1:] + d
        qv = qv[(qv ** 2).sum(1) < 1]
        rows.append(np.hstack([np.sqrt(1 - (qv ** 2).sum(1))[:, None], qv]))
    rows.append(scenes.random_quats(rng, n_random))
    rows.append(np.array([[0.0, 0.9, 0.5, 0.3]]))  # |q_vec| > 1: skipped row
    return np.vstack(rows)


def _assert_same(got, ref):
    ns, ni, C, med = got
    assert np.array_equal(ns, ref[:, 0].astype(np.int32))
    assert np.array_equal(ni, ref[:, 1].astype(np.int32))
    assert np.array_equal(_bits(C), _bits(ref[:, 2]))
    assert np.array_equal(_bits(med), _bits(ref[:, 3].astype(np.float32)))


@pytest.mark.parametrize("variant", ["plain", "tilted", "off_centre", "offsets", "beam_offsets", "sample_tilt", "irregular_rot", "beamstop"])
@pytest.mark.parametrize("method", METHODS, ids=["fast", "plain"])
def test_forward_comp_parity(engine, variant, method):
    s = _variant_scene(variant)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(5)
    q = _candidate_q(s, rng)
    for g, pos in ((0, s.centers[0]), (1, s.voxels[1][5])):
        ref = oracle.or_space_comp(s.geo, s.stack, pos, q, s.rot, s.hkl)
        got = engine.forward_comp(pos, q, method)
        _assert_same(got, ref)
        if variant not in ("offsets", "beam_offsets", "sample_tilt", "irregular_rot"):  # rendered before the change
            assert got[2][g] == 1.0  # the grain's own orientation at one of its voxels
        assert got[0][-1] == 0 and got[2][-1] == 0.0


def test_fast_and_plain_agree_and_are_deterministic(engine):
    s = _variant_scene("plain")
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    q = _candidate_q(s, np.random.default_rng(6), n_random=2000)
    a = engine.forward_comp(s.centers[2], q, _lib.METHOD_FAST)
    b = engine.forward_comp(s.centers[2], q, _lib.METHOD_FAST)
    c = engine.forward_comp(s.centers[2], q, _lib.METHOD_PLAIN)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(_bits(x) if x.dtype.kind == "f" else x, _bits(y) if y.dtype.kind == "f" else y)
        assert np.array_equal(_bits(x) if x.dtype.kind == "f" else x, _bits(z) if z.dtype.kind == "f" else z)


def _p1(n=40, vox=0.005, seed_comp=0.9, drop=0.02, maxdm=5.0):
    return np.array([1, 1, 1, 1, vox, 0, 0, 0, n, n, n, seed_comp, drop, maxdm], dtype=float)


@pytest.mark.parametrize("method", METHODS, ids=["fast", "plain"])
def test_grow_parity(engine, method):
    s = scenes.make_scene(seed=41, ngrains=2, spread_mm=0.02)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(3)
    idx = np.stack(np.meshgrid(np.arange(14, 27), np.arange(14, 27), np.arange(16, 25), indexing="ij"), -1)
    idx = idx.reshape(-1, 3).astype(float)
    vox = np.hstack([idx, rng.uniform(0, 0.6, (idx.shape[0], 1)), (rng.random((idx.shape[0], 1)) < 0.9) * 1.0,
                     rng.uniform(0, 30, (idx.shape[0], 1))])
    p1 = _p1()
    for g in range(2):
        ref = oracle.grow(s.geo, s.stack, vox, s.U[g], s.rot, s.hkl, p1)
        got = engine.grow_indexed_region(vox, s.U[g], p1, method)
        assert np.array_equal(_bits(got), _bits(ref))
        assert got[:, 6].sum() > 0


@pytest.mark.parametrize("method", METHODS, ids=["fast", "plain"])
def test_compete_parity(engine, method):
    s = scenes.make_scene(seed=43, ngrains=3, spread_mm=0.03)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(4)
    G = 6
    eu = np.vstack([rng.uniform([0, 0, 0], [360, 180, 360], (G, 3))])
    for g in range(3):  # three of the candidates are the true grains
        U = s.U[g]
        eu[g] = [math.degrees(math.atan2(U[0, 2], -U[1, 2])) % 360, math.degrees(math.acos(U[2, 2])),
                 math.degrees(math.atan2(U[2, 0], U[2, 1])) % 360]
    idx = np.stack(np.meshgrid(np.arange(14, 27, 2), np.arange(14, 27, 2), np.arange(16, 25, 2), indexing="ij"), -1)
    idx = idx.reshape(-1, 3).astype(float)
    ids, rows = [], []
    for v in idx:
        k = int(rng.integers(0, 5))
        cand = list(rng.integers(1, G + 1, size=k))
        start = len(ids) + 1
        ids += cand
        rows.append([*v, start, start + k - 1])
    rows.append([10 ** 5, 20, 20, 1, 3])  # far away: every C is NaN
    vox, ids = np.array(rows, float), np.array(ids, float)
    p1 = _p1()[:11]
    ref = oracle.compete(s.geo, s.stack, vox, ids, eu, s.rot, s.hkl, p1)
    got = engine.forward_comp_compete(vox, ids, eu, p1, method)
    assert np.array_equal(_bits(got), _bits(ref))
    assert np.isnan(got[-1, 5])


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def _legacy_inputs(s):
    geo = s.geo
    param = geo.legacy_param(s.hkl.shape[0], s.rot.size).astype(np.float32)
    # the oracle sees exactly what the engine sees: float32 values promoted to double
    g = type(geo)(**{k: getattr(geo, k) for k in geo.__dataclass_fields__})
    g.S, g.B, g.RotDet = _f32(geo.S), _f32(geo.B), _f32(geo.RotDet)
    for k, j in (("Lsam2sou", 2), ("Lsam2det", 3), ("P0y", 4), ("P0z", 5), ("dety00", 6), ("detz00", 7),
                 ("pixelysize", 8), ("pixelzsize", 9), ("detysize", 10), ("detzsize", 11), ("thetamax", 16),
                 ("lambda_min", 17), ("lambda_max", 18), ("RotAxisOffset", 19)):
        setattr(g, k, float(param[j]))
    g.BeamStopY = (float(param[12]), float(param[13]))
    g.BeamStopZ = (float(param[14]), float(param[15]))
    g.dety0 = g.detysize / 2 + 0.5
    g.detz0 = g.detzsize / 2 + 0.5
    g.minEucDis = 0.5 * ((g.pixelysize + g.pixelzsize) / 2)
    return g, param


def test_legacy_forward_comp_matches_oracle_on_rounded_inputs():
    s = scenes.make_scene(seed=51, ngrains=2)
    g, param = _legacy_inputs(s)
    q = _candidate_q(s, np.random.default_rng(8), n_random=200)
    pos = s.centers[1]
    # MATLAB passes matrices as reshape(X',1,9), i.e. numpy's row-major order
    out = E.cuda_forward_comp(pos, q.reshape(-1), np.asarray(s.geo.S).reshape(-1), np.asarray(s.geo.B).reshape(-1),
                              s.hkl.reshape(-1), E.stack_to_matlab_order(s.proj), s.rot,
                              np.asarray(s.geo.RotDet).reshape(-1), param)
    out = out.reshape(-1, 4)
    ref = oracle.or_space_comp(g, s.stack, _f32(pos), _f32(q), _f32(s.rot), _f32(s.hkl))
    assert np.array_equal(out[:, 0], ref[:, 0].astype(np.float32))
    assert np.array_equal(out[:, 1], ref[:, 1].astype(np.float32))
    assert np.array_equal(_bits(out[:, 2]), _bits(ref[:, 2].astype(np.float32)))
    assert np.array_equal(_bits(out[:, 3]), _bits(ref[:, 3].astype(np.float32)))


def test_legacy_grow_and_compete_match_oracle():
    s = scenes.make_scene(seed=52, ngrains=2, spread_mm=0.02)
    g, param = _legacy_inputs(s)
    idx = np.stack(np.meshgrid(np.arange(16, 25), np.arange(16, 25), np.arange(18, 23), indexing="ij"), -1)
    idx = idx.reshape(-1, 3).astype(float)
    vox = np.hstack([idx, np.full((idx.shape[0], 1), 0.25), np.ones((idx.shape[0], 1)), np.full((idx.shape[0], 1), 9.)])
    p1 = _p1()
    U = np.asarray(s.U[0], dtype=np.float32)
    out = E.cuda_grow_indexed_region(vox, U.reshape(-1), np.asarray(s.geo.S).reshape(-1),
                                     np.asarray(s.geo.B).reshape(-1), s.hkl.reshape(-1),
                                     E.stack_to_matlab_order(s.proj), s.rot, np.asarray(s.geo.RotDet).reshape(-1),
                                     param, p1).reshape(-1, 7)
    ref = oracle.grow(g, s.stack, _f32(vox), _f32(U), _f32(s.rot), _f32(s.hkl), _f32(p1))
    assert np.array_equal(_bits(out), _bits(ref.astype(np.float32)))
    eu = np.array([[10.0, 40.0, 70.0], [200.0, 100.0, 30.0]])
    ids = np.array([1, 2, 2, 1], float)
    vox5 = np.array([[20, 20, 20, 1, 2], [21, 20, 20, 3, 4], [22, 20, 20, 1, 4]], float)
    out = E.cuda_forward_comp_compete(vox5, ids, eu.reshape(-1), np.asarray(s.geo.S).reshape(-1),
                                      np.asarray(s.geo.B).reshape(-1), s.hkl.reshape(-1),
                                      E.stack_to_matlab_order(s.proj), s.rot, np.asarray(s.geo.RotDet).reshape(-1),
                                      param, p1[:11]).reshape(-1, 8)
    ref = oracle.compete(g, s.stack, _f32(vox5), _f32(ids), _f32(eu), _f32(s.rot), _f32(s.hkl), _f32(p1[:11]))
    assert np.array_equal(_bits(out), _bits(ref.astype(np.float32)))


def test_errors_are_reported_not_raised_from_c(engine):
    s = scenes.make_scene(seed=61, ngrains=1, det=128, pixel_mm=6.72e-3 * 8)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    with pytest.raises(_lib.GrsError):
        engine.set_distance_stack(np.full((64, 64, s.rot.size), 1.0, np.float32))
        engine.forward_comp(s.centers[0], s.q)  # stack smaller than the detector
    with pytest.raises(_lib.GrsError):
        engine.set_distance_stack(np.full((128, 128, s.rot.size), -1.0, np.float32))  # negative distances
    engine.set_distance_stack(s.proj)
    ns, ni, C, med = engine.forward_comp(s.centers[0], np.zeros((0, 4)))
    assert ns.size == 0
    out = engine.forward_comp_compete(np.array([[10, 10, 10, 1, 0]], float), np.zeros(0), np.zeros((0, 3)),
                                      _p1()[:11])
    assert out[0, 0] == 1 and (out[0, 1:] == 0).all()
    with pytest.raises(_lib.GrsError):
        engine.forward_comp_compete(np.array([[10, 10, 10, 1, 2]], float), np.array([1.0]), np.zeros((1, 3)),
                                    _p1()[:11])


def test_launch_configuration_errors_and_reset(engine):
    """grs_set_launch_config rejects configurations outside the tuner's set
    (waves per block other than 1, 2, 4, 8; a mask source for kind 0; an unknown
    source or mapping; an unknown kind) and, with 0 waves per block, forgets a
    pinned configuration so that grs_get_tuning reports the kind untuned."""
    from grainrecon_amd import _lib as L

    s = scenes.make_scene(seed=62, ngrains=1, det=128, pixel_mm=6.72e-3 * 8)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    for args in ((0, 3, L.MASK_BITS, False), (0, 4, L.MASK_KEYS, False), (1, 4, 3, False), (3, 4, L.MASK_BITS, False)):
        with pytest.raises(_lib.GrsError):
            engine.set_launch_config(*args)
    engine.set_launch_config(1, 2, L.MASK_TILED, True)
    t = engine.tuning(1)
    assert t == {"waves_per_block": 2, "mask_layout": "bits_tiled32", "xcd_map": True, "ms_per_eval": []}
    engine.set_launch_config(1, 0)
    assert engine.tuning(1) is None


def _expand_local(qc, nr, g):
    """The engine's local-sampling grid (q1 slowest), same IEEE operations."""
    half = (g * nr) / 2.0
    out = []
    for k1 in range(nr + 1):
        for k2 in range(nr + 1):
            for k3 in range(nr + 1):
                out.append([0.0, (qc[1] - half) + float(k1) * g, (qc[2] - half) + float(k2) * g,
                            (qc[3] - half) + float(k3) * g])
    return np.array(out)


def _matlab_argmax(C):
    best = -1
    for j, c in enumerate(C):
        if c == c and (best < 0 or c > C[best]):
            best = j
    return 0 if best < 0 and len(C) else best


def test_compete_csr_matches_oracle_argmax(engine):
    s = scenes.make_scene(seed=81, ngrains=3, spread_mm=0.03)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(12)
    nr, g = 2, 0.002
    pos, rows, cands = [], [0], []
    for r in range(14):
        k = int(rng.integers(0, 5)) if r != 3 else 0  # row 3 is empty
        grain = r % 3
        for _ in range(k):
            if rng.random() < 0.5:
                qc = s.q[grain] + np.concatenate([[0.0], rng.normal(scale=0.002, size=3)])
            else:
                qc = scenes.random_quats(rng, 1)[0]
            cands.append(qc)
        rows.append(len(cands))
        pos.append(s.voxels[grain][int(rng.integers(0, 27))])
    pos, rows, cands = np.array(pos), np.array(rows), np.array(cands)
    best, bq, ns, ni, C, med = engine.compete_csr(pos, rows, cands, nr, g)
    for r in range(len(pos)):
        if rows[r + 1] == rows[r]:
            assert best[r] == -1
            continue
        qs = np.vstack([_expand_local(cands[c], nr, g) for c in range(rows[r], rows[r + 1])])
        ref = oracle.or_space_comp(s.geo, s.stack, pos[r], qs, s.rot, s.hkl)
        j = _matlab_argmax(ref[:, 2])
        assert best[r] == j, r
        assert ns[r] == ref[j, 0] and ni[r] == ref[j, 1]
        assert np.array_equal(np.array([C[r]]).view(np.uint64), ref[j:j + 1, 2].view(np.uint64))
        assert np.float32(ref[j, 3]).view(np.uint32) == med[r].view(np.uint32)
        assert np.array_equal(bq[r, 1:], qs[j, 1:])
    assert (C[rows[1:] > rows[:-1]] > 0).any()


@pytest.mark.parametrize("variant", ["plain", "tilted_offsets"])
def test_resolution_stage_agrees_with_exact_pipeline(engine, variant):
    """Every event the fast kernel decides without the exact pipeline is re-run
    through it on the device: no disagreement, coordinates far inside the
    1e-4 pixel resolution margin, and only a small share sent to the exact path."""
    kw = {} if variant == "plain" else dict(tilts=(0.25, -0.15, 0.3), dety00=0.03, detz00=-0.02)
    s = scenes.make_scene(seed=61, ngrains=3, spread_mm=0.05, **kw)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(7)
    q = np.vstack([s.q, scenes.random_quats(rng, 4000)])
    ref = engine.forward_comp(s.voxels[0][13], q)
    engine.set_audit(True)
    try:
        out = engine.forward_comp(s.voxels[0][13], q)
        au = engine.audit()
    finally:
        engine.set_audit(False)
    for x, y in zip(out, ref):
        np.testing.assert_array_equal(x, y)
    assert au["examined"] == au["resolved_selected"] + au["resolved_rejected"] + au["exact"]
    assert au["resolved_selected"] > 10000
    assert au["mismatched"] == 0
    assert au["max_dev_px"] < 1e-6
    assert au["exact"] <= 0.01 * au["examined"]


def test_compete_csr_graph_replays_match_direct_launches(engine):
    """The batched compete runs as a captured HIP graph: replays with new data
    of the same shape, interleaved with other work, equal direct launches."""
    s = scenes.make_scene(seed=83, ngrains=3, spread_mm=0.03)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(14)
    nr, g = 2, 0.002
    rows = np.array([0, 3, 3, 5, 9, 10, 13, 16])   # fixed shape: 7 rows, 16 candidates (one empty row)
    datasets = []
    for _ in range(3):
        pos = np.array([s.voxels[r % 3][int(rng.integers(0, 27))] for r in range(7)])
        cands = np.array([s.q[c % 3] + np.concatenate([[0.0], rng.normal(scale=0.002, size=3)]) for c in range(16)])
        datasets.append((pos, cands))
    engine.set_graphs(False)
    try:
        ref = [engine.compete_csr(p, rows, c, nr, g) for p, c in datasets]
    finally:
        engine.set_graphs(True)
    cap0, rep0 = engine.graph_stats()
    for rnd in range(3):
        for k, (p, c) in enumerate(datasets):
            out = engine.compete_csr(p, rows, c, nr, g)
            for x, y in zip(out, ref[k]):
                np.testing.assert_array_equal(x, y)
            engine.forward_comp(s.voxels[0][0], s.q[:2])   # eager work between replays
            engine.synchronize()
    cap1, rep1 = engine.graph_stats()
    assert rep1 - rep0 >= 8 and cap1 - cap0 <= 1


def test_autotuned_launches_give_identical_results(engine):
    """The first large call of each kernel kind times its launch configurations
    on parts of the work; the outputs equal those of the untuned launch."""
    s = scenes.make_scene(seed=85, ngrains=3, spread_mm=0.03, nrot=19)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(21)
    q = np.vstack([s.q, scenes.random_quats(rng, 330000)])
    pos = s.voxels[1][13]
    rows = np.arange(0, 1921, 4)                   # 480 rows x 4 candidates x 729 samples
    cands = np.array([s.q[c % 3] + np.concatenate([[0.0], rng.normal(scale=0.004, size=3)]) for c in range(1920)])
    cpos = np.array([s.voxels[r % 3][int(rng.integers(0, 27))] for r in range(480)])
    engine.set_autotune(False)
    try:
        ref = engine.forward_comp(pos, q)
        ref_c = engine.compete_csr(cpos, rows, cands, 8, 0.002)
    finally:
        engine.set_autotune(True)
    out = engine.forward_comp(pos, q)
    out_c = engine.compete_csr(cpos, rows, cands, 8, 0.002)
    for x, y in zip(out, ref):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(out_c, ref_c):
        np.testing.assert_array_equal(x, y)
    t0, t1 = engine.tuning(0), engine.tuning(1)
    # kind 0: waves per block x XCD mapping; kind 1 also x bit mask / keys / tiled bit mask
    assert t0 is not None and t0["waves_per_block"] in (1, 2, 4, 8) and len(t0["ms_per_eval"]) == 8
    assert t0["mask_layout"] == "bits"
    assert t1 is not None and len(t1["ms_per_eval"]) == 24 and all(v > 0 for v in t1["ms_per_eval"])
    assert t1["mask_layout"] in ("bits", "keys", "bits_tiled32")
    # tuned launches afterwards, still identical
    for x, y in zip(engine.compete_csr(cpos, rows, cands, 8, 0.002), ref_c):
        np.testing.assert_array_equal(x, y)


def test_every_pinned_launch_configuration_gives_identical_results(engine):
    """Each launch configuration the autotuner can choose, pinned in turn:
    waves per block x XCD mapping for the median kernel, and also x bit mask /
    distance keys / tiled bit mask for the counts-only CSR and grid kernels;
    every one equals the default configuration bit for bit (forward_comp,
    compete_csr and index_seeds on a 256^2 x 19 scene, including a stack whose
    sides are not multiples of the 32-pixel tile)."""
    from grainrecon_amd import _lib as L

    for det in (256, 250):
        s = scenes.make_scene(seed=86, ngrains=3, spread_mm=0.03, nrot=19, det=det)
        engine.set_geometry(s.geo, s.hkl, s.rot)
        engine.set_distance_stack(s.proj)
        rng = np.random.default_rng(22)
        q = np.vstack([s.q, scenes.random_quats(rng, 3000)])
        q = q[q[:, 0] >= 0.5][:2000]
        pos = s.voxels[1][13]
        rows = np.arange(0, 41, 4)
        cands = np.array([s.q[c % 3] + np.concatenate([[0.0], rng.normal(scale=0.004, size=3)]) for c in range(40)])
        cpos = np.array([s.voxels[r % 3][int(rng.integers(0, 27))] for r in range(10)])
        engine.set_orientation_set(q)
        seeds = np.array([s.centers[0], s.voxels[2][4]])
        try:
            for k in range(3):
                engine.set_launch_config(k, 4, L.MASK_BITS, False)
            ref = engine.forward_comp(pos, q)
            ref_c = engine.compete_csr(cpos, rows, cands, 8, 0.002)
            ref_i = engine.index_seeds(seeds, nr_seed=6)
            print(f"det {det}: default configuration done", flush=True)
            for xcd in (False, True):
                for wpb in (8, 4, 2, 1):
                    engine.set_launch_config(0, wpb, L.MASK_BITS, xcd)
                    for x, y in zip(engine.forward_comp(pos, q), ref):
                        np.testing.assert_array_equal(x, y)
                    for mask in (L.MASK_BITS, L.MASK_KEYS, L.MASK_TILED):
                        engine.set_launch_config(1, wpb, mask, xcd)
                        engine.set_launch_config(2, wpb, mask, xcd)
                        for x, y in zip(engine.compete_csr(cpos, rows, cands, 8, 0.002), ref_c):
                            np.testing.assert_array_equal(x, y)
                        got = engine.index_seeds(seeds, nr_seed=6)
                        for key in ("indexing", "candidate", "seed_idx", "seed_C"):
                            np.testing.assert_array_equal(got[key], ref_i[key], err_msg=f"{key} {wpb} {mask} {xcd}")
                    print(f"det {det}: wpb {wpb} xcd {xcd} equal", flush=True)
        finally:
            for k in range(3):
                engine.set_launch_config(k, 0)


def test_small_launches_take_smaller_blocks(engine):
    """With 8 waves per block pinned, launches of fewer evaluations than a
    block has waves (1 to 9 orientations: 4 or 12 scratch slots) run in
    smaller blocks; every result equals the oracle, for the one-voxel median
    kernel, its allocation-free entry and the counts-only CSR kernel."""
    from grainrecon_amd import _lib as L

    s = scenes.make_scene(seed=87, ngrains=2, spread_mm=0.03, nrot=19, det=128, pixel_mm=6.72e-3 * 8)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(23)
    q_all = np.vstack([s.q, scenes.random_quats(rng, 16)])
    pos = s.voxels[0][3]
    try:
        for k in range(3):
            engine.set_launch_config(k, 8, L.MASK_BITS, False)
        for n in (1, 3, 5, 8, 9):
            q = q_all[:n]
            print(f"[small] {n} orientations", flush=True)
            ref = oracle.or_space_comp(s.geo, s.stack, pos, q, s.rot, s.hkl)
            _assert_same(engine.forward_comp(pos, q), ref)
            d_q = engine.to_device(q)
            bufs = [engine.device_alloc(4 * n), engine.device_alloc(4 * n), engine.device_alloc(8 * n),
                    engine.device_alloc(4 * n)]
            try:
                engine.forward_comp_device_noalloc(pos, d_q, n, *bufs)
                got = (engine.to_host(bufs[0], (n,), np.int32), engine.to_host(bufs[1], (n,), np.int32),
                       engine.to_host(bufs[2], (n,), np.float64), engine.to_host(bufs[3], (n,), np.float32))
                _assert_same(got, ref)
            finally:
                for b in bufs + [d_q]:
                    engine.device_free(b)
        rows = np.array([0, 1, 3])
        cands = np.array([s.q[c % 2] + np.concatenate([[0.0], rng.normal(scale=0.004, size=3)]) for c in range(3)])
        cpos = np.array([s.voxels[0][2], s.voxels[1][4]])
        got_c = engine.compete_csr(cpos, rows, cands, 2, 0.002)
        for k in range(3):
            engine.set_launch_config(k, 4, L.MASK_BITS, False)
        for x, y in zip(got_c, engine.compete_csr(cpos, rows, cands, 2, 0.002)):
            np.testing.assert_array_equal(x, y)
    finally:
        for k in range(3):
            engine.set_launch_config(k, 0)


@pytest.mark.parametrize("variant", ["plain", "offsets"])
def test_grow_many_voxels_equals_chunks_and_oracle(engine, variant):
    """Above 4096 voxels the omega-solve bounds come from one thread per voxel
    instead of one block per voxel: growth over 4,800 voxels equals the same
    voxels scored in chunks of 1,000 bit for bit, and a sample of them the
    oracle."""
    s = scenes.make_scene(seed=41, ngrains=2, spread_mm=0.02)
    if variant == "offsets":
        s.geo.P0y, s.geo.P0z, s.geo.RotAxisOffset = 0.02, -0.015, 0.01
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(8)
    idx = np.stack(np.meshgrid(np.arange(10, 30), np.arange(10, 30), np.arange(14, 26), indexing="ij"), -1)
    idx = idx.reshape(-1, 3).astype(float)
    vox = np.hstack([idx, rng.uniform(0, 0.6, (idx.shape[0], 1)), (rng.random((idx.shape[0], 1)) < 0.9) * 1.0,
                     rng.uniform(0, 30, (idx.shape[0], 1))])
    assert vox.shape[0] == 4800
    p1 = _p1()
    got = engine.grow_indexed_region(vox, s.U[0], p1)
    parts = np.vstack([engine.grow_indexed_region(vox[k:k + 1000], s.U[0], p1) for k in range(0, 4800, 1000)])
    assert np.array_equal(_bits(got), _bits(parts))
    assert got[:, 6].sum() > 0 and got[:, 3].sum() > 0
    sel = np.concatenate([np.flatnonzero(got[:, 6] > 0)[:40], rng.choice(4800, size=120, replace=False)])
    ref = oracle.grow(s.geo, s.stack, vox[sel], s.U[0], s.rot, s.hkl, p1)
    assert np.array_equal(_bits(got[sel]), _bits(ref))
