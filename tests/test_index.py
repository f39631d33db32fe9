"""Seed indexing (GrainIndexing_engine_gpu_comp.m:46-66 with
fit_OR_constrain_with_local_comp_cuda.m): the oracle's pieces pinned on the
CPU, the orientation bookkeeping of the C ABI against the oracle, and the
device pipeline (grs_index_seeds) against the oracle on the GPU."""
import math

import numpy as np
import pytest

import oracle
import scenes
from oracle import index_seed as IS
from oracle.orient import atan2, fit_orientation, tan


# ---- pins of the oracle -----------------------------------------------------------

def test_atan2_and_tan_against_libm():
    rng = np.random.default_rng(0)
    ys = np.concatenate([rng.normal(size=400), [0.0, 1.0, -1.0, 0.0, 3.0, -3.0]])
    xs = np.concatenate([rng.normal(size=400), [1.0, 0.0, 0.0, -1.0, -3.0, -3.0]])
    for y, x in zip(ys, xs):
        ref = math.atan2(y, x)
        assert abs(atan2(y, x) - ref) <= math.ulp(ref), (y, x)
    for t in rng.uniform(-1.5, 1.5, 300):
        assert abs(tan(t) - math.tan(t)) <= 2 * math.ulp(math.tan(t))


def _U_of(q):
    return scenes.U_from_qvec(*q[1:])


def test_orientation_chain_invariants():
    """quaternion -> ZXZ Euler -> Rodrigues -> quaternion returns the fitted
    rotation: the quaternion up to sign, U equal to quaternion2U of the input
    and to euler2u of the reported angles (which pins the ZXZ convention)."""
    rng = np.random.default_rng(1)
    for x in rng.uniform(-0.55, 0.55, (60, 3)):
        out = fit_orientation(x)
        assert out is not None
        o10, U = out
        q_in = np.array([math.sqrt(1 - x @ x), *x])
        q_out = o10[6:]
        assert min(np.abs(q_out - q_in).max(), np.abs(q_out + q_in).max()) < 1e-12
        np.testing.assert_allclose(U, _U_of(q_in), atol=1e-12)
        np.testing.assert_allclose(oracle.euler2u_deg(*o10[:3]), U, atol=1e-12)
        assert np.all((o10[:3] >= 0) & (o10[:3] < 360))
        np.testing.assert_allclose(o10[3:6], q_out[1:] / q_out[0], rtol=1e-12, atol=1e-14)
    # x = 0: rod2quat of a zero vector is the zero quaternion, so U = -I
    o10, U = fit_orientation([0.0, 0.0, 0.0])
    assert np.all(o10 == 0.0) and np.array_equal(U, -np.eye(3))
    assert fit_orientation([0.9, 0.5, 0.3]) is None


def test_sortrows_descend_and_max():
    rng = np.random.default_rng(2)
    for _ in range(50):
        C = rng.integers(0, 5, 40).astype(float) / 4.0
        C[rng.random(40) < 0.2] = np.nan
        ref = sorted(range(40), key=lambda i: (0 if math.isnan(C[i]) else 1, 0 if math.isnan(C[i]) else -C[i], i))
        assert list(IS.sort_desc(C)) == ref
        v = C[:17]
        ok = [i for i in range(17) if not math.isnan(v[i])]
        want = 0 if not ok else min(i for i in ok if v[i] == max(v[j] for j in ok))
        assert IS.matlab_max_index(v) == want


def test_qall_layout_matches_matlab_assignment():
    """The q_all rows of fit_OR_constrain_with_local_comp_cuda.m:19-26
    simulated with MATLAB's grow-on-assignment semantics."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        K = int(rng.integers(1, 9))
        lens = [int(v) for v in rng.choice([5, 6, 7], K)]
        rows = [0] * (K * lens[0])      # zeros(len_1*K, 5), column 5 only
        for j in range(1, K + 1):
            lo, hi = (j - 1) * lens[j - 1], j * lens[j - 1]
            if hi > len(rows):
                rows += [0] * (hi - len(rows))
            rows[lo:hi] = [j] * (hi - lo)
        assert list(IS.qall_owner(lens)) == rows


# ---- C ABI (host) -------------------------------------------------------------------

def test_fit_orientations_capi_equals_oracle():
    from grainrecon_amd.engine import fit_orientations

    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-0.6, 0.6, (300, 3)), [[0.0, 0.0, 0.0], [0.9, 0.5, 0.3], [0.01, -0.02, 0.03]]])
    out, U, ok = fit_orientations(x)
    for m in range(x.shape[0]):
        r = fit_orientation(x[m])
        if r is None:
            assert ok[m] == 0 and np.all(np.isnan(out[m]))
            continue
        assert ok[m] == 1
        np.testing.assert_array_equal(out[m], r[0])
        np.testing.assert_array_equal(U[m], r[1])


# ---- device pipeline ----------------------------------------------------------------

def _index_problem():
    s = scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)
    rng = np.random.default_rng(5)
    # an orientation set inside q0 >= 0.5 (the cubic fundamental zone has
    # q0 >= 0.86): the local-sampling grids stay inside the unit ball
    qs = scenes.random_quats(rng, 6000)
    qs = qs[qs[:, 0] >= 0.5][:2000]
    for g in range(3):
        qq = s.q[g].copy()
        qq[1:] += rng.uniform(-0.01, 0.01, 3)
        qs[100 + 500 * g] = qq / np.linalg.norm(qq)
    qs[7] = [0.1, 0.9, 0.5, 0.3]      # |q_vec| > 1: a skipped orientation (C = 0)
    pos = np.array([s.centers[0], s.voxels[1][5], s.centers[2], [0.04, -0.03, 0.02]])
    return s, qs, pos


def _same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.gpu
@pytest.mark.parametrize("nr_seed,n_fit", [(8, 1), (6, 3)])
def test_index_seeds_equals_oracle(engine, nr_seed, n_fit):
    s, qs, pos = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_orientation_set(qs)
    print(f"index_seeds: {pos.shape[0]} positions, N={qs.shape[0]}, nr_seed={nr_seed}, n_fit={n_fit}", flush=True)
    got = engine.index_seeds(pos, nr_seed=nr_seed, n_fit=n_fit)
    for p in range(pos.shape[0]):
        ref = IS.index_seed(s.geo, s.stack, pos[p], qs, s.rot, s.hkl, nr_seed=nr_seed, n_fit=n_fit)
        assert list(got["seed_idx"][p]) == list(ref["seed_idx"]), p
        assert _same(got["seed_C"][p], ref["seed_C"]), p
        f = ref["fit"]
        assert got["candidate"][p] == ref["candidate"], p
        assert _same(got["x0"][p], f["x0"]) and _same(got["x"][p], f["x"]), p
        assert _same(got["fval"][p], f["fval"]) and got["n_fun_evals"][p] == f["nf"] and got["n_iter"][p] == f["it"]
        row = ref["indexing"].copy()
        row[0] = p + 1
        assert _same(got["indexing"][p], row), (p, got["indexing"][p], row)
    # the true orientations are found at the grain centres
    assert np.all(got["indexing"][[0, 2], 6] > 0.9)


@pytest.mark.gpu
def test_index_seeds_graph_replay(engine):
    """A second batch of the same shape replays the captured pipeline and
    gives the results of an eager run."""
    s, qs, pos = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_orientation_set(qs)
    engine.set_graphs(True)
    a = engine.index_seeds(pos[:2], nr_seed=5)
    c0, r0 = engine.graph_stats()
    b = engine.index_seeds(pos[2:], nr_seed=5)
    c = engine.index_seeds(pos[:2], nr_seed=5)
    c1, r1 = engine.graph_stats()
    assert r1 >= r0 + 2
    engine.set_graphs(False)
    try:
        d = engine.index_seeds(pos, nr_seed=5)
    finally:
        engine.set_graphs(True)
    # column 1 numbers the seeds of each call
    assert list(a["indexing"][:, 0]) == [1, 2] and list(d["indexing"][:, 0]) == [1, 2, 3, 4]
    for r in (a, b, c, d):
        r["indexing"] = r["indexing"][:, 1:]
    for k in ("indexing", "candidate", "x", "seed_idx", "seed_C"):
        assert _same(np.concatenate([a[k], b[k]]), d[k]), k
        assert _same(c[k], a[k]), k


@pytest.mark.gpu
def test_index_seeds_small_set_and_more_fits_than_seeds(engine):
    """Nr_seed above the set size keeps the whole set (the mapping engine's
    rule), and Nr_fit above Nr_seed fits every seed."""
    s, qs, pos = _index_problem()
    small = qs[[100, 600, 1100, 3, 4]]
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_orientation_set(small)
    got = engine.index_seeds(pos[:2], nr_seed=10, n_fit=7)
    for p in range(2):
        ref = IS.index_seed(s.geo, s.stack, pos[p], small, s.rot, s.hkl, nr_seed=10, n_fit=7)
        assert list(got["seed_idx"][p][:5]) == list(ref["seed_idx"]) and np.all(got["seed_idx"][p][5:] == -1)
        assert np.all(np.isnan(got["seed_C"][p][5:]))
        row = ref["indexing"].copy()
        row[0] = p + 1
        assert _same(got["indexing"][p], row) and got["candidate"][p] == ref["candidate"]
