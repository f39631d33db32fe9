"""G-vector matching score (cuda_forward_Gt has no source in the reference):
calc_Gt_match (ForwardCalc_OR_space_Gt.m:117-205, the CPU path) and
calc_Gt_match_fun.m.  The oracle is pinned by the forward geometry and by
identities between the two semantics; the GPU (grs_forward_gt) equals the
oracle bit for bit, with and without a rotation-axis offset."""
import copy
import math

import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd import _lib
from oracle import gt_match as GM


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)


def _families(s):
    sq = sorted({int((h * h).sum()) for h in s.hkl.astype(int)})[:4]
    a = 2.8665
    return np.array(sq, float), np.array([a / math.sqrt(v) for v in sq])


def test_sind_known_values():
    x = np.array([0.0, 30.0, 45.0, 90.0, 135.0, 180.0, 270.0, 360.0, -90.0, 12.5])
    y = GM.sind(x)
    assert y[0] == 0.0 and y[3] == 1.0 and y[5] == 0.0 and y[6] == -1.0 and y[7] == 0.0 and y[8] == -1.0
    np.testing.assert_allclose(y, np.sin(np.radians(x)), atol=5e-16)


def _exact_spots(s, U, pos):
    """One spot at the unrounded detector position of every selected event."""
    _, _, _, ev = oracle.forward_comp(s.geo, s.stack, U, pos, s.rot, s.hkl, events=True)
    frames = [[] for _ in range(s.rot.size)]
    for om, h, dy, dz, y22, z22, dis, lam in ev:
        frames[int(om)].append([(-y22) / s.geo.pixelysize + s.geo.dety0, (-z22) / s.geo.pixelzsize + s.geo.detz0,
                                10.0])
    return [np.array(f).reshape(-1, 3) for f in frames], ev


def test_gt_oracle_inverts_the_forward_geometry(scene):
    """Spots at the exact positions of an orientation's events give angles of
    nearly zero for its own reflections (the forward model's spot construction
    and the Laue condition agree to about 0.002 degrees), and the
    expected-spot count equals the forward_comp selection (no tilt, no
    rotation-axis offset)."""
    s = scene
    assert s.geo.RotAxisOffset == 0.0 and np.array_equal(np.asarray(s.geo.RotDet), np.eye(3))
    fam_sq, dposs = _families(s)
    frames, ev = _exact_spots(s, s.U[0], s.centers[0])
    q = s.q[0][None, :]
    out = GM.gt_score(s.geo, s.centers[0], q, s.rot, s.hkl, frames, fam_sq, dposs, semantics="fun")
    ns, ni, med = oracle.forward_comp(s.geo, s.stack, s.U[0], s.centers[0], s.rot, s.hkl)
    fam_of = {int(v) for v in fam_sq}
    own = sum(1 for e in ev if int((s.hkl[int(e[1])] ** 2).sum()) in fam_of)
    assert out[0, 0] >= own > 20
    assert out[0, 2] == out[0, 0] / ns
    assert out[0, 1] < 0.01 and out[0, 3] < 0.05
    # calc_Gt_match: the same matches; the cost of near-zero angles is near zero
    m = GM.gt_score(s.geo, s.centers[0], q, s.rot, s.hkl, frames, fam_sq, dposs, semantics="match")
    assert m[0, 0] == out[0, 0] and m[0, 2] == out[0, 2] and m[0, 3] == out[0, 3]
    assert 0.0 <= m[0, 1] < 2.0 * (1.0 - math.cos(math.radians(0.05)))
    # a rotation of 3 degrees moves every spot away: few matches
    qq = s.q[0].copy()
    qq[1:] += 0.026
    far = GM.gt_score(s.geo, s.centers[0], qq[None, :], s.rot, s.hkl, frames, fam_sq, dposs)
    assert far[0, 0] < 0.2 * out[0, 0]


def _frames(s, rng):
    frames = [[] for _ in range(s.rot.size)]
    for g in range(len(s.U)):
        _, _, _, ev = oracle.forward_comp(s.geo, s.stack, s.U[g], s.centers[g], s.rot, s.hkl, events=True)
        for om, _h, dy, dz, *_ in ev:
            frames[int(om)].append([dy + rng.uniform(-1, 1), dz + rng.uniform(-1, 1), 20.0])
    det = int(s.geo.detysize)
    for f in frames:
        for _ in range(int(rng.integers(0, 8))):
            f.append([rng.uniform(1, det), rng.uniform(1, det), 20.0])
    frames[5] = []
    return [np.array(f).reshape(-1, 3) for f in frames]


def _with_offset(geo, rao):
    g = copy.deepcopy(geo)
    g.RotAxisOffset = rao
    return g


def test_match_cost_is_the_mean_squared_unit_vector_distance(scene):
    """Both vectors of a match are unit vectors, so |n - w|^2 = 2 - 2 cos(angle):
    the cost of calc_Gt_match equals the mean of 2 - 2 cos over the matched
    angles (to rounding), which holds only for the spot of the smallest angle
    and the division by Nr_match; no match gives 100 (:203)."""
    s = scene
    rng = np.random.default_rng(41)
    frames = _frames(s, rng)
    fam_sq, dposs = _families(s)
    q = np.concatenate([s.q, s.q + rng.normal(scale=0.002, size=(3, 4)) * [0, 1, 1, 1]])
    out, angles = GM.gt_score(s.geo, s.centers[0], q, s.rot, s.hkl, frames, fam_sq, dposs, details=True)
    for u in range(q.shape[0]):
        a = angles[u]
        assert out[u, 0] == a.size
        if a.size:
            ref = float(np.mean(2.0 - 2.0 * np.cos(np.radians(a))))
            assert abs(out[u, 1] - ref) <= 1e-9 * ref + 1e-15, (u, out[u, 1], ref)
            assert out[u, 3] == pytest.approx(float(np.mean(a)), rel=1e-12)
    assert out[0, 0] > 20
    none = GM.gt_score(s.geo, s.centers[0], q[:1], s.rot, s.hkl, [np.zeros((0, 3))] * s.rot.size, fam_sq, dposs)
    assert none[0, 0] == 0 and none[0, 1] == 100.0 and math.isnan(none[0, 3])


def test_match_semantics_shift_the_position_by_the_axis_offset(scene):
    """calc_Gt_match at pos with RotAxisOffset r scores exactly as
    calc_Gt_match_fun at pos - [0 r 0] (ForwardCalc_OR_space_Gt.m:135 is the
    only difference besides the output column): a sign error in the shift
    breaks the equality."""
    s = scene
    rng = np.random.default_rng(43)
    frames = _frames(s, rng)
    fam_sq, dposs = _families(s)
    g = _with_offset(s.geo, 0.02)
    q = np.concatenate([s.q, scenes.random_quats(rng, 4)])
    pos = np.asarray(s.centers[0], float)
    m = GM.gt_score(g, pos, q, s.rot, s.hkl, frames, fam_sq, dposs, semantics="match")
    f = GM.gt_score(g, pos - [0.0, 0.02, 0.0], q, s.rot, s.hkl, frames, fam_sq, dposs, semantics="fun")
    f0 = GM.gt_score(g, pos, q, s.rot, s.hkl, frames, fam_sq, dposs, semantics="fun")
    assert np.array_equal(m[:, [0, 2]], f[:, [0, 2]])
    assert np.array_equal(m[:, 3], f[:, 3], equal_nan=True)
    assert not np.array_equal(m[:, 3], f0[:, 3], equal_nan=True)


@pytest.mark.gpu
@pytest.mark.parametrize("rao", [0.0, 0.02])
def test_forward_gt_equals_oracle(engine, scene, rao):
    s = scene
    geo = _with_offset(s.geo, rao)
    engine.set_geometry(geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(31)
    frames = _frames(s, rng)
    engine.set_frame_spots(frames)
    fam_sq, dposs = _families(s)
    q = np.concatenate([s.q, s.q + rng.normal(scale=0.003, size=(3, 4)) * [0, 1, 1, 1], scenes.random_quats(rng, 40),
                        [[0.1, 0.9, 0.5, 0.3]]])
    for sem, name in ((_lib.GT_MATCH, "match"), (_lib.GT_MATCH_FUN, "fun")):
        print(f"forward_gt {name}, RotAxisOffset {rao}: {q.shape[0]} orientations, "
              f"{sum(len(f) for f in frames)} spots", flush=True)
        got = engine.forward_gt(s.centers[0], q, fam_sq, dposs, sem)
        ref = GM.gt_score(geo, s.centers[0], q, s.rot, s.hkl, frames, fam_sq, dposs, semantics=name)
        same = (got == ref) | (np.isnan(got) & np.isnan(ref))
        assert same.all(), (name, np.argwhere(~same)[:5])
        assert got[0, 0] > 20 and np.all(got[-1] == 0.0)


@pytest.mark.gpu
def test_forward_gt_dense_spots_and_full_scan(engine, scene):
    """Spots crowded around every event (pixel offsets up to 60, i.e. angles
    spread across the 1-degree threshold; some spots duplicated, so the
    smallest angle is tied and the first spot must be taken) exercise the
    z-window search of the sorted spots; a projection with more spots than
    the sort holds (4200 extra) switches to the full scan.  Both equal the
    oracle in both semantics, with a rotation-axis offset."""
    s = scene
    geo = _with_offset(s.geo, -0.015)
    engine.set_geometry(geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(37)
    frames = _frames(s, rng)
    dense = []
    for f in frames:
        extra = [[y + rng.uniform(-60, 60), z + rng.uniform(-60, 60), 20.0] for y, z, _ in f for _ in range(6)]
        dup = [list(r) for r in f[::3]]
        dense.append(np.array(list(f) + extra + dup).reshape(-1, 3))
    fam_sq, dposs = _families(s)
    q = np.concatenate([s.q, s.q + rng.normal(scale=0.004, size=(3, 4)) * [0, 1, 1, 1], scenes.random_quats(rng, 10)])
    det = int(s.geo.detysize)
    crowded = list(dense)
    crowded[0] = np.vstack([dense[0], np.column_stack([rng.uniform(1, det, 4200), rng.uniform(1, det, 4200),
                                                       np.full(4200, 20.0)])])
    for name, fr in (("dense", dense), ("full scan", crowded)):
        engine.set_frame_spots(fr)
        for sem, sname in ((_lib.GT_MATCH, "match"), (_lib.GT_MATCH_FUN, "fun")):
            print(f"forward_gt {name} {sname}: {q.shape[0]} orientations, {sum(len(f) for f in fr)} spots, "
                  f"largest projection {max(len(f) for f in fr)}", flush=True)
            got = engine.forward_gt(s.centers[1], q, fam_sq, dposs, sem)
            ref = GM.gt_score(geo, s.centers[1], q, s.rot, s.hkl, fr, fam_sq, dposs, semantics=sname)
            same = (got == ref) | (np.isnan(got) & np.isnan(ref))
            assert same.all(), (name, sname, np.argwhere(~same)[:5])
            assert got[:3, 0].min() > 10
