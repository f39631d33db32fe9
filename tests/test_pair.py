"""Spot pairing (forward_comp_pair_spots.m:3-93): the oracle pinned on
constructed spot lists, the GPU (grs_pair_spots) against the oracle."""
import numpy as np
import pytest

import oracle
import scenes
from oracle.pair_spots import pair_spots


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)


def _frames_at_events(s, ev, offset=(0.0, 0.0), area=4.0):
    frames = [[] for _ in range(s.rot.size)]
    for om, _h, dy, dz, *_ in ev:
        frames[int(om)].append([dy + offset[0], dz + offset[1], area])
    return [np.array(f).reshape(-1, 3) for f in frames]


def test_oracle_pairing_known_answers(scene):
    s = scene
    U, p = s.U[0], s.centers[0]
    ns, ni, med, ev = oracle.forward_comp(s.geo, s.stack, U, p, s.rot, s.hkl, events=True)
    thr = s.geo.minEucDis / ((s.geo.pixelysize + s.geo.pixelzsize) / 2.0)
    hits = ev[ev[:, 6] <= thr]
    assert ni == hits.shape[0] > 20
    # a spot on every event pixel: every hit pairs at distance 0
    _, _, _, rows = pair_spots(s.geo, s.stack, U, p, s.rot, s.hkl, _frames_at_events(s, ev))
    assert rows.shape == (hits.shape[0], 23)
    np.testing.assert_array_equal(rows[:, 17:19], rows[:, 20:22])
    np.testing.assert_array_equal(rows[:, 20:22], hits[:, 2:4])
    np.testing.assert_array_equal(rows[:, 0:9], np.float32(U.flatten(order="F")).astype(float)[None, :].repeat(len(rows), 0))
    np.testing.assert_allclose(rows[:, 10], np.float32(p[1] - s.geo.RotAxisOffset))
    # Energy from the wavelength of each event (keV, 12.398/lambda in single precision)
    np.testing.assert_array_equal(rows[:, 15], np.float32(12.398 / hits[:, 7]).astype(float))
    # spots 5 px away: paired when sqrt(area) > 5, not when sqrt(area) = 4 or the distance reaches 15
    _, _, _, r36 = pair_spots(s.geo, s.stack, U, p, s.rot, s.hkl, _frames_at_events(s, ev, (3.0, 4.0), 36.0))
    _, _, _, r16 = pair_spots(s.geo, s.stack, U, p, s.rot, s.hkl, _frames_at_events(s, ev, (3.0, 4.0), 16.0))
    _, _, _, rfar = pair_spots(s.geo, s.stack, U, p, s.rot, s.hkl, _frames_at_events(s, ev, (9.0, 12.0), 900.0))
    assert r36.shape[0] == hits.shape[0] and r16.shape[0] == 0 and rfar.shape[0] == 0


def _frames(s, rng):
    """Spots near the events of every grain centre plus random ones."""
    frames = [[] for _ in range(s.rot.size)]
    for g in range(len(s.U)):
        _, _, _, ev = oracle.forward_comp(s.geo, s.stack, s.U[g], s.centers[g], s.rot, s.hkl, events=True)
        for om, _h, dy, dz, *_ in ev:
            frames[int(om)].append([dy + rng.uniform(-3, 3), dz + rng.uniform(-3, 3), rng.uniform(1, 60)])
    det = int(s.geo.detysize)
    for f in frames:
        for _ in range(int(rng.integers(0, 6))):
            f.append([rng.uniform(1, det), rng.uniform(1, det), rng.uniform(1, 400)])
    frames[3] = []  # a projection without spots
    return [np.array(f).reshape(-1, 3) for f in frames]


@pytest.mark.gpu
def test_pair_spots_equals_oracle(engine, scene):
    s = scene
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    rng = np.random.default_rng(11)
    frames = _frames(s, rng)
    engine.set_frame_spots(frames)
    pos = np.array([s.centers[0], s.centers[1], s.voxels[2][3], s.centers[2]])
    U = np.stack([s.U[0], s.U[1], s.U[2], scenes.U_from_qvec(*(s.q[2, 1:] + 0.004))])
    print(f"pair_spots: {pos.shape[0]} orientations, {sum(len(f) for f in frames)} spots", flush=True)
    ns, ni, med, rows = engine.pair_spots(pos, U)
    total = 0
    for m in range(pos.shape[0]):
        rns, rni, rmed, rrows = pair_spots(s.geo, s.stack, U[m], pos[m], s.rot, s.hkl, frames)
        assert (ns[m], ni[m], med[m]) == (rns, rni, np.float32(rmed)), m
        assert rows[m].shape == rrows.shape, m
        np.testing.assert_array_equal(rows[m], rrows)
        total += rrows.shape[0]
    assert total > 50
