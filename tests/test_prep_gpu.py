"""GPU input preparation: spot rendering and the exact distance transform."""
import numpy as np
import pytest
from scipy import ndimage

import oracle
import scenes

pytestmark = pytest.mark.gpu


def test_render_matches_oracle_rendering(engine):
    s = scenes.make_scene(seed=71, ngrains=3, dilate=0)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    pos = np.vstack(s.voxels)
    grain = np.repeat(np.arange(3), [len(v) for v in s.voxels])
    engine.render_spots(pos, grain, np.stack(s.U))
    got = engine.get_mask(s.mask.shape).astype(bool)
    assert got.sum() > 100
    assert np.array_equal(got, s.mask)


def test_edt_matches_scipy_bitwise(engine):
    rng = np.random.default_rng(3)
    s = scenes.make_scene(seed=72, ngrains=1, det=128, pixel_mm=6.72e-3 * 8, nrot=7)
    mask = rng.random((128, 96, 7)) < 0.01
    mask[:, :, 3] = False  # an empty projection
    mask[5, 7, 5] = True
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_binary_stack(mask)
    engine.distance_from_mask()
    d = engine.get_distance_stack(mask.shape)
    for i in range(mask.shape[2]):
        if mask[:, :, i].any():
            ref = ndimage.distance_transform_edt(~mask[:, :, i]).astype(np.float32)
            assert np.array_equal(d[:, :, i].view(np.uint32), ref.view(np.uint32)), i
        else:
            assert np.isinf(d[:, :, i]).all()


def test_rendered_scene_scores_one_at_truth(engine):
    s = scenes.make_scene(seed=73, ngrains=3, dilate=1)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    pos = np.vstack(s.voxels)
    grain = np.repeat(np.arange(3), [len(v) for v in s.voxels])
    engine.render_spots(pos, grain, np.stack(s.U))
    engine.distance_from_mask()
    ns, ni, C, med = engine.forward_comp(s.centers[0], s.q)
    assert C[0] == 1.0 and med[0] == 0.0
    # the stack built on the GPU scores exactly as the oracle on the same stack
    stack = oracle.Stack(engine.get_distance_stack(s.mask.shape))
    ref = oracle.or_space_comp(s.geo, stack, s.centers[0], s.q, s.rot, s.hkl)
    assert np.array_equal(ns, ref[:, 0]) and np.array_equal(C.view(np.uint64), ref[:, 2].view(np.uint64))
