"""The compete engine of GrainGrow_engine_compete_comp.m on the device
stages against the same loop over the oracle's neighbour search and compete."""
import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd.mapping import MappingDS


@pytest.mark.gpu
def test_compete_engine_equals_oracle_loop(engine):
    from grainrecon_amd.compete import grain_grow_compete_engine
    from grainrecon_amd.engine import euler_to_rod_quat, u2euler

    s = scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    n, vox = 12, 0.005
    p1 = np.array([1, 1, 1, 0, vox, *s.centers[0], n, n, n], float)
    mask = np.zeros((n, n, n))
    mask[2:10, 2:10, 2:10] = 1
    ds = MappingDS(mask)
    eul = [u2euler(U.flatten(order="F")[None, :])[0] for U in s.U]
    # grain 1 (grain 0's orientation) around the centre voxel, grain 2 (grain 1's) in a corner
    for (lo, hi, g) in (((5, 8), (5, 8), 0), ((2, 4), (2, 4), 1)):
        ds.GrainId[lo[0]:lo[1], hi[0]:hi[1], 4:7] = g + 1
        ds.EulerZXZ[:, lo[0]:lo[1], hi[0]:hi[1], 4:7] = eul[g][:, None, None, None]
    ref = MappingDS(mask)
    for k in ("GrainId", "EulerZXZ"):
        setattr(ref, k, getattr(ds, k).copy())
    print("compete engine: 12^3 volume, 2 grains", flush=True)
    n_rev = grain_grow_compete_engine(engine, ds, p1, search_radius=4, nr_voxel_per_iter=97)
    # the loop over the oracle
    flat = ref.GrainId.reshape(-1, order="F")
    lin = np.flatnonzero((flat == 0) & (ref.Mask.reshape(-1, order="F") == 1))
    wait = np.stack(np.unravel_index(lin, ref.GrainId.shape, order="F"), 1) + 1
    assert n_rev == wait.shape[0] > 100
    euler = np.array([eul[0], eul[1]])
    ids, ind = oracle.find_neighbor_ors(ref.GrainId.astype(np.int32), wait, 4)
    out = oracle.compete(s.geo, s.stack, np.column_stack([wait, ind]), ids, euler, s.rot, s.hkl, p1)
    x, y, z = wait[:, 0] - 1, wait[:, 1] - 1, wait[:, 2] - 1
    np.testing.assert_array_equal(ds.GrainId[x, y, z], out[:, 4])
    np.testing.assert_array_equal(ds.Completeness[x, y, z], out[:, 5])
    np.testing.assert_array_equal(ds.Ninter[x, y, z], out[:, 6])
    np.testing.assert_array_equal(ds.Dismedian[x, y, z], out[:, 7])
    np.testing.assert_array_equal(ds.EulerZXZ[:, x, y, z], out[:, 1:4].T)
    np.testing.assert_array_equal(ds.Rodrigues[:, x, y, z], euler_to_rod_quat(out[:, 1:4])[0].T)
    assert np.all(ds.VisitFlag[x, y, z] == 0.5)
    assert set(np.unique(out[:, 4])) <= {1.0, 2.0}
