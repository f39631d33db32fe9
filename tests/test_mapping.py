"""The seed -> index -> fit -> grow -> re-centre loop
(GrainMapping_engine_gpu_comp.m) on the device stages: runs end to end on a
small scene, indexes the grains at their centres, grows them into the volume,
and gives the same result twice."""
import numpy as np
import pytest

import oracle
import scenes


def _frames(s):
    frames = [[] for _ in range(s.rot.size)]
    for g in range(len(s.U)):
        _, _, _, ev = oracle.forward_comp(s.geo, s.stack, s.U[g], s.centers[g], s.rot, s.hkl, events=True)
        for om, _h, dy, dz, *_ in ev:
            frames[int(om)].append([dy, dz, 30.0])
    return [np.array(f).reshape(-1, 3) for f in frames]


@pytest.mark.gpu
def test_grain_mapping_engine_end_to_end(engine):
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine
    from test_index import _index_problem

    s, qs, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_frames(s))
    n, vox = 14, 0.005
    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=2.0, max_dmedian=20.0, nr_seed=6, first_grain_id=1,
              center=s.centers[0], vox_size=vox, rvp_first=[1, 1, 1], tomo_dim=(n, n, n))
    pos = np.array([s.centers[0], s.centers[0] + [0.005, 0.0, 0.0]])
    runs = []
    for _ in range(2):
        ds = MappingDS(np.ones((n, n, n)))
        print("grain_mapping_engine: 2 seeds on a 14^3 volume", flush=True)
        out, pos_new, nvox, _ = grain_mapping_engine(engine, pos, qs, ds, **kw)
        runs.append((out, pos_new, nvox, ds))
    (o1, p1, n1, d1), (o2, p2, n2, d2) = runs
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(p1, p2)
    for k in ("Completeness", "GrainId", "EulerZXZ", "VisitFlag", "Dismedian", "Ninter"):
        np.testing.assert_array_equal(getattr(d1, k), getattr(d2, k), err_msg=k)
    # the first seed indexes grain 0 (C = 1) and grows it; the second seed lies in the grown region
    assert o1[0, 6] > 0.95 and n1[0] > 1 and d1.GrainId[6, 6, 6] == 1
    assert d1.VisitFlag[6, 6, 6] == 1.0
    assert (d1.GrainId == 1).sum() == n1[0]


@pytest.mark.gpu
def test_grain_mapping_engine_small_spots_grows_by_connected_voxels(engine):
    """Spot areas of 3 px^2 (half length sqrt(3/pi) < 2 px) send the growth to
    the connected-voxel loop (GrainMapping_engine_gpu_comp.m:207-214): the
    grain grows from its centre, the engine's volumes equal the host's after
    the run, and a second run gives the same result."""
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine
    from test_index import _index_problem

    s, qs, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    frames = [f.copy() for f in _frames(s)]
    for f in frames:
        if f.shape[0]:
            f[:, 2] = 3.0
    engine.set_frame_spots(frames)
    n, vox = 14, 0.005
    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=2.0, max_dmedian=20.0, nr_seed=6, first_grain_id=1,
              center=s.centers[0], vox_size=vox, rvp_first=[1, 1, 1], tomo_dim=(n, n, n))
    runs = []
    for _ in range(2):
        ds = MappingDS(np.ones((n, n, n)))
        print("grain_mapping_engine: connected growth, 1 seed", flush=True)
        out, pos_new, nvox, _ = grain_mapping_engine(engine, s.centers[0][None, :], qs, ds, **kw)
        for k in ("Completeness", "Dismedian", "Ninter"):
            np.testing.assert_array_equal(engine.get_volume(k), getattr(ds, k), err_msg=k)
        runs.append((out, nvox, ds))
    (o1, n1, d1), (o2, n2, d2) = runs
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_array_equal(d1.GrainId, d2.GrainId)
    assert o1[0, 6] > 0.95 and n1[0] > 1 and d1.GrainId[6, 6, 6] == 1


def _vols(ds):
    return {k: getattr(ds, k).copy(order="F") for k in ("Completeness", "Mask", "GrainId", "Rodrigues", "IPF001",
                                                          "EulerZXZ", "Dismedian", "Ninter", "VisitFlag")}


def _same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.gpu
@pytest.mark.parametrize("area", [30.0, 3.0])
def test_grain_mapping_engine_equals_oracle_loop(engine, area):
    """The whole loop against oracle/mapping.py (GrainMapping_engine_gpu_comp.m:54-331
    composed from the stage oracles) on four seeds: the indexing rows, the new
    seed positions, the grown voxel counts, the SeedIndexing rows and every DS
    volume bit for bit.  Spot areas of 30 px^2 grow inside the bounding sphere
    (grow_indexed_region_cuda.m), 3 px^2 by connected voxels
    (grow_indexed_region.m)."""
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine
    from oracle.mapping import grain_mapping
    from test_index import _index_problem

    s, qs, _ = _index_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    frames = [f.copy() for f in _frames(s)]
    for f in frames:
        if f.shape[0]:
            f[:, 2] = area
    engine.set_frame_spots(frames)
    n, vox = 14, 0.005
    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=2.0, max_dmedian=20.0, nr_seed=6, first_grain_id=1,
              center=s.centers[0], vox_size=vox, rvp_first=[1, 1, 1], tomo_dim=(n, n, n))
    pos = np.array([s.centers[0], s.centers[0] + [0.005, 0.0, 0.0], s.centers[1], s.centers[2],
                    s.centers[0] + [-0.02, 0.015, 0.01]])
    mask = np.ones((n, n, n))
    mask[:, :, 0] = 0.0
    ds = MappingDS(mask)
    ds.GrainId[0, 0, 1:3] = 9.0  # indexed before the call
    vols = _vols(ds)
    print(f"grain_mapping_engine vs oracle loop: {pos.shape[0]} seeds, spot area {area}", flush=True)
    got = grain_mapping_engine(engine, pos, qs, ds, **kw)
    ref = grain_mapping(s.geo, s.stack, frames, s.rot, s.hkl, pos, qs, vols, **kw)
    for name, a, b in zip(("indexing_final", "pos_list_new", "Nr_indexed_voxel", "SeedIndexing"), got, ref):
        assert _same(a, b), (name, a, b)
    for k, v in vols.items():
        assert _same(getattr(ds, k), v), k
    assert got[2][0] > 1 and got[0][0, 6] > 0.95
