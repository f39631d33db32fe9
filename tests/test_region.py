"""Region growth around an indexed seed (grow_indexed_region_cuda.m:29-151):
the oracle's bookkeeping pinned with controlled growth decisions, the GPU
(grs_grow_region on resident volumes) against the oracle."""
import math

import numpy as np
import pytest

import oracle.grow_region as GR
import scenes


class _Geo:
    pixelysize = pixelzsize = 0.01
    Lsam2sou = 10.0
    Lsam2det = 10.0


def _vols(n=9):
    return {"Completeness": np.zeros((n, n, n)), "Mask": np.ones((n, n, n)), "Dismedian": np.zeros((n, n, n)),
            "Ninter": np.zeros((n, n, n))}


def _run_with(monkeypatch, accept, seed=(5, 5, 5), spot_L=1.0, maxD=1.0):
    """grow_region with the operator replaced by `accept(x, y, z) -> C or None`."""
    def fake_grow(geo, stack, rows, U, rot, hkl, p1, nthreads=8):
        out = np.zeros((rows.shape[0], 7))
        out[:, :3] = rows[:, :3]
        for k, (x, y, z) in enumerate(rows[:, :3].astype(int)):
            c = accept(x, y, z)
            if c is not None:
                out[k, 3:7] = [100.0, round(c * 100), 0.5, 1.0]
        return out
    monkeypatch.setattr(GR, "grow", fake_grow)
    v = _vols()
    p1 = [1, 1, 1, 0, 0.1, 0, 0, 0, 9, 9, 9]  # box half width = ceil(spot_L*0.01/(1*0.1)) voxels
    return GR.grow_region(_Geo, None, v, seed, 0.9, np.eye(3), spot_L * 10, 0.1, maxD, 5.0, p1, None, None), v


def test_region_box_and_sphere(monkeypatch):
    (center, rc, idx, nchk), v = _run_with(monkeypatch, lambda x, y, z: None, spot_L=2.0)
    # Box_L = 2: the 5^3 box around (5,5,5), sphere of radius 2 about its centre point
    ref = sum(1 for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3) if a * a + b * b + c * c <= 4)
    assert nchk == ref == 33
    assert rc == 0 and np.array_equal(center, [5.0, 5.0, 5.0]) and idx.tolist() == [[5, 5, 5]]
    assert v["Completeness"][4, 4, 4] == 0.9


def test_region_symmetric_growth_keeps_centre(monkeypatch):
    (center, rc, idx, _), v = _run_with(monkeypatch, lambda x, y, z: 0.8 if abs(x - 5) + abs(y - 5) + abs(z - 5) == 1
                                        else None)
    assert idx.shape == (7, 3) and rc == 0 and np.allclose(center, 5.0)
    assert v["Completeness"][5, 4, 4] == 0.8 and v["Ninter"][5, 4, 4] == 80 and v["Dismedian"][5, 4, 4] == 0.5


def test_region_one_sided_growth_moves_centre(monkeypatch):
    (center, rc, idx, _), _ = _run_with(monkeypatch, lambda x, y, z: 0.8 if x > 5 else None, spot_L=2.0, maxD=0.5)
    # seed counted twice (initial term and in J), the accepted voxels once, in find() order
    cx = 0.9 * 5 + 0.9 * 5
    s = 0.9 + 0.9
    for (x, y, z) in idx.tolist():
        if (x, y, z) != (5, 5, 5):
            cx += 0.8 * x
            s += 0.8
    assert rc == 1 and center[0] == round(cx / s) and center[0] > 5
    # find() order: x fastest, then y, then z
    lin = [(x - 1) + 9 * ((y - 1) + 9 * (z - 1)) for x, y, z in idx.tolist()]
    assert lin == sorted(lin)


@pytest.mark.gpu
def test_grow_region_equals_oracle(engine):
    from grainrecon_amd._lib import GrsError

    s = scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    n = 14
    vox = 0.005
    p1 = np.array([1, 1, 1, 0, vox, *s.centers[0], n, n, n], float)  # voxel (7,7,7) sits on grain 0's centre
    rng = np.random.default_rng(21)
    vols = {"Completeness": rng.uniform(0, 0.3, (n, n, n)), "Mask": (rng.random((n, n, n)) > 0.05).astype(float),
            "Dismedian": rng.uniform(0, 5, (n, n, n)), "Ninter": np.zeros((n, n, n))}
    for k, v in vols.items():
        engine.set_volume(k, v)
    g = s.geo
    spot_L = 3.0 * ((g.Lsam2det / g.Lsam2sou) * vox) / ((g.pixelysize + g.pixelzsize) / 2.0)
    for seed, comp in (((7, 7, 7), 0.9), ((8, 6, 7), 0.7)):
        print(f"grow_region: seed {seed}", flush=True)
        got = engine.grow_region(seed, comp, s.U[0], spot_L, 0.5, 1.0, 10.0, p1)
        c, rc, idx, nchk = GR.grow_region(g, s.stack, vols, seed, comp, s.U[0], spot_L, 0.5, 1.0, 10.0, p1, s.rot,
                                          s.hkl)
        assert got["n_checked"] == nchk and nchk > 50
        assert got["indexed"].tolist() == idx.tolist() and idx.shape[0] > 3
        assert got["replace_center"] == rc and np.array_equal(got["center"], c)
        for k, v in vols.items():
            np.testing.assert_array_equal(engine.get_volume(k), v, err_msg=k)
        # the device gather at the grown voxels (and a few others, repeated
        # and out of order) reads the same values
        vox = np.concatenate([idx, [[1, 1, 1], [n, n, n], [1, n, 2]], idx[::-1]])
        for k, v in vols.items():
            want = np.array([v[a - 1, b - 1, c - 1] for a, b, c in vox.tolist()])
            np.testing.assert_array_equal(engine.get_volume_at(k, vox), want, err_msg=k)
    assert engine.get_volume_at("Completeness", np.zeros((0, 3), int)).shape == (0,)
    for bad in ([[n + 1, 1, 1]], [[0, 1, 1]], [[1, 1, n + 1]]):
        with pytest.raises(GrsError):
            engine.get_volume_at("Completeness", bad)
    # the device scatter writes exactly the listed voxels
    vox = np.array([[1, 1, 1], [n, n, n], [3, 5, 7], [7, 5, 3]])
    for k in ("Completeness", "Dismedian", "Ninter"):
        want = engine.get_volume(k).copy()
        vals = np.array([0.125, -2.5, 7.0, np.nan])
        engine.set_volume_at(k, vox, vals)
        for (a, b, c), w in zip(vox.tolist(), vals):
            want[a - 1, b - 1, c - 1] = w
        np.testing.assert_array_equal(engine.get_volume(k), want, err_msg=k)
    for bad in ([[n + 1, 1, 1]], [[2, 2, 2], [2, 2, 2]]):
        with pytest.raises(GrsError):
            engine.set_volume_at("Completeness", bad, np.zeros(len(bad)))


def _connected_problem():
    s = scenes.make_scene(seed=83, ngrains=3, spread_mm=0.05, nrot=37)
    n, vox = 14, 0.005
    p1 = np.array([1, 1, 1, 0, vox, *s.centers[0], n, n, n], float)  # voxel (7,7,7) sits on grain 0's centre
    rng = np.random.default_rng(23)
    vols = {"Completeness": rng.uniform(0, 0.3, (n, n, n)), "Mask": (rng.random((n, n, n)) > 0.05).astype(float),
            "Dismedian": rng.uniform(0, 5, (n, n, n)), "Ninter": rng.integers(0, 40, (n, n, n)).astype(float)}
    vols["Dismedian"][rng.random((n, n, n)) < 0.05] = np.nan
    return s, p1, vols


def test_grow_connected_oracle_grows_the_grain():
    """The connected growth from grain 0's centre stays connected (every grown
    voxel has an 18-neighbour in the region) and covers the grain's voxels."""
    from oracle import grow_connected as GC

    s, p1, vols = _connected_problem()
    c, rc, idx = GC.grow_connected(s.geo, s.stack, vols, (7, 7, 7), 0.95, s.U[0], 0.2, 1.0, 10.0, p1, s.rot, s.hkl)
    cells = {tuple(v) for v in idx.tolist()}
    assert (7, 7, 7) in cells and len(cells) > 5
    offs = [o for o in GC.NEIGHBOURS]
    for v in cells:
        if v != (7, 7, 7):
            assert any((v[0] + a, v[1] + b, v[2] + d) in cells for a, b, d in offs)
    assert vols["Completeness"][6, 6, 6] == 0.95


@pytest.mark.gpu
def test_grow_connected_equals_oracle(engine):
    """grow_indexed_region.m on the engine (grainrecon_amd.mapping.grow_connected)
    against the oracle: the same region in find() order, centre decision and
    written-back volumes."""
    from grainrecon_amd.mapping import MappingDS, grow_connected
    from oracle import grow_connected as GC

    s, p1, vols = _connected_problem()
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    for seed, comp, drop in (((7, 7, 7), 0.95, 0.2), ((8, 6, 7), 0.7, 0.5)):
        ds = MappingDS(vols["Mask"])
        ds.Completeness, ds.Dismedian, ds.Ninter = (vols[k].copy() for k in ("Completeness", "Dismedian", "Ninter"))
        print(f"grow_connected: seed {seed}", flush=True)
        c, rc, idx = grow_connected(engine, ds, seed, comp, s.U[0], drop, 1.0, 10.0, p1)
        ref = {k: v.copy() for k, v in vols.items()}
        c2, rc2, idx2 = GC.grow_connected(s.geo, s.stack, ref, seed, comp, s.U[0], drop, 1.0, 10.0, p1, s.rot, s.hkl)
        assert idx.tolist() == idx2.tolist() and idx.shape[0] > 3
        assert rc == rc2 and np.array_equal(c, c2)
        for k, a in (("Completeness", ds.Completeness), ("Dismedian", ds.Dismedian), ("Ninter", ds.Ninter)):
            np.testing.assert_array_equal(a, ref[k], err_msg=k)
