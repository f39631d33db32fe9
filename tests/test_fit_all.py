"""The fit_all seed indexing of fit_OR_constrain_with_local_comp_cuda_2.m
(two rounds of finer local gridding, five fits): the engine composition
(grainrecon_amd.indexing.index_seed_fit_all) against the oracle
(oracle.index_seed.index_seed_fit_all), bit for bit."""
import numpy as np
import pytest

import oracle
from oracle import index_seed as IS


def _same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def test_fit_all_needs_five_seeds():
    """Nr_fit = 5 fits index the first five rows (MATLAB raises below five)."""
    from test_index import _index_problem

    s, qs, pos = _index_problem()
    with pytest.raises(IndexError):
        IS.index_seed_fit_all(s.geo, s.stack, pos[0], qs[:40], s.rot, s.hkl, nr_seed=4)


@pytest.mark.gpu
def test_fit_all_equals_oracle(engine):
    from grainrecon_amd.indexing import index_seed_fit_all
    from test_index import _index_problem
    from test_mapping import _frames

    s, qs, pos = _index_problem()
    q = qs[:400]
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_frames(s))
    for p in pos[:2]:
        print(f"fit_all at {p}: {q.shape[0]} orientations, Nr_seed 8", flush=True)
        got = index_seed_fit_all(engine, p, q, 8)
        ref = IS.index_seed_fit_all(s.geo, s.stack, p, q, s.rot, s.hkl, 8)
        assert _same(got["output_max"], ref["output_max"])
        assert got["candidate"] == ref["candidate"]
        for g, r in zip(got["fits"], ref["fits"]):
            assert _same(g["x0"], r["x0"]) and _same(g["x"], r["x"])
            assert g["fval"] == r["fval"] and g["n_fun_evals"] == r["nf"] and g["n_iter"] == r["it"]
            assert _same(g["C"], r["C"]) and g["group"] == r["group"]
        assert _same(got["indexing"][1:], ref["indexing"][1:])
        # the chosen fit's pairs are those of forward_comp_pair_spots for its U
        ns, ni, med = oracle.forward_comp(s.geo, s.stack, engine_U(got["indexing"]), p, s.rot, s.hkl)
        assert ns == got["indexing"][4] and ni == got["indexing"][5]


def engine_U(row):
    from grainrecon_amd.geometry import quaternion2U

    return quaternion2U(np.asarray(row[14:18], float))


@pytest.mark.gpu
def test_mapping_takes_fit_all_for_seeds_away_from_indexed_voxels(engine):
    """GrainMapping_engine_gpu_comp.m:46-53,119-127: with check_fit_all, a
    seed more than 2 voxels from every voxel of Completeness >= TrustComp is
    indexed by the fit_all variant, a nearer one by the plain one."""
    from grainrecon_amd.indexing import index_seed_fit_all
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine
    from test_index import _index_problem
    from test_mapping import _frames

    s, qs, _ = _index_problem()
    q = qs[:400]
    engine.set_geometry(s.geo, s.hkl, s.rot)
    engine.set_distance_stack(s.proj)
    engine.set_frame_spots(_frames(s))
    n, vox = 14, 0.005
    # maxD large: the centre-of-mass fit never replaces the indexed row
    kw = dict(trust_comp=0.9, min_comp=0.5, drop_off=0.2, maxD=1e9, max_dmedian=20.0, nr_seed=8, first_grain_id=1,
              center=s.centers[0], vox_size=vox, rvp_first=[1, 1, 1], tomo_dim=(n, n, n), check_fit_all=True)
    pos = s.centers[0][None, :]
    # the two variants give different rows here, so the check below tells
    # which one ran
    fa = index_seed_fit_all(engine, s.centers[0], q, 8)["indexing"]
    engine.set_orientation_set(q)
    pl = engine.index_seeds(s.centers[0][None, :], nr_seed=8)["indexing"][0]
    assert not _same(fa[1:], pl[1:])
    # the seed voxel: pos2ind of the centre is (7, 7, 7) (1-based)
    for far, marked in ((True, (1, 1, 1)), (False, (7, 7, 8))):
        ds = MappingDS(np.ones((n, n, n)))
        ds.Completeness[marked[0] - 1, marked[1] - 1, marked[2] - 1] = 0.95
        print(f"mapping with check_fit_all, indexed voxel at {marked}", flush=True)
        out, _, _, _ = grain_mapping_engine(engine, pos, q, ds, **kw)
        assert _same(out[0, 1:], (fa if far else pl)[1:])
