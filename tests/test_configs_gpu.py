"""GPU parity on the benchmark configurations themselves, against the oracle.

Every configuration is built as the benchmark builds it (grainrecon_amd.synth:
seeded Voronoi cylinder, spots forward-simulated on the GPU, exact distance
transform on the GPU) and a seeded sample of its work is checked bit for bit
against the C oracle (counts, completeness bits, single-precision median,
argmax index):

  cfg1  BCC Fe, 100 um cylinder at 5 um (6 grains), 1024^2 detector at
        6.72 um, 181 omega: >= 10^4 (voxel, orientation) pairs through the
        fast and the plain kernel (forward_comp.m:4-90);
  cfg3  the cfg2 volume (3.2 M voxels, 300 grains, 2048^2 at 3.36 um) with
        CSR rows of Geometric(1/8) neighbour orientations truncated to 1..64,
        each expanded to 7^3 local samples (+-1.5 deg in 0.5 deg steps):
        the longest row (64 candidates, 21,952 evaluations) and random rows,
        argmax over the whole expanded row (grow_unindexed_compete_comp_cuda.m:39-48,
        revise_all_unindexed_voxel.m:138-153);
  cfg4  the cfg2 grains' matched spot centroids (300 grains, 0.5 px noise)
        at the perturbed start geometry: residual and adjoint gradient
        against the torch FP64 restatement (geo_fitting.m, dis_calc_spotspair.m);
  cfg5  FCC Al, 1000 grains, 2.5 M voxels in 8 uneven subvolumes, 2048^2 at
        3.36 um, 361 omega of 1 deg: one counts-and-argmax launch over every
        voxel's neighbour-grain list, rows sampled at random, the longest rows
        and every subvolume's first and last voxel; plus >= 10^4 pairs
        through the fast and plain kernels.
"""
import gc
import os

import numpy as np
import pytest

import oracle
from grainrecon_amd import _lib, synth

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _shape(cfg):
    return int(cfg.geo.detzsize), int(cfg.geo.detysize), cfg.rot.size


def _turned(q, deg, rng):
    """q turned by `deg` degrees about random axes (one row per angle)."""
    deg = np.asarray(deg, float)
    ax = rng.normal(size=(deg.size, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    h = np.deg2rad(deg) / 2
    t = np.column_stack([np.cos(h), np.sin(h)[:, None] * ax])
    out = synth._qmul(np.broadcast_to(q, t.shape), t)
    out[out[:, 0] < 0] *= -1
    return out


def _pairs_parity(engine, cfg, stack, pos_list, q_rows):
    """Fast and plain kernels against the oracle for every voxel's rows."""
    n = 0
    for pos, q in zip(pos_list, q_rows):
        ref = oracle.or_space_comp(cfg.geo, stack, pos, q, cfg.rot, cfg.hkl, nthreads=THREADS)
        for method in (_lib.METHOD_FAST, _lib.METHOD_PLAIN):
            ns, ni, C, med = engine.forward_comp(pos, q, method)
            assert np.array_equal(ns, ref[:, 0].astype(np.int32)), (pos, method)
            assert np.array_equal(ni, ref[:, 1].astype(np.int32)), (pos, method)
            assert np.array_equal(_bits(C), _bits(ref[:, 2])), (pos, method)
            assert np.array_equal(_bits(med), _bits(ref[:, 3].astype(np.float32))), (pos, method)
        n += q.shape[0]
    return n


def _row_argmax_parity(cfg, stack, pos, qs, best, ns, ni, C, med):
    """MATLAB max over the row's expanded list (first maximum, NaN skipped)."""
    ref = oracle.or_space_comp(cfg.geo, stack, pos, qs, cfg.rot, cfg.hkl, nthreads=THREADS)
    Cr = ref[:, 2]
    j = int(np.nanargmax(Cr)) if np.isfinite(Cr).any() else 0
    assert best == j, (best, j)
    assert ns == ref[j, 0] and ni == ref[j, 1]
    assert _bits(np.float64(C)) == _bits(np.float64(Cr[j]))
    assert _bits(np.float32(med)) == _bits(np.float32(ref[j, 3]))
    return ref


_CFG2 = {}


def _cfg2():
    if "cfg" not in _CFG2:
        _CFG2["cfg"] = synth.fe_config("cfg2")
    return _CFG2["cfg"]


def test_cfg1_pairs_parity(engine):
    """cfg1: 12 voxels x 896 orientations (each voxel's own grain turned by
    0.02-1.5 deg, the six grain orientations, and a seeded slice of a cubic
    fundamental-zone set of 9,216 orientations, the size of the 3 deg set)."""
    cfg = synth.fe_config("cfg1")
    smp = cfg.sample
    print(f"[cfg1] {smp.pos.shape[0]} voxels, {smp.q.shape[0]} grains, detector "
          f"{int(cfg.geo.detysize)}^2, {cfg.rot.size} omega", flush=True)
    synth.load_on_engine(engine, cfg)
    stack = oracle.Stack(engine.get_distance_stack(_shape(cfg)))
    rng = np.random.default_rng(3001)
    fz = synth.fz_reduce_cubic(synth.haar_quaternions(np.random.default_rng(2001), 9216))
    vox = rng.choice(smp.pos.shape[0], size=12, replace=False)
    pos_list, q_rows = [], []
    for k, v in enumerate(vox):
        own = smp.q[smp.grain[v]]
        q = np.vstack([own[None, :], smp.q, _turned(own, rng.uniform(0.02, 1.5, 89), rng),
                       fz[k * 800:(k + 1) * 800]])
        pos_list.append(smp.pos[v])
        q_rows.append(q)
    n = _pairs_parity(engine, cfg, stack, pos_list, q_rows)
    assert n >= 10_000
    # known answer: the voxel's own orientation projects onto its own rendered spots
    for pos, q in zip(pos_list, q_rows):
        _, _, C, _ = engine.forward_comp(pos, q[:1])
        assert C[0] >= 0.98, C[0]
    del stack
    gc.collect()


def _csr_rows(smp, rng, n_voxels, force_longest=True):
    """CSR rows of the batched compete: each voxel's own grain plus
    Geometric(1/8) - 1 other grains (truncated to 1..64 in all), ascending
    grain id as find_neighbor_ORs returns them; row 0 has 64 candidates."""
    G = smp.q.shape[0]
    vox = rng.choice(smp.pos.shape[0], size=n_voxels, replace=False)
    k = np.minimum(rng.geometric(1.0 / 8.0, size=n_voxels), 64)
    if force_longest:
        k[0] = 64
    rows, cands = [0], []
    for i, vi in enumerate(vox):
        own = int(smp.grain[vi])
        others = rng.choice(np.delete(np.arange(G), own), size=k[i] - 1, replace=False)
        ids = np.sort(np.concatenate([[own], others]))
        cands.append(smp.q[ids])
        rows.append(rows[-1] + ids.size)
    return vox, np.array(rows, np.int64), np.vstack(cands)


def _expand_local(qc, nr, g):
    """The local-sampling grid of one candidate (q1 slowest, q0 recomputed by
    the scorer): (qc_i - g*nr/2) + k*g, k = 0..nr, same IEEE operations as
    the device."""
    half = (g * nr) / 2.0
    k = np.arange(nr + 1, dtype=float) * g
    a, b, c = (qc[1] - half) + k, (qc[2] - half) + k, (qc[3] - half) + k
    A, B, C = np.meshgrid(a, b, c, indexing="ij")
    return np.stack([np.zeros(A.size), A.reshape(-1), B.reshape(-1), C.reshape(-1)], axis=1)


def test_cfg3_batched_csr_parity(engine):
    """cfg3: 1024 CSR rows scored in one batched launch; the longest row
    (64 candidates x 343 samples) and 7 random rows checked against the
    oracle's argmax over the whole expanded row."""
    cfg = _cfg2()
    smp = cfg.sample
    print(f"[cfg3] cfg2 volume: {smp.pos.shape[0]} voxels, {smp.q.shape[0]} grains", flush=True)
    synth.load_on_engine(engine, cfg)
    rng = np.random.default_rng(3003)
    vox, rp, cq = _csr_rows(smp, rng, 1024)
    nr, g = 6, float(np.sin(np.deg2rad(0.25)))
    best, bq, ns, ni, C, med = engine.compete_csr(smp.pos[vox], rp, cq, nr, g)
    sizes = np.diff(rp)
    assert sizes.max() == 64 and sizes.min() >= 1
    rows = [0] + [int(r) for r in rng.choice(np.arange(1, vox.size), size=7, replace=False)]
    stack = oracle.Stack(engine.get_distance_stack(_shape(cfg)))
    multi = 0
    for r in rows:
        qs = np.vstack([_expand_local(cq[c], nr, g) for c in range(rp[r], rp[r + 1])])
        print(f"[cfg3] row {r}: {sizes[r]} candidates, {qs.shape[0]} evaluations", flush=True)
        _row_argmax_parity(cfg, stack, smp.pos[vox[r]], qs, int(best[r]), ns[r], ni[r], C[r], med[r])
        multi += int(best[r] >= 343)
        # the winner's quaternion is the expanded row's entry with q0 recomputed
        w = qs[int(best[r])]
        assert np.array_equal(bq[r, 1:], w[1:])
    # the winning sample lies outside the first candidate's block in at least one checked row
    assert multi >= 1
    # every row's winner indexes its own expanded list
    assert np.all((best >= 0) & (best < sizes * 343))
    del stack
    gc.collect()


def test_cfg4_adjoint_all_grains(engine):
    """cfg4: the 300 grains' matched centroids (0.5 px noise) at the perturbed
    start (Lsam2sou -0.2 mm, Lsam2det +0.3 mm, dety00/detz00 +-20 px, tilts
    +-0.2 deg; grains as indexed): residual and gradient of every geometry
    value, Rodrigues vector and centre of mass against torch autograd in FP64."""
    from grainrecon_amd import fit as F
    from oracle import residual_ref

    cfg = _cfg2()
    smp, geo = cfg.sample, cfg.geo
    engine.set_geometry(geo, cfg.hkl, cfg.rot)
    rng = np.random.default_rng(4004)
    G = smp.q.shape[0]
    cnt = np.bincount(smp.grain, minlength=G).astype(float)
    com = np.stack([np.bincount(smp.grain, weights=smp.pos[:, k], minlength=G) for k in range(3)], 1) / cnt[:, None]
    px = geo.pixelysize
    g7_true = np.array([geo.Lsam2sou, geo.Lsam2det, geo.dety00, geo.detz00, 0.0, 0.0, 0.0])
    gr6_true = np.hstack([smp.q[:, 1:] / smp.q[:, :1], com])
    sp = F.simulate_matched_spots(engine, g7_true, com, smp.U, rng=rng, noise_px=0.5)
    fit = F.SpotFit(engine, cfg.hkl[sp["hkl_idx"]], cfg.rot[sp["omega_idx"]], sp["meas"], sp["grain"], G)
    print(f"[cfg4] {G} grains, {fit.n} matched spots", flush=True)
    assert fit.n > 100_000 and np.unique(sp["grain"]).size == G
    spots = dict(hkl=fit.hkl, omega=fit.omega, meas=fit.meas, grain=fit.grain)
    g7_0 = g7_true + np.array([-0.2, 0.3, 20 * px, -20 * px, 0.2, -0.2, 0.2])
    gr6_0 = gr6_true + np.hstack([rng.normal(scale=1e-4, size=(G, 3)), rng.normal(scale=2e-3, size=(G, 3))])
    for g7, gr6 in ((g7_0, gr6_0), (g7_true, gr6_0)):
        r, d7, d6 = fit.value_and_grad(g7, gr6)
        _, t7, t6, rt = residual_ref.loss_and_grad(geo, spots, g7, gr6)
        assert np.abs(r - rt).max() <= 1e-8 * max(1.0, np.abs(rt).max())
        assert np.linalg.norm(d7 - t7) <= 1e-8 * np.linalg.norm(t7)
        assert np.linalg.norm(d6 - t6) <= 1e-8 * np.linalg.norm(t6)
        # per grain too: no grain's block hides under the others' norm
        per = np.linalg.norm(d6 - t6, axis=1) <= 1e-8 * np.maximum(np.linalg.norm(t6, axis=1), 1e-300) + 1e-12
        assert per.all()


def test_cfg5_al_one_launch_parity(engine):
    """cfg5: one counts-and-argmax launch over 2.5 M voxels in 8 subvolumes,
    then >= 10^4 (voxel, orientation) pairs through the fast and plain kernels."""
    cfg = synth.al_config()
    smp = cfg.sample
    slab = synth.subvolumes(smp, [0.02, 0.05, 0.08, 0.10, 0.13, 0.16, 0.20, 0.26])
    order = np.argsort(slab, kind="stable")
    gid = synth.grain_id_volume(smp)
    print(f"[cfg5] {smp.pos.shape[0]} voxels, {smp.q.shape[0]} grains, subvolumes "
          f"{np.bincount(slab).tolist()}, {cfg.rot.size} omega, {cfg.hkl.shape[0]} hkl", flush=True)
    synth.load_on_engine(engine, cfg)
    vox = smp.index[order]
    ids, ind = engine.find_neighbor_ors(gid, vox, 2, 30)
    row_ptr = np.concatenate([[0], ind[:, 1]]).astype(np.int64)
    assert np.array_equal(ind[:, 0] - 1, row_ptr[:-1])
    cand_q = smp.q[ids - 1]
    pos = smp.pos[order]
    M = pos.shape[0]
    assert M > 2_400_000
    best, bq, ns, ni, C, med = engine.compete_csr(pos, row_ptr, cand_q, 0, 0.0)
    sizes = np.diff(row_ptr)
    rng = np.random.default_rng(3005)
    starts = np.concatenate([[0], np.cumsum(np.bincount(slab))])
    rows = set(int(r) for r in rng.choice(M, size=40, replace=False))
    rows |= set(int(r) for r in np.argsort(sizes, kind="stable")[-4:])
    for s in range(8):
        rows |= {int(starts[s]), int(starts[s + 1] - 1)}
    rows = sorted(rows)
    # neighbour lists of these rows against the oracle's find_neighbor_ORs
    oids, oind = oracle.find_neighbor_ors(gid, vox[rows], 2, 30)
    for k, r in enumerate(rows):
        a = ids[ind[r, 0] - 1:ind[r, 1]]
        b = oids[oind[k, 0] - 1:oind[k, 1]]
        assert np.array_equal(a, b), r
    stack = oracle.Stack(engine.get_distance_stack(_shape(cfg)))
    for r in rows:
        qs = cand_q[row_ptr[r]:row_ptr[r + 1]]
        _row_argmax_parity(cfg, stack, pos[r], qs, int(best[r]), ns[r], ni[r], C[r], med[r])
    print(f"[cfg5] {len(rows)} rows checked (longest {sizes.max()} candidates)", flush=True)
    # >= 10^4 pairs: 8 voxels x 1280 orientations (own grain turned, neighbours, random)
    fz = synth.fz_reduce_cubic(synth.haar_quaternions(np.random.default_rng(2005), 8 * 1200))
    pos_list, q_rows = [], []
    for k, v in enumerate(rng.choice(M, size=8, replace=False)):
        own = smp.q[smp.grain[order[v]]]
        q = np.vstack([own[None, :], _turned(own, rng.uniform(0.02, 1.5, 79), rng), fz[k * 1200:(k + 1) * 1200]])
        pos_list.append(pos[v])
        q_rows.append(q)
    n = _pairs_parity(engine, cfg, stack, pos_list, q_rows)
    assert n >= 10_000
    del stack
    gc.collect()
