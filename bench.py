#!/usr/bin/env python3
"""Benchmark of the forward-completeness hot path on one MI355X.

Workload (config 2, "plain method at scale"): a 3.2M-voxel BCC Fe cylinder
(400 um at 2.5 um, 300 Voronoi grains) forward-simulated onto a 2048^2
detector (3.36 um pixels) over 181 omega steps; the distance stack is built on
the GPU.  One step = the seeding call of cuda_forward_comp: one seed voxel
scored against a 262144-orientation cubic fundamental-zone set (the size of
the 1 degree set), orientations and outputs resident in HBM.

value      evals/s (voxel x orientation) of the fast (batched) kernel over the
           timed steps; the plain thread-per-evaluation kernel is timed on
           the same workload and reported beside it.
roofline   FP64 vector bound: the FP64 operations of the pruned algorithm the
           seeding kernel runs (per evaluation F_SETUP + F_SOLVE*H + F_CAND*
           candidates + F_SPOT*window passes + F_EXACT*exact-pipeline events,
           the per-unit counts taken from the kernel source and the units
           counted by an audit launch of the same workload), times the
           evaluations of one launch, over the launch's HIP-event duration on
           the engine's stream, against the 78.6 TFLOP/s FP64 peak; beside it
           the forward_comp.m-equivalent rate (Nomega*(118*H + 124) + 90 + 15*H
           per evaluation) and, from the committed counter profile of the same
           build, executed FP64 operations and HBM traffic.
cpu_baseline  the C oracle (FP64 restatement of forward_comp.m, OpenMP) on a
           bounded sample of the same workload, which is also checked bit for
           bit against the GPU results.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "completeness evals/s (voxel×OR, bit-exact) batched vs plain; geo-fit wall time"
PEAK_FP64_TFLOPS = 78.6


def flop_per_eval(nrot: int, H: int) -> int:
    """The full forward_comp.m work of one evaluation (every event through
    the whole pipeline): the work the reference's algorithm does."""
    return nrot * (118 * H + 124) + 90 + 15 * H


# FP64 operations of the algorithm k_fast runs, per unit, counted from the
# kernel source (grs_kernels.hip; + - * / sqrt rsq rcp acos atan2 = 1 each, a
# fused multiply-add = 2; comparisons and integer work are not counted):
F_SETUP = 136    # per evaluation: q0, U(q), S*U*B (k_sub_from_q)
F_SOLVE = 182    # per (evaluation, reflection): G, |G|, thresholds, the cos(beta) window with the detector
                 # rows and columns bound, om_ranges (a refined rcp/rsq counts its Newton step: 4 and 6)
F_CAND = 20      # per candidate event: Gt, cos(beta), lambda and the window tests (b1_classify; the
                 # thresholds come precomputed per block)
F_SPOT = 79      # per candidate inside the windows: spot, tilted-detector intersection (K_out
                 # unnormalised), pixel
F_EXACT = 242    # per event sent to the exact pipeline: per-(voxel, omega) terms 124 + event 118


def pruned_flop_per_eval(H, cand, in_window, exact):
    """FP64 operations per evaluation of the pruned algorithm, from the
    average candidates, window passes and exact-pipeline events per
    evaluation (the audit counters of the same workload)."""
    return F_SETUP + F_SOLVE * H + F_CAND * cand + F_SPOT * in_window + F_EXACT * exact


# the translation unit of k_fast: its counters change only with these files
KFAST_SOURCES = ("grs_geom.h", "grs_internal.h", "grs_kernels.hip", "grs_math.h")


def kernel_sources_sha():
    """sha256 (first 16 hex digits) of the sources k_fast is compiled from:
    names a build so that a counter profile can be matched to it."""
    import hashlib

    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "grainrecon_amd", "csrc")
    for name in KFAST_SOURCES:
        h.update(name.encode())
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def counter_profile(name="k_fast_seed_cfg2_counters.md"):
    """The machine-readable line of a committed counter profile
    (`counters_json: {...}`), or None.  Its sources_sha names the build it was
    measured on."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("counters_json:"):
            return json.loads(line.split(":", 1)[1])
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def roofline_leg(eng, cfg, q, seeds, kernel_ms, N, config, n_audit):
    """Roofline of the seeding launch: the FP64 work of the pruned algorithm
    k_fast runs (pruned_flop_per_eval, with its units per evaluation counted
    by the audit launch on n_audit orientations at two of the timed seed
    voxels) over the launch time and the FP64 vector peak.  The executed FP64
    instructions and the HBM traffic come from the committed counter profile
    when its sources_sha names this build, else they are null."""
    H = cfg.hkl.shape[0]
    qa = q[:n_audit]
    tot = {"examined": 0, "in_window": 0, "exact": 0, "mismatched": 0}
    sel = 0
    for p in seeds:
        eng.set_audit(True)
        try:
            ns, _, _, _ = eng.forward_comp(p, qa)
            a = eng.audit()
        finally:
            eng.set_audit(False)
        for k in tot:
            tot[k] += a[k]
        sel += int(ns.sum())
    E = qa.shape[0] * len(seeds)
    units = {"reflections": H, "candidates": tot["examined"] / E, "in_window": tot["in_window"] / E,
             "exact": tot["exact"] / E, "selected": sel / E}
    fp = pruned_flop_per_eval(H, units["candidates"], units["in_window"], units["exact"])
    t = kernel_ms * 1e-3
    achieved = N * fp / t / 1e12
    gather = 4.0 * units["selected"] * N
    out = {"bound": "fp64-valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
           "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
           "achieved_note": ("FP64 operations of the pruned algorithm k_fast runs (omega solve per reflection, "
                             "window tests per candidate, spot construction per candidate inside the windows, exact "
                             "pipeline per undecided event), counted from the kernel source, per launch, over the "
                             "launch time"),
           "kernel": "k_fast", "kernel_ms": round(kernel_ms, 4), "evals_per_launch": N,
           "flop_per_eval": round(fp, 1), "units_per_eval": {k: round(v, 3) for k, v in units.items()},
           "audit_sample": f"{len(seeds)} timed seed voxels x {qa.shape[0]} orientations, "
                           f"{tot['mismatched']} disagreements with the exact pipeline",
           "gather_bytes_per_launch": int(gather), "gather_GB_per_s": round(gather / t / 1e9, 1),
           "forward_comp_equivalent_tflops": round(N * flop_per_eval(cfg.rot.size, H) / t / 1e12, 3),
           "forward_comp_flop_per_eval": flop_per_eval(cfg.rot.size, H)}
    prof = counter_profile()
    sha = kernel_sources_sha()
    out["build_sources_sha"] = sha
    if prof and config == "cfg2" and prof.get("sources_sha") == sha:
        ex = prof["fp64_ops_per_eval"]
        out["executed_fp64_per_eval"] = ex
        out["executed_tflops"] = round(N * ex / t / 1e12, 3)
        out["executed_frac"] = round(N * ex / t / 1e12 / PEAK_FP64_TFLOPS, 4)
        out["traffic"] = int(round(prof["hbm_bytes_per_eval"] * N))
        out["traffic_GB_per_s"] = round(out["traffic"] / t / 1e9, 1)
        out["traffic_over_gather"] = round(out["traffic"] / gather, 2)
        if "valu_insts_per_eval" in prof:
            # wave instructions x 4 cycles (64 lanes on a 16-wide SIMD) over
            # the 1,024 SIMDs at 2.4 GHz during the launch
            out["executed_valu_busy"] = round(N * prof["valu_insts_per_eval"] * 4 / (1024 * 2.4e9 * t), 3)
        out["counter_profile"] = "profiles/k_fast_seed_cfg2_counters.md (this build)"
    else:
        out["counter_profile"] = ("no counter profile of this build (profiles/k_fast_seed_cfg2_counters.md names "
                                  f"sources {prof.get('sources_sha') if prof else None})")
    log(f"[bench] roofline: {fp:.0f} FP64 ops per evaluation (pruned algorithm; units {units}) -> "
        f"{achieved:.3f} TFLOP/s = {achieved / PEAK_FP64_TFLOPS:.4f} of FP64 peak; counters: {out['counter_profile']}")
    return out


def cpu_baseline(cfg, proj_host, pos, q, gpu_out, seconds, threads):
    """Times the oracle on a bounded sample and checks the GPU results on it."""
    import oracle

    stack = oracle.Stack(proj_host)
    # the rate of a first sample (threads started, caches warm) sizes the timed one
    n0 = 1024
    oracle.or_space_comp(cfg.geo, stack, pos, q[:64], cfg.rot, cfg.hkl, nthreads=threads)
    t = time.perf_counter()
    oracle.or_space_comp(cfg.geo, stack, pos, q[:n0], cfg.rot, cfg.hkl, nthreads=threads)
    rate0 = n0 / (time.perf_counter() - t)
    n = int(min(q.shape[0], max(n0, rate0 * seconds)))
    t = time.perf_counter()
    ref = oracle.or_space_comp(cfg.geo, stack, pos, q[:n], cfg.rot, cfg.hkl, nthreads=threads)
    dt = time.perf_counter() - t
    ns, ni, C, med = (a[:n] for a in gpu_out)
    exact = (np.array_equal(ns, ref[:, 0]) and np.array_equal(ni, ref[:, 1])
             and np.array_equal(C.view(np.uint64), ref[:, 2].view(np.uint64))
             and np.array_equal(med.view(np.uint32), ref[:, 3].astype(np.float32).view(np.uint32)))
    return n / dt, n, dt, exact


def host_io_leg(eng, cfg, q, pos, calls, legacy_calls):
    """The same seeding call with transfers included: grs_forward_comp with
    host buffers (orientations up, results down every call), and the legacy
    single-precision entry grs_cuda_forward_comp, which like the reference
    wrapper (ForwardCalc_OR_space_comp_cuda.m:17-42) receives the geometry and
    the whole distance stack with every call and uploads them."""
    from grainrecon_amd import engine as E

    N = q.shape[0]
    eng.forward_comp(pos, q)
    t = time.perf_counter()
    for _ in range(calls):
        eng.forward_comp(pos, q)
    dt = time.perf_counter() - t
    res = {"host_buffers": {"value": round(calls * N / dt, 1), "unit": "evals/s",
                            "ms_per_call": round(dt / calls * 1e3, 3), "calls": calls,
                            "bytes_per_call": int(N * (32 + 20))}}
    log(f"[bench] host buffers: {calls} calls -> {calls * N / dt:.4g} evals/s ({dt / calls * 1e3:.2f} ms per call)")
    if legacy_calls > 0:
        geo, nrot, H = cfg.geo, cfg.rot.size, cfg.hkl.shape[0]
        proj = eng.get_distance_stack((int(geo.detzsize), int(geo.detysize), nrot)).reshape(-1, order="F")
        args = (np.asarray(pos, float), q.reshape(-1), np.asarray(geo.S).reshape(-1), np.asarray(geo.B).reshape(-1),
                cfg.hkl.reshape(-1), proj, cfg.rot, np.asarray(geo.RotDet).reshape(-1), geo.legacy_param(H, nrot))
        E.cuda_forward_comp(*args)
        t = time.perf_counter()
        for _ in range(legacy_calls):
            out = E.cuda_forward_comp(*args)
        ldt = time.perf_counter() - t
        res["legacy_f32"] = {"value": round(legacy_calls * N / ldt, 1), "unit": "evals/s",
                             "ms_per_call": round(ldt / legacy_calls * 1e3, 1), "calls": legacy_calls,
                             "stack_bytes_per_call": int(proj.nbytes),
                             "orientations_with_events": int((out.reshape(-1, 4)[:, 0] > 0).sum())}
        log(f"[bench] legacy f32 entry: {legacy_calls} calls -> {legacy_calls * N / ldt:.4g} evals/s "
            f"({ldt / legacy_calls * 1e3:.0f} ms per call, {proj.nbytes / 1e9:.2f} GB stack uploaded each)")
        del proj
    return res


def geofit_leg(eng, cfg, n_grains, nm_grains, seed=4004):
    """Geometry and grain refinement on matched spot centroids (0.5 px
    Gaussian noise, perturbed start): adjoint LM on the GPU and, when
    nm_grains > 0, the reference's procedure -- bounded Nelder-Mead over the 7
    geometry values with the grains held -- timed on the CPU on a subset of
    grains."""
    from grainrecon_amd import fit as F

    smp, geo = cfg.sample, cfg.geo
    rng = np.random.default_rng(seed)
    px = geo.pixelysize
    g7_true = np.array([geo.Lsam2sou, geo.Lsam2det, geo.dety00, geo.detz00, 0.0, 0.0, 0.0])
    G = min(n_grains, smp.q.shape[0])
    com = np.stack([smp.pos[smp.grain == g].mean(axis=0) for g in range(G)])
    gr6_true = np.hstack([smp.q[:G, 1:] / smp.q[:G, :1], com])
    sp = F.simulate_matched_spots(eng, g7_true, com, smp.U[:G], rng=rng, noise_px=0.5)
    hkl, omega = cfg.hkl[sp["hkl_idx"]], cfg.rot[sp["omega_idx"]]
    # start: Lsam2sou -0.2 mm, Lsam2det +0.3 mm, dety00/detz00 +-20 px, tilts +-0.2 deg;
    # grains as indexed (orientation ~0.01 deg, centre of mass ~2 um off)
    g7_0 = g7_true + np.array([-0.2, 0.3, 20 * px, -20 * px, 0.2, -0.2, 0.2])
    gr6_0 = gr6_true + np.hstack([rng.normal(scale=1e-4, size=(G, 3)), rng.normal(scale=2e-3, size=(G, 3))])

    def rounded(fit, r):
        return F.mean_rounded_px(r + fit.meas, fit.meas)

    full = F.SpotFit(eng, hkl, omega, sp["meas"], sp["grain"], G)
    r0 = full.residual(g7_0, gr6_0)
    eng.synchronize()
    joint = F.levenberg_marquardt(full, g7_0, gr6_0, max_iter=100)
    geo_only = F.levenberg_marquardt(full, g7_0, gr6_0, max_iter=100, fix_grains=True)
    rj = full.residual(joint.g7, joint.gr6)
    out = {"grains": G, "spots": int(full.n), "noise_px": 0.5, "start_rms_px": round(F.rms_px(r0), 3),
           "lm_joint_gpu": {"unknowns": 7 + 6 * G, "rms_px": round(joint.rms, 4),
                            "mean_rounded_px": round(rounded(full, rj), 4), "seconds": round(joint.seconds, 3),
                            "iterations": joint.iterations,
                            "geometry_error": [float(x) for x in np.round(joint.g7 - g7_true, 7)]},
           "lm_geometry_gpu": {"unknowns": 7, "rms_px": round(geo_only.rms, 4), "seconds": round(geo_only.seconds, 3),
                               "iterations": geo_only.iterations}}
    log(f"[bench] geo-fit: {G} grains, {full.n} spots, RMS {F.rms_px(r0):.2f} px at start; joint LM "
        f"({7 + 6 * G} unknowns) -> {joint.rms:.4f} px in {joint.seconds:.2f}s ({joint.iterations} it); "
        f"geometry-only LM -> {geo_only.rms:.4f} px in {geo_only.seconds:.2f}s")

    if nm_grains <= 0:
        return out
    # CPU baseline: the reference's procedure, FitAllOnce of L_shift_tilt_fitting.m:23-50
    # (GrainMapping_fit_geo.m:190-210, iter_fit = 1): bounded Nelder-Mead on geo_fitting.m's
    # mean rounded-pixel distance, re-centred while the result lies within 0.1 of a bound
    from oracle.geo_fit import GeoObjective, fit_all_once

    Gs = min(nm_grains, G)
    sel = sp["grain"] < Gs
    sub = F.SpotFit(eng, hkl[sel], omega[sel], sp["meas"][sel], sp["grain"][sel], Gs)
    threads = min(16, os.cpu_count() or 1)
    obj = GeoObjective(geo, sub.hkl, sub.omega, sub.meas, sub.grain, gr6_0[:Gs], nthreads=threads)
    t = time.perf_counter()
    x_nm, f_nm, nfe, rounds = fit_all_once(obj, g7_0, px)
    nm_t = time.perf_counter() - t
    nm_rms = obj.rms(x_nm)
    lm = F.levenberg_marquardt(sub, g7_0, gr6_0[:Gs], max_iter=100, fix_grains=True)
    lm_rounded = obj(lm.g7)
    out["fminsearchbnd_reference"] = {
        "grains": Gs, "spots": int(sub.n), "unknowns": 7,
        "cpu": {"rms_px": round(nm_rms, 4), "mean_rounded_px": round(f_nm, 4), "seconds": round(nm_t, 2),
                "fun_evals": nfe, "rounds": rounds, "threads": threads,
                "procedure": "FitAllOnce (L_shift_tilt_fitting.m:23-50) on geo_fitting.m, grains held"},
        "gpu_lm": {"rms_px": round(lm.rms, 4), "mean_rounded_px": round(lm_rounded, 4),
                   "seconds": round(lm.seconds, 3), "iterations": lm.iterations,
                   "procedure": "adjoint Levenberg-Marquardt on the unrounded residual, grains held"},
        "rms_gap_px": round(nm_rms - lm.rms, 4), "mean_rounded_gap_px": round(f_nm - lm_rounded, 4),
        "speedup": round(nm_t / max(lm.seconds, 1e-9), 1),
    }
    log(f"[bench] geometry fit on {Gs} grains ({sub.n} spots): FitAllOnce fminsearchbnd (CPU, {threads} threads) "
        f"{nm_rms:.4f} px RMS / {f_nm:.4f} px rounded mean in {nm_t:.1f}s ({nfe} evaluations, {rounds} rounds); "
        f"adjoint LM (GPU) {lm.rms:.4f} px RMS / {lm_rounded:.4f} px rounded mean in {lm.seconds:.3f}s")
    return out


def expand_local(qc, nr, g):
    """The engine's local-sampling grid for one candidate (q1 slowest; q0 is
    recomputed by the scorer), same IEEE operations as the device."""
    half = (g * nr) / 2.0
    k = np.arange(nr + 1, dtype=float) * g
    a, b, c = (qc[1] - half) + k, (qc[2] - half) + k, (qc[3] - half) + k
    A, B, C = np.meshgrid(a, b, c, indexing="ij")
    return np.stack([np.zeros(A.size), A.reshape(-1), B.reshape(-1), C.reshape(-1)], axis=1)


def batched_leg(eng, cfg, n_voxels, steps, warmup, seed=2003):
    """Batched compete with local sampling (CSR rows of neighbour orientations,
    each expanded to 7^3 samples spanning +-1.5 deg in 0.5 deg steps) in one
    persistent launch per batch."""
    smp = cfg.sample
    rng = np.random.default_rng(seed)
    G = smp.q.shape[0]
    nr, gridsize = 6, float(np.sin(np.deg2rad(0.25)))
    batches = []
    vox = rng.choice(smp.pos.shape[0], size=n_voxels * (steps + warmup), replace=False)
    for b in range(steps + warmup):
        v = vox[b * n_voxels:(b + 1) * n_voxels]
        k = np.minimum(rng.geometric(1.0 / 8.0, size=n_voxels), 64)
        rows, cands = [0], []
        for i, vi in enumerate(v):
            own = int(smp.grain[vi])
            others = rng.choice(np.delete(np.arange(G), own), size=k[i] - 1, replace=False)
            ids = np.sort(np.concatenate([[own], others]))   # find_neighbor_ORs order: ascending grain id
            cands.append(smp.q[ids])
            rows.append(rows[-1] + ids.size)
        batches.append((smp.pos[v], np.array(rows), np.vstack(cands)))
    S = (nr + 1) ** 3
    for b in range(warmup):
        eng.compete_csr(*batches[b], nr, gridsize)
    eng.synchronize()
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    outs = []
    for b in range(warmup, warmup + steps):
        outs.append(eng.compete_csr(*batches[b], nr, gridsize))
    eng.synchronize()
    dt = time.perf_counter() - t
    kms, kn = eng.kernel_time(reset=True)
    evals = sum(int(batches[b][1][-1]) * S for b in range(warmup, warmup + steps))
    C = np.concatenate([o[4] for o in outs])
    res = {"value": round(evals / (kms * 1e-3), 1), "unit": "evals/s", "wall_evals_per_s": round(evals / dt, 1),
           "evals": evals, "voxels_per_step": n_voxels, "steps": steps, "samples_per_candidate": S,
           "mean_candidates_per_voxel": round(evals / S / (n_voxels * steps), 3), "kernel_ms_total": round(kms, 3),
           "launches": kn, "winners_C1": int(np.sum(C == 1.0)), "kernel": "k_fast<counts-only, CSR>"}
    log(f"[bench] batched: {evals} evals in {kms:.1f} ms of scoring kernels ({res['value']:.4g} evals/s; wall "
        f"{dt:.2f}s incl. host CSR/winner pass), {res['winners_C1']} of {n_voxels * steps} voxels at C = 1")
    return res, batches[warmup], outs[0], (nr, gridsize)


def check_batched_rows(cfg, proj, batch, out, nrg, rows, threads):
    """Oracle argmax over the expanded lists of a few rows."""
    import oracle

    pos, rp, cq = batch
    nr, g = nrg
    stack = oracle.Stack(proj)
    ok = True
    for r in rows:
        qs = np.vstack([expand_local(cq[c], nr, g) for c in range(rp[r], rp[r + 1])])
        ref = oracle.or_space_comp(cfg.geo, stack, pos[r], qs, cfg.rot, cfg.hkl, nthreads=threads)
        Cr = ref[:, 2]
        j = int(np.nanargmax(Cr)) if np.isfinite(Cr).any() else 0
        best, _, ns, ni, C, med = out
        ok &= bool(best[r] == j and ns[r] == ref[j, 0] and ni[r] == ref[j, 1] and C[r] == Cr[j]
                   and np.float32(med[r]) == np.float32(ref[j, 3]))
    return ok


def al_leg(eng, steps, warmup, cpu_rows, threads, fit_grains):
    """Config 5: FCC Al, 1000 grains, 2.5 M voxels in 8 uneven subvolumes
    scored in one batched launch against their neighbour grains' orientations
    (find_neighbor_ORs on the grain-id volume, radius 2 voxels)."""
    from grainrecon_amd import synth

    t0 = time.perf_counter()
    cfg = synth.al_config()
    smp = cfg.sample
    slab = synth.subvolumes(smp, [0.02, 0.05, 0.08, 0.10, 0.13, 0.16, 0.20, 0.26])
    order = np.argsort(slab, kind="stable")
    gid = synth.grain_id_volume(smp)
    t_gen = time.perf_counter() - t0
    t = time.perf_counter()
    synth.load_on_engine(eng, cfg)
    eng.synchronize()
    t_prep = time.perf_counter() - t
    t = time.perf_counter()
    ids, ind = eng.find_neighbor_ors(gid, smp.index[order], 2, 30)
    t_nb = time.perf_counter() - t
    row_ptr = np.concatenate([[0], ind[:, 1]]).astype(np.int64)
    cand_q = smp.q[ids - 1]
    pos = smp.pos[order]
    M = pos.shape[0]
    for _ in range(warmup):
        eng.compete_csr(pos, row_ptr, cand_q, 0, 0.0)
    eng.synchronize()
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    for _ in range(steps):
        out = eng.compete_csr(pos, row_ptr, cand_q, 0, 0.0)
    eng.synchronize()
    wall = time.perf_counter() - t
    kms, kn = eng.kernel_time(reset=True)
    evals = int(row_ptr[-1]) * steps
    rate = evals / (kms * 1e-3)
    fe = flop_per_eval(cfg.rot.size, cfg.hkl.shape[0])
    best, bq, ns, ni, C, med = out
    # units of the pruned algorithm on this workload: the audit launch (one-voxel
    # kernel) over 8 of the voxels x the 1000 grain orientations; the counts-only
    # kernel runs the same omega solve, window tests and spot construction
    rng_a = np.random.default_rng(3105)
    tot = {"examined": 0, "in_window": 0, "exact": 0, "mismatched": 0}
    for r in rng_a.choice(M, size=8, replace=False):
        eng.set_audit(True)
        try:
            eng.forward_comp(pos[r], smp.q)
            a = eng.audit()
        finally:
            eng.set_audit(False)
        for k in tot:
            tot[k] += a[k]
    E = 8 * smp.q.shape[0]
    H = cfg.hkl.shape[0]
    units = {"reflections": H, "candidates": tot["examined"] / E, "in_window": tot["in_window"] / E,
             "exact": tot["exact"] / E}
    fp = pruned_flop_per_eval(H, units["candidates"], units["in_window"], units["exact"])
    own = smp.grain[order]
    winner_grain = ids[(row_ptr[:-1] + best).clip(0, ids.size - 1)] - 1
    res = {"voxels": int(M), "subvolumes": np.bincount(slab).tolist(), "grains": int(smp.q.shape[0]),
           "omega": int(cfg.rot.size), "hkl": int(cfg.hkl.shape[0]), "candidates": int(row_ptr[-1]),
           "evals_per_launch": int(row_ptr[-1]), "launches": kn, "kernel_ms_per_launch": round(kms / kn, 3),
           "value": round(rate, 1), "unit": "evals/s", "wall_voxels_per_s": round(M * steps / wall, 1),
           "roofline": {"bound": "fp64-valu", "achieved": round(rate * fp / 1e12, 3), "peak": PEAK_FP64_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(rate * fp / 1e12 / PEAK_FP64_TFLOPS, 4),
                        "flop_per_eval": round(fp, 1), "units_per_eval": {k: round(v, 3) for k, v in units.items()},
                        "audit_sample": f"8 voxels x {smp.q.shape[0]} grain orientations, "
                                        f"{tot['mismatched']} disagreements with the exact pipeline"},
           "forward_comp_equivalent_tflops": round(rate * fe / 1e12, 3), "forward_comp_flop_per_eval": fe,
           "own_grain_wins": round(float(np.mean(winner_grain == own)), 4),
           "setup_s": {"sample": round(t_gen, 2), "render_and_edt": round(t_prep, 2), "neighbour_lists": round(t_nb, 2)}}
    log(f"[bench] Al: {M} voxels, {row_ptr[-1]} candidates in one launch of {kms / kn:.1f} ms -> {rate:.4g} evals/s; "
        f"own grain wins on {res['own_grain_wins']:.4f} of voxels; neighbour lists {t_nb:.2f}s, render+EDT {t_prep:.1f}s")
    if cpu_rows > 0:
        import oracle

        shape = (int(cfg.geo.detzsize), int(cfg.geo.detysize), cfg.rot.size)
        stack = oracle.Stack(eng.get_distance_stack(shape))
        rng = np.random.default_rng(3005)
        rows = rng.choice(M, size=cpu_rows, replace=False)
        ok = True
        t = time.perf_counter()
        for r in rows:
            qs = cand_q[row_ptr[r]:row_ptr[r + 1]]
            ref = oracle.or_space_comp(cfg.geo, stack, pos[r], qs, cfg.rot, cfg.hkl, nthreads=threads)
            Cr = ref[:, 2]
            j = int(np.nanargmax(Cr)) if np.isfinite(Cr).any() else 0
            ok &= bool(best[r] == j and ns[r] == ref[j, 0] and ni[r] == ref[j, 1] and C[r] == Cr[j]
                       and np.float32(med[r]) == np.float32(ref[j, 3]))
        res["parity"] = {"rows_checked": int(cpu_rows), "argmax_and_counts_equal": ok}
        log(f"[bench] Al parity on {cpu_rows} rows: {ok} ({time.perf_counter() - t:.1f}s)")
    if fit_grains > 0:
        res["grain_refinement"] = geofit_leg(eng, cfg, fit_grains, 0)
    return res


def index_leg(eng, q, smp, n_pos, nr_seed, steps, seed=5005, turn_deg=0.05):
    """Seed indexing (GrainIndexing_engine_gpu_comp.m:46-66 with
    fit_OR_constrain_with_local_comp_cuda.m) of n_pos seed positions per call
    against the resident orientation set: device pipeline (scores of the set,
    top nr_seed, local sampling, scores, maxima) replayed as a HIP graph, then
    the batched Nelder-Mead fits and the verification counts.  The set is the
    random one plus every grain's orientation turned by `turn_deg` about a
    random axis (the completeness peak of this noise-free synthetic stack is
    narrower than the set's spacing: 0.3 degrees off the truth already scores
    at background level); an indexed orientation is counted as found when it
    lies within 0.1 degrees of its grain's (cubic symmetry), and separately
    when it lies that close to any grain's (a seed near a grain boundary
    scores the neighbour's orientation nearly as high)."""
    from grainrecon_amd import synth

    rng = np.random.default_rng(seed)
    G = smp.q.shape[0]
    ax = rng.normal(size=(G, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    h = np.deg2rad(turn_deg) / 2
    turn = np.column_stack([np.full(G, np.cos(h)), np.sin(h) * ax])
    qset = np.concatenate([q, synth.fz_reduce_cubic(synth._qmul(smp.q, turn))])
    eng.set_orientation_set(qset)
    idx = rng.choice(smp.pos.shape[0], size=(steps + 1) * n_pos, replace=False)
    picks = smp.pos[idx]
    c0, r0 = eng.graph_stats()
    t = time.perf_counter()
    eng.index_seeds(picks[:n_pos], nr_seed=nr_seed)  # first call: tuning of the grid kernel, graph capture
    first = time.perf_counter() - t
    eng.synchronize()
    res = []
    t = time.perf_counter()
    for k in range(steps):
        res.append(eng.index_seeds(picks[(k + 1) * n_pos:(k + 2) * n_pos], nr_seed=nr_seed))
    eng.synchronize()
    dt = time.perf_counter() - t
    c1, r1 = eng.graph_stats()
    C = np.concatenate([r["indexing"][:, 6] for r in res])
    nf = np.concatenate([r["n_fun_evals"] for r in res])
    qfit = np.concatenate([r["indexing"][:, 14:18] for r in res])
    qtrue = smp.q[smp.grain[idx[n_pos:]]]
    sym = synth.cubic_symmetry_quats()
    dots = np.abs(np.einsum("nk,nsk->ns", qfit, synth._qmul(qtrue[:, None, :], sym[None, :, :]))).max(1)
    mis = np.degrees(2 * np.arccos(np.clip(dots, -1, 1)))
    allq = synth._qmul(smp.q[:, None, :], sym[None, :, :]).reshape(-1, 4)
    mis_any = np.degrees(2 * np.arccos(np.clip(np.abs(qfit @ allq.T).max(1), -1, 1)))
    N = qset.shape[0]
    evals_dev = steps * n_pos * (N + nr_seed * 730)
    # why a seed misses its own grain: the completeness of its own grain's
    # orientation at the seed position against the fitted one, and how far the
    # seed lies from its grain's boundary (the distance to the nearest Voronoi
    # bisector plane)
    pos_s = smp.pos[idx[n_pos:]]
    own = smp.grain[idx[n_pos:]]
    own_C = np.array([float(eng.forward_comp(p, smp.q[g][None, :])[2][0]) for p, g in zip(pos_s, own)])
    s1 = smp.seeds[own]
    d_b = np.full(pos_s.shape[0], np.inf)
    for k in range(pos_s.shape[0]):
        s2 = np.delete(smp.seeds, own[k], axis=0)
        n2 = np.linalg.norm(s2 - s1[k], axis=1)
        d_b[k] = float(np.min((np.sum((pos_s[k] - s2) ** 2, 1) - np.sum((pos_s[k] - s1[k]) ** 2)) / (2 * n2)))
    miss = mis >= 0.1
    interior = d_b > 0.02
    # what the procedure sees of the own grain: the best completeness on the
    # local-sampling grid (nr 8, step 0.004) around the own grain's entry of the
    # set, against the completeness of the start the fit of the chosen group
    # began from (the best grid sample of the winning seed)
    x0 = np.concatenate([r["x0"] for r in res])
    own_grid_best = np.full(mis.size, np.nan)
    start_C = np.full(mis.size, np.nan)
    for k in np.flatnonzero(miss):
        ls = eng.local_sampling(qset[q.shape[0] + own[k]][None, :], 8, 0.004)[0]
        own_grid_best[k] = float(np.nanmax(eng.forward_comp(pos_s[k], ls)[2]))
        v = x0[k]
        start_C[k] = float(eng.forward_comp(pos_s[k], np.concatenate([[np.sqrt(max(0.0, 1 - v @ v))], v])[None, :])[2][0])
    out = {"positions_per_call": n_pos, "calls": steps, "nr_seed": nr_seed, "orientations": N,
           "ms_per_call": round(dt / steps * 1e3, 2), "seeds_per_s": round(steps * n_pos / dt, 2),
           "first_call_s": round(first, 3), "graph_replays": int(r1 - r0), "graph_captures": int(c1 - c0),
           "device_evals_per_call": int(evals_dev / steps), "fit_evals_mean": round(float(nf.mean()), 1),
           "fitted_C_median": round(float(np.nanmedian(C)), 4), "fitted_C_min": round(float(np.nanmin(C)), 4),
           "seeds": int(mis.size), "own_grain_within_0.1deg": int((~miss).sum()),
           "interior_seeds": int(interior.sum()),
           "interior_own_grain_within_0.1deg": int((~miss & interior).sum()),
           "misses_with_fitted_C_ge_own_grain_C": int((miss & (C >= own_C)).sum()),
           "misses_with_fitted_C_below_own_grain_C": int((miss & (C < own_C)).sum()),
           "misses_whose_own_grid_best_C_is_below_the_chosen_start_C": int((miss & (own_grid_best < start_C)).sum()),
           "misses_own_grid_best_C": [round(float(v), 4) for v in own_grid_best[miss]],
           "misses_chosen_start_C": [round(float(v), 4) for v in start_C[miss]],
           "any_grain_within_0.1deg": int((mis_any < 0.1).sum()),
           "boundary_distance_mm_of_misses": [round(float(v), 4) for v in d_b[miss]],
           "misorientation_deg_median": round(float(np.median(mis)), 4), "set_turn_deg": turn_deg,
           "note": ("the completeness peak of this noise-free stack is narrower than the local-sampling step "
                    "(0.004, about 0.46 deg): a miss is a seed whose own grain's best grid sample scores below the "
                    "start the procedure chose (a neighbour grain's sample nearer its own peak), so the fit of "
                    "fit_OR_constrain_with_local_comp_cuda.m starts on the neighbour; own_grain C is the true "
                    "orientation's, which no grid sample reaches; interior seeds lie more than 20 um from every "
                    "grain boundary")}
    log(f"[bench] index: {steps} calls x {n_pos} seeds in {dt:.3f}s -> {out['seeds_per_s']} seeds/s, "
        f"median fitted C {out['fitted_C_median']}, own grain within 0.1 deg: {out['own_grain_within_0.1deg']}/"
        f"{mis.size} (interior {out['interior_own_grain_within_0.1deg']}/{out['interior_seeds']}), misses with "
        f"fitted C >= own grain's C: {out['misses_with_fitted_C_ge_own_grain_C']}, graph replays "
        f"{out['graph_replays']}")
    return out


def _grain_com(smp):
    G = smp.q.shape[0]
    cnt = np.bincount(smp.grain, minlength=G).astype(float)
    return np.stack([np.bincount(smp.grain, weights=smp.pos[:, k], minlength=G) for k in range(3)], 1) / cnt[:, None]


def gt_leg(eng, cfg, n_orient, cpu, threads, seed=6006):
    """G-vector matching score (calc_Gt_match, ForwardCalc_OR_space_Gt.m:117-205)
    at one grain's centre of mass: spot centroids of every grain simulated
    with the engine (one spot per selected event at its unrounded pixel
    position), n_orient orientations (the grain's local sampling and random
    ones); the calc_Gt_match_fun.m semantics timed beside it."""
    from grainrecon_amd import fit as F
    from grainrecon_amd import synth

    smp, rng = cfg.sample, np.random.default_rng(seed)
    com = _grain_com(smp)
    sp = F.simulate_matched_spots(eng, None, com, smp.U)
    nrot = cfg.rot.size
    frames = []
    for i in range(nrot):
        yz = sp["meas"][sp["omega_idx"] == i]
        frames.append(np.column_stack([yz, np.full(yz.shape[0], 20.0)]))
    eng.set_frame_spots(frames)
    B = np.asarray(cfg.geo.B)
    sq, dposs = [], []
    for h in cfg.hkl:
        v = int(round((h * h).sum()))
        if v not in sq:
            sq.append(v)
            dposs.append(2 * np.pi / np.linalg.norm(B @ h))
    sq, dposs = np.array(sq[:4], float), np.array(dposs[:4])
    g = 7
    qt = synth.fz_reduce_cubic(smp.q[g:g + 1])[0]
    local = eng.local_sampling(qt, 8, 0.004)[0]
    q = np.concatenate([local, synth.fz_reduce_cubic(synth.haar_quaternions(rng, n_orient - local.shape[0]))])
    eng.kernel_time(reset=True)
    out = eng.forward_gt(com[g], q, sq, dposs)
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    out = eng.forward_gt(com[g], q, sq, dposs)
    wall = time.perf_counter() - t
    kms, _ = eng.kernel_time(reset=True)
    from grainrecon_amd import _lib

    t = time.perf_counter()
    out_fun = eng.forward_gt(com[g], q, sq, dposs, _lib.GT_MATCH_FUN)
    wall_fun = time.perf_counter() - t
    kms_fun, _ = eng.kernel_time(reset=True)
    best = int(np.argmax(out[:, 0]))
    res = {"orientations": int(q.shape[0]), "spots": int(sum(f.shape[0] for f in frames)),
           "kernel_ms": round(kms, 3), "wall_s": round(wall, 3),
           "orientations_per_s": round(q.shape[0] / (kms * 1e-3), 1),
           "best_in_local_sampling": bool(best < local.shape[0]), "best_Nr_match": int(out[best, 0]),
           "best_matched_cost": float(out[best, 1]), "best_mean_angle_deg": float(out[best, 3]),
           "calc_Gt_match_fun": {"kernel_ms": round(kms_fun, 3), "wall_s": round(wall_fun, 3),
                                 "best_median_angle_deg": float(out_fun[best, 1])}}
    if cpu:
        from oracle import gt_match

        pick = [best, local.shape[0] + 3]
        t = time.perf_counter()
        ref = gt_match.gt_score(cfg.geo, com[g], q[pick], cfg.rot, cfg.hkl, frames, sq, dposs, semantics="match")
        cdt = time.perf_counter() - t
        ref_fun = gt_match.gt_score(cfg.geo, com[g], q[pick], cfg.rot, cfg.hkl, frames, sq, dposs, semantics="fun")
        same = (out[pick] == ref) | (np.isnan(out[pick]) & np.isnan(ref))
        same_fun = (out_fun[pick] == ref_fun) | (np.isnan(out_fun[pick]) & np.isnan(ref_fun))
        res["cpu_baseline"] = {"value": round(len(pick) / cdt, 3), "unit": "orientations/s", "cores": 1,
                               "kind": "port", "sample": f"{len(pick)} orientations (oracle/gt_match.py, numpy)"}
        res["parity"] = {"orientations_checked": len(pick), "bit_exact": bool(same.all()),
                         "calc_Gt_match_fun_bit_exact": bool(same_fun.all())}
    log(f"[bench] gt: {q.shape[0]} orientations x {res['spots']} spots in {kms:.2f} ms of kernel time; best row "
        f"{best} (Nr_match {res['best_Nr_match']})")
    return res


def audit_leg(eng, smp, q, n_evals, n_seeds=2, seed=3002):
    """Resolution-stage audit of the seeding kernel: for n_seeds seed voxels
    x the first n_evals orientations, the audit launch also runs the exact
    forward_comp.m pipeline on every event the resolution stage decides and
    counts candidates, decisions, disagreements and the largest coordinate
    difference; the counts and C must equal those of the plain launch."""
    qa = q[:n_evals]
    pos = smp.pos[np.random.default_rng(seed).choice(smp.pos.shape[0], n_seeds, replace=False)]
    rows = []
    for k, p in enumerate(pos):
        ref = eng.forward_comp(p, qa)
        eng.set_audit(True)
        try:
            out = eng.forward_comp(p, qa)
            a = dict(eng.audit())
        finally:
            eng.set_audit(False)
        same = all(bool(np.array_equal(x, y, equal_nan=True)) for x, y in zip(out, ref))
        per = {kk: (v / qa.shape[0] if kk != "max_dev_px" else v) for kk, v in a.items()}
        per["equal_to_plain_launch"] = same
        rows.append(per)
        log(f"[bench] audit seed {k}: {json.dumps(per)}")
    return {"evals_per_seed": int(qa.shape[0]), "seeds": rows, "note": "per-evaluation averages"}


def mapping_leg(eng, cfg, q, n_seeds, nr_seed, seed=8008, refine=("nelder-mead", "adjoint")):
    """The seed -> index -> pair -> centre-of-mass fit -> grow -> re-centre
    loop of GrainMapping_engine_gpu_comp.m (grainrecon_amd.mapping) on the
    Fe 1,000-grain sample: seeds at random voxels, the 262,144-orientation set
    plus the grains' orientations turned by 0.05 degrees, spot centroids of
    all grains with an area of 800 px^2 (a grain-sized spot, so the growth box
    spans about 25 voxels).  Run once per refine mode from the same empty
    volumes: the reference's Nelder-Mead fits, then the adjoint LM fits."""
    from grainrecon_amd import fit as F
    from grainrecon_amd import synth
    from grainrecon_amd.mapping import MappingDS, grain_mapping_engine

    smp, geo = cfg.sample, cfg.geo
    rng = np.random.default_rng(seed)
    com = _grain_com(smp)
    sp = F.simulate_matched_spots(eng, None, com, smp.U)
    frames = []
    for i in range(cfg.rot.size):
        yz = sp["meas"][sp["omega_idx"] == i]
        frames.append(np.column_stack([yz, np.full(yz.shape[0], 800.0)]))
    eng.set_frame_spots(frames)
    G = smp.q.shape[0]
    ax = rng.normal(size=(G, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    h = np.deg2rad(0.05) / 2
    turn = np.column_stack([np.full(G, np.cos(h)), np.sin(h) * ax])
    qset = np.concatenate([q, synth.fz_reduce_cubic(synth._qmul(smp.q, turn))])
    nx, ny, nz = smp.dims
    mask = np.zeros((nx, ny, nz))
    mask[smp.index[:, 0] - 1, smp.index[:, 1] - 1, smp.index[:, 2] - 1] = 1.0
    picks = smp.pos[rng.choice(smp.pos.shape[0], size=n_seeds, replace=False)]
    gid = np.full((nx, ny, nz), -1, np.int32)
    gid[smp.index[:, 0] - 1, smp.index[:, 1] - 1, smp.index[:, 2] - 1] = smp.grain
    geo7 = np.array([geo.Lsam2sou, geo.Lsam2det, geo.dety00, geo.detz00, 0.0, 0.0, 0.0])
    sym = synth.cubic_symmetry_quats()
    res = {"seeds": n_seeds, "nr_seed": nr_seed}
    for mode in refine:
        ds = MappingDS(mask)
        fit_log = []
        t = time.perf_counter()
        out, _, nvox, _ = grain_mapping_engine(eng, picks, qset, ds, trust_comp=0.9, min_comp=0.6, drop_off=0.05,
                                            maxD=2.0, max_dmedian=10.0, nr_seed=nr_seed, first_grain_id=1,
                                            center=[0, 0, 0], vox_size=smp.vox_mm, rvp_first=[1, 1, 1],
                                            tomo_dim=(nx, ny, nz), refine=mode, geo7=geo7, fit_log=fit_log)
        wall = time.perf_counter() - t
        grown = ds.GrainId > 0
        r = {"wall_s": round(wall, 3), "seeds_per_s": round(n_seeds / wall, 2),
             "seeds_indexed": int((out[:, 6] >= 0.6).sum()), "voxels_assigned": int(grown.sum()),
             "median_C": round(float(np.nanmedian(out[:, 6])), 4)}
        # the assigned voxels' true grains: the share in the grain most assigned per seed id
        same = 0
        for sid in np.unique(ds.GrainId[grown]):
            g = gid[ds.GrainId == sid]
            g = g[g >= 0]
            if g.size:
                same += int(np.bincount(g).max())
        r["voxels_in_majority_grain"] = same
        # orientation error of the indexed rows against the true grain at the row's position
        err = []
        for row in out[out[:, 6] >= 0.6]:
            vx = np.rint((row[1:4] - smp.pos[0]) / smp.vox_mm).astype(int) + (smp.index[0] - 1)
            if np.all(vx >= 0) and np.all(vx < smp.dims) and gid[tuple(vx)] >= 0:
                qt = smp.q[gid[tuple(vx)]]
                d = np.abs(synth._qmul(np.repeat(qt[None, :], sym.shape[0], 0), sym) @ row[14:18]).max()
                err.append(np.degrees(2 * np.arccos(min(1.0, d))))
        r["orientation_error_deg_median"] = round(float(np.median(err)), 6) if err else None
        if fit_log:
            r["lm_fits"] = len(fit_log)
            r["lm_rms_px_median"] = float(np.nanmedian([e["rms_px"] for e in fit_log]))
        res[mode] = r
        log(f"[bench] mapping ({mode}): {n_seeds} seeds in {wall:.2f}s ({r['seeds_per_s']} seeds/s), "
            f"{r['voxels_assigned']} voxels assigned, {same} in their seed's majority grain, median orientation "
            f"error {r['orientation_error_deg_median']} deg")
    return res


def grow_leg(eng, cfg, box_l, cpu, threads, seed=7007):
    """Region growth (grow_indexed_region_cuda.m) around a voxel of a grain
    with the reconstruction volumes resident on the device, bounding sphere of
    radius box_l voxels."""
    smp, geo = cfg.sample, cfg.geo
    nx, ny, nz = smp.dims
    rng = np.random.default_rng(seed)
    vol_c = np.zeros((nx, ny, nz))
    mask = np.zeros((nx, ny, nz))
    mask[smp.index[:, 0] - 1, smp.index[:, 1] - 1, smp.index[:, 2] - 1] = 1.0
    dm = np.full((nx, ny, nz), 10.0)
    eng.set_volume("Completeness", vol_c)
    eng.set_volume("Mask", mask)
    eng.set_volume("Dismedian", dm)
    eng.set_volume("Ninter", np.zeros((nx, ny, nz)))
    p1 = np.array([1, 1, 1, 0, smp.vox_mm, 0, 0, 0, nx, ny, nz], float)
    com = _grain_com(smp)
    g = 11
    k = int(np.argmin(((smp.pos - com[g]) ** 2).sum(1)))
    seed_vox = smp.index[k]
    spot_l = box_l * ((geo.Lsam2det / geo.Lsam2sou) * smp.vox_mm) / ((geo.pixelysize + geo.pixelzsize) / 2)
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    r = eng.grow_region(seed_vox, 1.0, smp.U[g], spot_l, 0.05, 2.0, 10.0, p1)
    wall = time.perf_counter() - t
    kms, _ = eng.kernel_time(reset=True)
    gid = np.full((nx, ny, nz), -1, np.int32)
    gid[smp.index[:, 0] - 1, smp.index[:, 1] - 1, smp.index[:, 2] - 1] = smp.grain
    iv = r["indexed"] - 1
    n_ok = int((gid[iv[:, 0], iv[:, 1], iv[:, 2]] == smp.grain[k]).sum())
    res = {"voxels_checked": r["n_checked"], "voxels_grown": int(r["indexed"].shape[0]),
           "grown_in_own_grain": n_ok, "box_half_width": box_l, "scoring_kernel_ms": round(kms, 3),
           "wall_s": round(wall, 3), "voxels_per_s": round(r["n_checked"] / wall, 1),
           "replace_center": r["replace_center"]}
    if cpu:
        import oracle

        rows = []
        for v in r["indexed"][:64].tolist():
            rows.append([*v, 0.0, 1.0, 10.0])
        rows = np.array(rows, float)
        param1 = np.concatenate([p1, [1.0, 0.05, 10.0]])
        stack = oracle.Stack(eng.get_distance_stack((int(geo.detzsize), int(geo.detysize), cfg.rot.size)))
        t = time.perf_counter()
        ref = oracle.grow(geo, stack, rows, smp.U[g], cfg.rot, cfg.hkl, param1, nthreads=threads)
        cdt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": round(rows.shape[0] / cdt, 2), "unit": "voxels/s", "cores": threads,
                               "kind": "port", "sample": f"{rows.shape[0]} grown voxels re-scored by oracle.grow"}
        res["parity"] = {"voxels_checked": int(rows.shape[0]), "all_accepted_by_oracle": bool((ref[:, 6] == 1).all())}
    log(f"[bench] grow: {r['n_checked']} voxels checked, {r['indexed'].shape[0]} grown ({n_ok} in the seed's "
        f"grain) in {wall:.3f}s")
    return res


def growth_all_leg(eng, cfg, n_ors, cpu, threads, seed=7107):
    """Config 2's growth stage: an indexed orientation scored against every
    voxel of the volume (cuda_grow_indexed_region over the 3.2M voxels,
    grow_indexed_region_cuda.m:105-107 with the whole volume as the list) for
    n_ors grains' orientations; kernel time by HIP events, 256 voxels of the
    first orientation re-scored by the oracle."""
    smp, geo = cfg.sample, cfg.geo
    nx, ny, nz = smp.dims
    M = smp.index.shape[0]
    rows = np.column_stack([smp.index.astype(float), np.zeros(M), np.ones(M), np.full(M, 10.0)])
    param1 = np.array([1, 1, 1, 0, smp.vox_mm, 0, 0, 0, nx, ny, nz, 1.0, 0.05, 10.0], float)
    rng = np.random.default_rng(seed)
    grains = rng.choice(smp.q.shape[0], size=n_ors, replace=False)
    eng.grow_indexed_region(rows[:4096], smp.U[grains[0]], param1)  # warm-up
    eng.synchronize()
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    first = None
    for g in grains:
        out = eng.grow_indexed_region(rows, smp.U[g], param1)
        if first is None:
            first = out
    wall = time.perf_counter() - t
    kms, kn = eng.kernel_time(reset=True)
    rate = n_ors * M / (kms * 1e-3)
    res = {"orientations": int(n_ors), "voxels": int(M), "evals": int(n_ors * M), "kernel_ms_per_orientation":
           round(kms / max(kn, 1), 2), "value": round(rate, 1), "unit": "evals/s",
           "wall_s_per_orientation": round(wall / n_ors, 3),
           "note": "wall includes the 154 MB of voxel rows up and 180 MB of results down per call"}
    if cpu:
        import oracle

        sel = rng.choice(M, size=256, replace=False)
        stack = oracle.Stack(eng.get_distance_stack((int(geo.detzsize), int(geo.detysize), cfg.rot.size)))
        ref = oracle.grow(geo, stack, rows[sel], smp.U[grains[0]], cfg.rot, cfg.hkl, param1, nthreads=threads)
        got = first[sel]
        ok = bool(np.array_equal(got[:, 3:5], ref[:, 3:5]) and np.array_equal(
            got[:, 5].astype(np.float32).view(np.uint32), ref[:, 5].astype(np.float32).view(np.uint32))
                  and np.array_equal(got[:, 6], ref[:, 6]))
        res["parity"] = {"voxels_checked": 256, "counts_median_flag_equal": ok}
    log(f"[bench] growth against all {M} voxels: {n_ors} orientations, {kms / max(kn, 1):.1f} ms of kernel each "
        f"-> {rate:.4g} evals/s (wall {wall / n_ors:.3f} s per orientation incl. transfers)")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--orientations", type=int, default=262144)
    ap.add_argument("--config", default="cfg2", choices=["cfg1", "cfg2"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-voxels", type=int, default=16384, help="voxels per batched compete step; 0 = off")
    ap.add_argument("--batched-steps", type=int, default=2)
    ap.add_argument("--growth-ors", type=int, default=4,
                    help="orientations scored against every voxel in the growth leg; 0 = off")
    ap.add_argument("--al", type=int, default=1, help="run the FCC Al 2.5 M-voxel leg (1) or not (0)")
    ap.add_argument("--al-steps", type=int, default=2)
    ap.add_argument("--al-rows", type=int, default=16, help="Al voxels checked against the oracle")
    ap.add_argument("--host-io-calls", type=int, default=3, help="seeding calls with host buffers; 0 = off")
    ap.add_argument("--legacy-calls", type=int, default=2, help="calls of the legacy f32 entry; 0 = off")
    ap.add_argument("--geo-fit-grains", type=int, default=300, help="grains in the config-4 fit leg; 0 = off")
    ap.add_argument("--geo-fit-nm-grains", type=int, default=300,
                    help="grains in the comparison with the reference FitAllOnce Nelder-Mead procedure; 0 = off")
    ap.add_argument("--index-positions", type=int, default=8, help="seed positions per indexing call; 0 = off")
    ap.add_argument("--index-steps", type=int, default=4)
    ap.add_argument("--index-nr-seed", type=int, default=500)
    ap.add_argument("--gt-orientations", type=int, default=8192, help="orientations of the Gt-score leg; 0 = off")
    ap.add_argument("--grow-box", type=int, default=24, help="sphere radius (voxels) of the region-growth leg; 0 = off")
    ap.add_argument("--mapping-seeds", type=int, default=6, help="seeds of the mapping-loop leg; 0 = off")
    ap.add_argument("--roofline-audit-evals", type=int, default=32768,
                    help="orientations per seed of the audit launch that counts the roofline's units")
    ap.add_argument("--audit-evals", type=int, default=0,
                    help="orientations per seed of the resolution-stage audit leg; 0 = off")
    ap.add_argument("--copy-probe-mb", type=int, default=256,
                    help="size of a plain device copy timed (and visible to counter runs) for calibration; 0 = off")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by voxel or "
              "orientation with no exchange beyond gathering results)", file=sys.stderr)
        sys.exit(2)

    from grainrecon_amd import _lib
    from grainrecon_amd.engine import Engine
    from grainrecon_amd import synth

    t0 = time.perf_counter()
    cfg = synth.fe_config(args.config)
    smp = cfg.sample
    log(f"[bench] sample: {smp.pos.shape[0]} voxels, {smp.q.shape[0]} grains ({time.perf_counter() - t0:.1f}s)")
    eng = Engine(0)
    t = time.perf_counter()
    synth.load_on_engine(eng, cfg)
    eng.synchronize()
    log(f"[bench] spots rendered + distance transform on GPU in {time.perf_counter() - t:.1f}s")

    rng = np.random.default_rng(2002)
    q = synth.fz_reduce_cubic(synth.haar_quaternions(rng, args.orientations))
    N = q.shape[0]
    d_q = eng.to_device(q)
    d_ns = eng.device_alloc(4 * N)
    d_ni = eng.device_alloc(4 * N)
    d_C = eng.device_alloc(8 * N)
    d_med = eng.device_alloc(4 * N)
    srng = np.random.default_rng(3002)
    total = args.warmup + args.steps + args.plain_steps
    seeds = smp.pos[srng.choice(smp.pos.shape[0], size=total, replace=False)]

    eng.set_timing(True)
    copy_probe = None
    if args.copy_probe_mb > 0:
        nb = args.copy_probe_mb << 20
        a_, b_ = eng.device_alloc(nb), eng.device_alloc(nb)
        eng.device_copy(b_, a_, nb)  # first touch
        eng.kernel_time(reset=True)
        eng.device_copy(b_, a_, nb)
        cms, _ = eng.kernel_time(reset=True)
        eng.device_free(a_)
        eng.device_free(b_)
        copy_probe = {"bytes_read": nb, "bytes_written": nb, "ms": round(cms, 4),
                      "GB_per_s": round(2 * nb / (cms * 1e-3) / 1e9, 1)}
        log(f"[bench] copy probe: {2 * nb / 1e6:.0f} MB moved in {cms:.3f} ms ({copy_probe['GB_per_s']} GB/s)")
    for k in range(args.warmup):
        eng.forward_comp_device(seeds[k], d_q, N, d_ns, d_ni, d_C, d_med, _lib.METHOD_FAST)
    eng.synchronize()
    eng.kernel_time(reset=True)
    t = time.perf_counter()
    for k in range(args.steps):
        eng.forward_comp_device(seeds[args.warmup + k], d_q, N, d_ns, d_ni, d_C, d_med, _lib.METHOD_FAST)
    eng.synchronize()
    dt = time.perf_counter() - t
    kms, kn = eng.kernel_time(reset=True)
    value = args.steps * N / dt
    log(f"[bench] fast: {args.steps} steps in {dt:.3f}s -> {value:.4g} evals/s; kernel {kms / kn:.3f} ms/launch")

    # plain kernel on the same workload
    t = time.perf_counter()
    for k in range(args.plain_steps):
        eng.forward_comp_device(seeds[args.warmup + args.steps + k], d_q, N, d_ns, d_ni, d_C, d_med,
                                _lib.METHOD_PLAIN)
    eng.synchronize()
    pdt = time.perf_counter() - t
    pms, pn = eng.kernel_time(reset=True)
    plain_value = args.plain_steps * N / pdt if args.plain_steps else float("nan")
    log(f"[bench] plain: {args.plain_steps} steps in {pdt:.3f}s -> {plain_value:.4g} evals/s; "
        f"kernel {pms / max(pn, 1):.3f} ms/launch")

    nrot, H = cfg.rot.size, cfg.hkl.shape[0]
    roofline = roofline_leg(eng, cfg, q, seeds[args.warmup:args.warmup + 2], kms / kn, N, args.config,
                            args.roofline_audit_evals)

    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": (f"synthetic: forward-simulated BCC Fe cylinder ({smp.pos.shape[0]} voxels, {smp.q.shape[0]} "
                 f"Voronoi grains), {int(cfg.geo.detzsize)}^2 x {nrot} distance stack built on the GPU, "
                 f"{N}-orientation cubic fundamental-zone set"),
        "config": {"workload": f"seeding call of cuda_forward_comp at {args.config}: 1 seed voxel x {N} "
                               f"orientations per step", "voxels": int(smp.pos.shape[0]), "grains": int(smp.q.shape[0]),
                   "orientations": N, "omega": nrot, "hkl": H, "detector": int(cfg.geo.detzsize)},
        "plain": {"value": round(plain_value, 1), "unit": "evals/s", "kernel_ms": round(pms / max(pn, 1), 3),
                  "steps": args.plain_steps},
        "speedup_vs_plain": round(value / plain_value, 2) if args.plain_steps else None,
        "roofline": roofline,
    }
    if copy_probe:
        result["copy_probe"] = copy_probe
    if args.host_io_calls > 0:
        result["with_transfers"] = host_io_leg(eng, cfg, q, seeds[0], args.host_io_calls, args.legacy_calls)
    batched = None
    if args.batched_voxels > 0:
        bres, bbatch, bout, nrg = batched_leg(eng, cfg, args.batched_voxels, args.batched_steps, 1)
        if args.plain_steps:
            bres["speedup_vs_plain"] = round(bres["value"] / plain_value, 2)
        result["batched"] = bres
        batched = (bbatch, bout, nrg)

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        pos0 = seeds[args.warmup]
        qn = q[:65536]
        gout = eng.forward_comp(pos0, qn, _lib.METHOD_FAST)
        shape = (int(cfg.geo.detzsize), int(cfg.geo.detysize), nrot)
        proj = eng.get_distance_stack(shape)
        rate, n, cdt, exact = cpu_baseline(cfg, proj, pos0, qn, gout, args.cpu_seconds, threads)
        log(f"[bench] cpu oracle: {n} evals in {cdt:.1f}s on {threads} threads -> {rate:.4g} evals/s; "
            f"GPU bit-exact on that sample: {exact}")
        result["cpu_baseline"] = {"value": round(rate, 2), "unit": "evals/s", "cores": threads, "kind": "port",
                                  "sample": f"{n} orientations x 1 seed voxel of the same workload "
                                            f"(oracle/grs_oracle.c, gcc -O2, OpenMP)"}
        result["parity"] = {"evals_checked": n, "bit_exact": bool(exact)}
        # the same oracle on one core, on a smaller sample
        import oracle

        n1 = 1024
        t1 = time.perf_counter()
        oracle.or_space_comp(cfg.geo, oracle.Stack(proj), pos0, qn[:n1], cfg.rot, cfg.hkl, nthreads=1)
        r1 = n1 / (time.perf_counter() - t1)
        result["cpu_baseline_1core"] = {"value": round(r1, 2), "unit": "evals/s", "cores": 1, "kind": "port",
                                        "sample": f"{n1} orientations of the same sample"}
        log(f"[bench] cpu oracle on 1 core: {r1:.4g} evals/s")
        if batched:
            bbatch, bout, nrg = batched
            sizes = np.diff(bbatch[1])
            # the longest row and five rows drawn at random
            longest = int(np.argmax(sizes))
            rest = np.delete(np.arange(sizes.size), longest)
            rows = [longest] + [int(r) for r in np.random.default_rng(3003).choice(rest, size=5, replace=False)]
            t = time.perf_counter()
            ok = check_batched_rows(cfg, proj, bbatch, bout, nrg, rows, threads)
            result["batched"]["parity"] = {"rows_checked": rows, "evals_checked": int(sizes[rows].sum()) * 343,
                                           "argmax_and_counts_equal": ok}
            log(f"[bench] batched parity on rows {rows}: {ok} ({time.perf_counter() - t:.1f}s)")
    for d in (d_q, d_ns, d_ni, d_C, d_med):
        eng.device_free(d)
    if args.index_positions > 0:
        result["index"] = index_leg(eng, q, smp, args.index_positions, args.index_nr_seed, args.index_steps)
    threads = min(16, os.cpu_count() or 1)
    if args.gt_orientations > 0:
        result["gt_score"] = gt_leg(eng, cfg, args.gt_orientations, not args.no_cpu, threads)
    if args.grow_box > 0:
        result["grow_region"] = grow_leg(eng, cfg, args.grow_box, not args.no_cpu, threads)
    if args.growth_ors > 0:
        result["growth_all_voxels"] = growth_all_leg(eng, cfg, args.growth_ors, not args.no_cpu, threads)
    # the whole jobs of configs 2 and 3 at the measured rates
    est = {}
    n_vox = int(smp.pos.shape[0])
    est["config2_seeding_10k_seeds_s"] = round(1e4 * N / value, 1)
    if "growth_all_voxels" in result:
        est["config2_growth_300_ors_all_voxels_s"] = round(smp.q.shape[0] * n_vox / result["growth_all_voxels"]["value"], 1)
    if "batched" in result:
        b = result["batched"]
        evals3 = n_vox * b["mean_candidates_per_voxel"] * b["samples_per_candidate"]
        est["config3_evals"] = int(evals3)
        est["config3_all_voxels_s"] = round(evals3 / b["value"], 1)
    est["note"] = ("measured rates extrapolated to the full configurations: seeding 10k seed voxels x the "
                   f"{N}-orientation set; growth every grain's orientation against all {n_vox} voxels; batched "
                   "compete over every voxel's candidate list")
    result["full_config_estimates"] = est
    if args.mapping_seeds > 0:
        result["mapping"] = mapping_leg(eng, cfg, q, args.mapping_seeds, args.index_nr_seed)
    if args.audit_evals > 0:
        result["audit"] = audit_leg(eng, smp, q, args.audit_evals)
    result["autotune"] = {"median_kernel": eng.tuning(0), "counts_kernel": eng.tuning(1),
                          "counts_grid_kernel": eng.tuning(2)}
    if args.geo_fit_grains > 0:
        result["geo_fit"] = geofit_leg(eng, cfg, args.geo_fit_grains, 0 if args.no_cpu else args.geo_fit_nm_grains)
    if args.al:
        result["al_cfg5"] = al_leg(eng, args.al_steps, 1, 0 if args.no_cpu else args.al_rows,
                                   min(16, os.cpu_count() or 1), 1000 if args.geo_fit_grains > 0 else 0)
    eng.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
