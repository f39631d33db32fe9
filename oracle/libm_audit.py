"""Margin audit of the oracle's continuous decisions -- TEST INFRASTRUCTURE.

forward_comp.m decides per event with four window tests (beta > pi/2,
beta < beta_max, lambda_min < lambda < lambda_max, :61-62) and two roundings
(:58-59, which fix the pixel, the detector bounds and the beamstop test).
The MATLAB built-ins behind them (acos, sin, cos, atan) cannot run here, and
the oracle and engine use the project's own FP64 functions (< 1 ulp).  This
audit recomputes every event with the C library's functions as well and
reports, per workload:
  * the events whose selection or pixel differ and the orientations whose
    counts or median differ between the two function sets;
  * the largest difference of beta, lambda and the raw pixel coordinates
    (how far two < 1-ulp implementations move the decided quantities);
  * the smallest margin of each decision over the events it decides.
A margin many times the difference means any other < 1-ulp library (MATLAB's)
takes the same decisions on that workload.

Usage: python -m oracle.libm_audit [--orientations N] [--threads T]
prints a markdown table for the seeding sample of config 2 (Fe, 2048^2 at
3.36 um, 181 omega; the bench's seed voxel and first N orientations of its
262,144-orientation set) and for config 5 (Al, 361 omega; 8 voxels of the
sample x random orientations).  Decisions do not depend on the distance
stack, so a zero stack stands in for it (counts then differ only if a
selection or pixel does).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import oracle  # noqa: E402


def zero_stack(nz: int, ny: int, nrot: int) -> oracle.Stack:
    """A zero distance stack in MATLAB order (np.zeros maps untouched pages
    lazily, so only the pages the events read are ever materialised)."""
    return oracle.Stack(np.zeros((nz, ny, nrot), np.float32, order="F"))


def audit(geo, rot, hkl, pos_list, q_rows, threads):
    st = zero_stack(int(geo.detzsize), int(geo.detysize), rot.size)
    tot = None
    for pos, q in zip(pos_list, q_rows):
        r = oracle.margin_audit(geo, st, pos, q, rot, hkl, nthreads=threads)
        if tot is None:
            tot = dict(r)
            continue
        for k, v in r.items():
            if k.startswith("max_"):
                tot[k] = max(tot[k], v)
            elif k.startswith("min_"):
                tot[k] = min(tot[k], v)
            else:
                tot[k] += v
    return tot


def cfg2_sample(n):
    """The bench's seeding sample: seed voxel seeds[warmup] (3 warmup steps)
    and the first n orientations of its set."""
    from grainrecon_amd import synth

    cfg = synth.fe_config("cfg2")
    q = synth.fz_reduce_cubic(synth.haar_quaternions(np.random.default_rng(2002), 262144))[:n]
    srng = np.random.default_rng(3002)
    seeds = cfg.sample.pos[srng.choice(cfg.sample.pos.shape[0], size=3 + 20 + 1, replace=False)]
    return cfg, [seeds[3]], [q]


def cfg5_sample(n_vox, n_or):
    from grainrecon_amd import synth

    cfg = synth.al_config()
    rng = np.random.default_rng(3005)
    vox = rng.choice(cfg.sample.pos.shape[0], size=n_vox, replace=False)
    qs = synth.fz_reduce_cubic(synth.haar_quaternions(np.random.default_rng(2005), n_vox * n_or))
    return cfg, [cfg.sample.pos[v] for v in vox], [qs[k * n_or:(k + 1) * n_or] for k in range(n_vox)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--orientations", type=int, default=65536)
    ap.add_argument("--al-voxels", type=int, default=8)
    ap.add_argument("--al-orientations", type=int, default=2048)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    rows = []
    for name, (cfg, pos, q) in (
            (f"cfg2 seeding: 1 voxel x {args.orientations} orientations", cfg2_sample(args.orientations)),
            (f"cfg5 Al: {args.al_voxels} voxels x {args.al_orientations} orientations",
             cfg5_sample(args.al_voxels, args.al_orientations))):
        t = time.perf_counter()
        r = audit(cfg.geo, cfg.rot, cfg.hkl, pos, q, args.threads)
        r["seconds"] = time.perf_counter() - t
        rows.append((name, r))
        print(f"[audit] {name}: {r['seconds']:.1f}s", file=sys.stderr, flush=True)
    keys = list(rows[0][1])
    print("| quantity | " + " | ".join(n for n, _ in rows) + " |")
    print("|---|" + "---|" * len(rows))
    for k in keys:
        print(f"| {k} | " + " | ".join(f"{r[k]:.6g}" for _, r in rows) + " |")
    print("\nmin_ratio_*: the smallest, over the events a decision decides, of its margin divided by "
          "max(|project - C library|, 1 ulp) of the decided quantity at that event.")


if __name__ == "__main__":
    main()
