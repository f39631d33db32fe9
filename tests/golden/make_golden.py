"""Writes the brute-force golden vectors of forward_comp.m for tiny inputs.

The reference holds no golden vectors, known-answer files or fixtures for the
completeness path (its arithmetic lives in MATLAB built-ins, and the MATLAB
sources cannot run here).  These vectors are enumerated by two independent
restatements of forward_comp.m:15-90 / ForwardCalc_OR_space_comp.m:24-53:
the scalar C oracle (oracle/grs_oracle.c, own FP64 transcendentals) and the
vectorised numpy one (oracle/np_restatement.py, numpy's libm).  A case is
written only if both agree on every count, every selected event (omega, hkl,
dety, detz) and every median; the script fails otherwise.

Sizes: 16 x 16 detector, 8 omega steps, 6 reflections, 3 voxel positions x
36 orientation rows per case.  Cases: untilted; tilted detector with
dety00/detz00, P0y/P0z and RotAxisOffset; beamstop.  Orientation rows include
the two rendered grains' own orientations (C = 1 at their centres), random
ones, rows with |q_vec| > 1 (zero output rows) and a voxel whose every spot
leaves the detector (Nr_simu = 0, C = NaN, dis_median = 2000).

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
import sys

import numpy as np
from scipy import ndimage

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle.np_restatement import forward_comp_np  # noqa: E402
from grainrecon_amd.geometry import cubic_hkl, labdct_geometry  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "forward_comp_tiny.json")
DET, NROT, PIX = 16, 8, 6.72e-3 * 160
GEO_FIELDS = ("S", "B", "RotDet", "Lsam2sou", "Lsam2det", "P0y", "P0z", "RotAxisOffset", "dety00", "detz00",
              "pixelysize", "pixelzsize", "dety0", "detz0", "detysize", "detzsize", "BeamStopY", "BeamStopZ",
              "thetamax", "lambda_min", "lambda_max", "minEucDis")


def _U(q):
    q1, q2, q3 = q[1], q[2], q[3]
    ss = (q1 * q1 + q2 * q2) + q3 * q3
    q0 = math.sqrt(1.0 - ss)
    return np.array([
        [2.0 * (q0 * q0 + q1 * q1) - 1.0, 2.0 * (q1 * q2 - q0 * q3), 2.0 * (q1 * q3 + q0 * q2)],
        [2.0 * (q1 * q2 + q0 * q3), 2.0 * (q0 * q0 + q2 * q2) - 1.0, 2.0 * (q2 * q3 - q0 * q1)],
        [2.0 * (q1 * q3 - q0 * q2), 2.0 * (q2 * q3 + q0 * q1), 2.0 * (q0 * q0 + q3 * q3) - 1.0],
    ])


def _quats(rng, n):
    u1, u2, u3 = rng.random((3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    q[q[:, 0] < 0] *= -1
    return q


def make_case(name, seed, tilts=(0.0, 0.0, 0.0), dety00=0.0, detz00=0.0, P0y=0.0, P0z=0.0, RAO=0.0,
              beamstop=((0.0, -1.0), (0.0, -1.0))):
    rng = np.random.default_rng(seed)
    geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=DET, pixel_mm=PIX, tilts=tilts, dety00=dety00,
                          detz00=detz00, beamstop=beamstop)
    geo = dataclasses.replace(geo, P0y=P0y, P0z=P0z, RotAxisOffset=RAO)
    hkl = cubic_hkl("I", 2)[::3][:6]
    rot = np.arange(NROT) * 45.0
    qg = _quats(rng, 2)
    centres = rng.uniform(-0.05, 0.05, (2, 3))
    mask = np.zeros((DET, DET, NROT), bool)
    empty = oracle.Stack(np.zeros((DET, DET, NROT), np.float32))
    for g in range(2):
        _, _, _, ev = oracle.forward_comp(geo, empty, _U(qg[g]), centres[g], rot, hkl, events=True)
        for om, _h, dy, dz, *_ in ev:
            mask[int(dz) - 1, int(dy) - 1, int(om)] = True
    proj = np.full((DET, DET, NROT), np.float32(DET))
    for i in range(NROT):
        if mask[:, :, i].any():
            proj[:, :, i] = ndimage.distance_transform_edt(~mask[:, :, i]).astype(np.float32)
    stack = oracle.Stack(proj)
    qs = np.vstack([qg, _quats(rng, 32)])
    bad = np.array([[0.0, 0.8, 0.5, 0.4], [0.1, -0.9, 0.3, 0.4]])   # |q_vec| > 1
    qs = np.vstack([qs, bad])
    positions = np.vstack([centres, [[0.0, 0.0, 20.0]]])            # the last one sends every spot off the detector
    rows = []
    for p in positions:
        out = oracle.or_space_comp(geo, stack, p, qs, rot, hkl, nthreads=1)
        for k, q in enumerate(qs):
            if np.dot(q[1:], q[1:]) > 1.0:
                assert not out[k].any()
                continue
            U = _U(q)
            ns, ni, med, ev = oracle.forward_comp(geo, stack, U, p, rot, hkl, events=True)
            ns2, ni2, med2, info = forward_comp_np(geo, proj, U, p, rot, hkl)
            sel = info["sel"]
            assert (ns, ni) == (ns2, ni2), (name, p, k)
            assert np.float32(med) == np.float32(med2), (name, p, k)
            assert np.array_equal(ev[:, :4], sel[:, :4]), (name, p, k)
            assert (out[k, 0], out[k, 1]) == (ns, ni) and np.float32(out[k, 3]) == np.float32(med)
        rows.append(out)
    expected = np.stack(rows)                                      # (P, N, 4)
    assert np.isnan(expected[2, :34, 2]).all() and (expected[2, :34, 3] == 2000).all()
    assert (expected[0, 0, 2] == 1.0) and (expected[1, 1, 2] == 1.0)
    truth_events = []
    for g in range(2):
        _, _, _, ev = oracle.forward_comp(geo, stack, _U(qg[g]), centres[g], rot, hkl, events=True)
        truth_events.append(ev[:, :4].astype(int).tolist())
    gd = {}
    for f in GEO_FIELDS:
        v = getattr(geo, f)
        gd[f] = np.asarray(v, float).reshape(-1).tolist() if f in ("S", "B", "RotDet", "BeamStopY", "BeamStopZ") \
            else float(v)
    return {
        "name": name, "seed": seed, "geometry": gd, "hkl": hkl.tolist(), "rot": rot.tolist(),
        "proj_shape": list(proj.shape), "proj_fortran_order": [float(x) for x in proj.reshape(-1, order="F")],
        "positions": positions.tolist(), "q": qs.tolist(),
        "expected": [[[None if math.isnan(x) else float(x) for x in r] for r in rowp] for rowp in expected],
        "truth_events": truth_events,
    }


def main():
    oracle.build()
    cases = [
        make_case("untilted", 5001),
        make_case("tilted_offsets", 5002, tilts=(0.25, -0.15, 0.3), dety00=0.2, detz00=-0.15, P0y=0.03, P0z=-0.02,
                  RAO=0.01),
        make_case("beamstop", 5003, beamstop=((6.0, 10.0), (6.0, 10.0))),
    ]
    with open(OUT, "w") as f:
        json.dump({"source": "forward_comp.m brute force, C oracle == numpy restatement", "cases": cases}, f)
    print(OUT, os.path.getsize(OUT), "bytes")
    for c in cases:
        e = np.array([[[np.nan if x is None else x for x in r] for r in rp] for rp in c["expected"]])
        print(c["name"], "Nr_simu total", int(np.nansum(e[..., 0])), "hits", int(np.nansum(e[..., 1])),
              "C=1 rows", int(np.sum(e[..., 2] == 1.0)), "events(truth)", [len(t) for t in c["truth_events"]])


if __name__ == "__main__":
    main()
