"""GPU parity at the detector sizes and pixel pitches of the benchmark
workloads, against the C oracle.

The scene tests render small detectors with coarse pixels; here the detector
is 2048^2 at 3.36 um (BCC Fe, 181 omega of 2 deg) or 1024^2 at 6.72 um (FCC
Al, 361 omega of 1 deg), where a pixel is smallest against the sample and the
candidate bounds of the fast kernel (omega windows, detector cone and rows)
have the least room.  The distance stack is random (bwdist-like values
sqrt(k), about a quarter zeros), the orientations are random, and the voxel
positions span the sample cylinder (radius and half height 0.2 mm) including
its rim; one case adds beam offsets, detector tilts and offsets.  Counts,
completeness and the single-precision median must be bitwise equal.
"""
import gc

import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd import _lib
from grainrecon_amd.geometry import cubic_hkl, labdct_geometry

pytestmark = pytest.mark.gpu

CASES = {
    "fe_2048": dict(a=2.8665, lattice="I", det=2048, px=3.36e-3, step=2.0),
    "fe_2048_offsets_tilts": dict(a=2.8665, lattice="I", det=2048, px=3.36e-3, step=2.0, tilts=(0.2, -0.15, 0.1),
                                  dety00=0.05, detz00=-0.04, P0y=0.1, P0z=-0.08, RotAxisOffset=0.02),
    "al_1024_361": dict(a=4.0495, lattice="F", det=1024, px=6.72e-3, step=1.0),
}


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _stack(rng, det, nrot):
    """A (det, det, nrot) float32 stack in MATLAB (column-major) order, built
    in place so that neither the engine nor the oracle copies it."""
    n = det * det * nrot
    flat = rng.integers(0, 400, n, dtype=np.uint16).astype(np.float32)
    np.subtract(flat, 100.0, out=flat)
    np.maximum(flat, 0.0, out=flat)
    np.sqrt(flat, out=flat)
    return flat.reshape((det, det, nrot), order="F")


def _positions(rng):
    r = 0.2 * np.sqrt(rng.random(3))
    t = rng.uniform(0, 2 * np.pi, 3)
    pts = np.stack([r * np.cos(t), r * np.sin(t), rng.uniform(-0.2, 0.2, 3)], 1)
    return np.vstack([pts, [[0.2, 0.0, 0.2]], [[-0.1414, 0.1414, -0.2]]])


@pytest.mark.parametrize("case", list(CASES))
def test_fullsize_parity(engine, case):
    c = dict(CASES[case])
    rng = np.random.default_rng(7100 + list(CASES).index(case))
    geo = labdct_geometry(a=c["a"], lattice=c["lattice"], det_pixels=c["det"], pixel_mm=c["px"],
                          tilts=c.get("tilts", (0.0, 0.0, 0.0)), dety00=c.get("dety00", 0.0),
                          detz00=c.get("detz00", 0.0))
    for k in ("P0y", "P0z", "RotAxisOffset"):
        if k in c:
            setattr(geo, k, c[k])
    rot = np.arange(0.0, 360.0 + c["step"] / 2, c["step"])
    hkl = cubic_hkl(c["lattice"], 6)
    print(f"[fullsize] {case}: detector {c['det']}^2, {rot.size} omega, {hkl.shape[0]} hkl", flush=True)
    proj = _stack(rng, c["det"], rot.size)
    engine.set_geometry(geo, hkl, rot)
    engine.set_distance_stack(proj)
    stack = oracle.Stack(proj)
    q = scenes.random_quats(rng, 48)
    for pos in _positions(rng):
        ref = oracle.or_space_comp(geo, stack, pos, q, rot, hkl)
        assert ref[:, 0].min() > 0
        for method in (_lib.METHOD_FAST, _lib.METHOD_PLAIN):
            ns, ni, C, med = engine.forward_comp(pos, q, method)
            assert np.array_equal(ns, ref[:, 0].astype(np.int32)), (case, pos, method)
            assert np.array_equal(ni, ref[:, 1].astype(np.int32)), (case, pos, method)
            assert np.array_equal(_bits(C), _bits(ref[:, 2])), (case, pos, method)
            assert np.array_equal(_bits(med), _bits(ref[:, 3].astype(np.float32))), (case, pos, method)
    del stack, proj
    gc.collect()
