"""grs_forward_comp_device_noalloc (the allocation-free one-voxel entry)
against grs_forward_comp_device, through the C ABI, bit for bit.

Two calls in a row share the library's static scratch: the first scores
40,000 orientations (three chunks of the scratch's S*U*B region), the second
a smaller set at another voxel, so stale contents of the first call must not
leak into the second.  The scene selects ~4,000 events per evaluation (512^2
detector of 27 um pixels, 361 omega, 140 reflections), so every wave keeps
its median keys in its scratch key list.  A geometry whose events per
evaluation leave room for fewer than 4 key lists is rejected before any
launch.
"""
import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd import _lib
from grainrecon_amd.geometry import cubic_hkl, labdct_geometry

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _run(engine, fn, pos, d_q, N, n_out):
    d_ns, d_ni, d_C, d_med = (engine.device_alloc(b * n_out) for b in (4, 4, 8, 4))
    try:
        fn(pos, d_q, N, d_ns, d_ni, d_C, d_med)
        engine.synchronize()
        return (engine.to_host(d_ns, (N,), np.int32), engine.to_host(d_ni, (N,), np.int32),
                engine.to_host(d_C, (N,), np.float64), engine.to_host(d_med, (N,), np.float32))
    finally:
        for d in (d_ns, d_ni, d_C, d_med):
            engine.device_free(d)


def test_noalloc_equals_device_entry_on_two_calls(engine):
    rng = np.random.default_rng(9191)
    det = 512
    geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=det, pixel_mm=6.72e-3 * 4)
    hkl = cubic_hkl("I", 8)
    rot = np.arange(0.0, 360.5, 1.0)
    flat = rng.integers(0, 400, det * det * rot.size, dtype=np.uint16).astype(np.float32)
    np.subtract(flat, 100.0, out=flat)
    np.maximum(flat, 0.0, out=flat)
    np.sqrt(flat, out=flat)
    proj = flat.reshape((det, det, rot.size), order="F")
    engine.set_geometry(geo, hkl, rot)
    engine.set_distance_stack(proj)
    print(f"[noalloc] {hkl.shape[0]} hkl x {rot.size} omega on {det}^2", flush=True)
    q = scenes.random_quats(rng, 40_000)
    q[7, 1:] = [0.8, 0.6, 0.3]  # |q_vec| > 1: a skipped row
    d_q = engine.to_device(q)
    try:
        for call, (pos, N) in enumerate((([0.01, -0.02, 0.005], 40_000), ([-0.03, 0.01, 0.02], 2_500))):
            print(f"[noalloc] call {call}: {N} orientations", flush=True)
            ref = _run(engine, lambda *a: engine.forward_comp_device(*a, _lib.METHOD_FAST), pos, d_q, N, N)
            got = _run(engine, engine.forward_comp_device_noalloc, pos, d_q, N, N)
            for r, g in zip(ref, got):
                assert np.array_equal(_bits(r), _bits(g)), call
            assert ref[0].max() > 1024  # the median reads the scratch key lists
            assert ref[0][7] == 0
        # a few rows of the last call against the oracle as well
        ora = oracle.or_space_comp(geo, oracle.Stack(proj), pos, q[:8], rot, hkl, nthreads=8)
        assert np.array_equal(got[0][:8], ora[:, 0].astype(np.int32))
        assert np.array_equal(got[1][:8], ora[:, 1].astype(np.int32))
        assert np.array_equal(_bits(got[2][:8]), _bits(ora[:, 2]))
        assert np.array_equal(_bits(got[3][:8]), _bits(ora[:, 3].astype(np.float32)))
    finally:
        engine.device_free(d_q)


def test_noalloc_rejects_geometry_beyond_scratch(engine):
    """240 reflections x 18,000 omega = 4.32e6 events per evaluation: fewer
    than 4 key lists fit in the scratch, so the call fails before launching."""
    geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=1, pixel_mm=6.72e-3)
    hkl = np.tile([[1.0, 1.0, 0.0]], (240, 1))
    rot = np.linspace(0.0, 360.0, 18_000)
    engine.set_geometry(geo, hkl, rot)
    engine.set_distance_stack(np.zeros((1, 1, rot.size), np.float32))
    d_q = engine.to_device(np.array([[1.0, 0.0, 0.0, 0.0]]))
    bufs = [engine.device_alloc(8) for _ in range(4)]
    try:
        with pytest.raises(_lib.GrsError) as e:
            engine.forward_comp_device_noalloc([0.0, 0.0, 0.0], d_q, 1, *bufs)
        assert e.value.code == _lib.GRS_ERR_ARG
        assert "allocation-free scratch" in str(e.value)
    finally:
        for d in bufs + [d_q]:
            engine.device_free(d)
