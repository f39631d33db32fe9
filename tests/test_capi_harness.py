"""The legacy cuda_forward_comp entry called from plain C (tests/capi_harness.c,
linked against libgrscore.so with gcc, only include/grscore.h), with the row
vectors of ForwardCalc_OR_space_comp_cuda.m:17-33; results equal the oracle on
the same float inputs promoted to double."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import scenes
from grainrecon_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("capi") / "capi_harness")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-O2", "-std=c99", "-I", os.path.dirname(_lib.HEADER_PATH), "-o", exe,
                    os.path.join(HERE, "capi_harness.c"), "-L", libdir, "-lgrscore", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True)
    return exe


def test_harness_builds_and_links(harness):
    out = subprocess.run(["nm", "-u", harness], capture_output=True, text=True).stdout
    assert "grs_cuda_forward_comp" in out and "grs_last_error" in out


@pytest.mark.gpu
def test_c_caller_equals_oracle(harness, tmp_path):
    from test_engine_gpu import _bits, _candidate_q, _f32, _legacy_inputs

    s = scenes.make_scene(seed=57, ngrains=2)
    g, param = _legacy_inputs(s)
    q = _candidate_q(s, np.random.default_rng(12), n_random=150)
    pos = s.centers[0]
    arrays = [np.asarray(pos), q.reshape(-1), np.asarray(s.geo.S).reshape(-1), np.asarray(s.geo.B).reshape(-1),
              s.hkl.reshape(-1), np.asarray(s.proj, np.float32).reshape(-1, order="F"), s.rot,
              np.asarray(s.geo.RotDet).reshape(-1), param]
    arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in arrays]
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([a.size for a in arrays], np.int64).tobytes())
        for a in arrays:
            f.write(a.tobytes())
    print(f"capi_harness: {q.shape[0]} orientations", flush=True)
    r = subprocess.run([harness, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(outp, np.float32).reshape(-1, 4)
    ref = oracle.or_space_comp(g, s.stack, _f32(pos), _f32(q), _f32(s.rot), _f32(s.hkl))
    assert np.array_equal(out[:, 0], ref[:, 0].astype(np.float32))
    assert np.array_equal(out[:, 1], ref[:, 1].astype(np.float32))
    assert np.array_equal(_bits(out[:, 2]), _bits(ref[:, 2].astype(np.float32)))
    assert np.array_equal(_bits(out[:, 3]), _bits(ref[:, 3].astype(np.float32)))
    assert out[0, 2] > 0.9  # grain 0's own orientation at its centre
