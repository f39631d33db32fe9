"""The C-ABI library loads and exports every entry point include/grscore.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import subprocess

import numpy as np

from grainrecon_amd import _lib
from grainrecon_amd.geometry import CGeometry, labdct_geometry


def test_library_exports_every_declared_function():
    L = _lib.lib()
    declared = _lib.header_functions()
    assert "grs_cuda_forward_comp" in declared and "grs_forward_comp" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared) == set(_lib._SIGS), "ctypes signatures out of sync with include/grscore.h"
    assert L.grs_version().decode().startswith("grscore")


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in _lib.header_functions():
        assert n in names, n


def test_geometry_struct_matches_header_layout(tmp_path):
    """Field offsets of the ctypes image equal those gcc computes from grscore.h."""
    fields = [f[0] for f in CGeometry._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu\\n", offsetof(grs_geometry, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "grscore.h"\n'
                   f'int main(void){{printf("%zu\\n", sizeof(grs_geometry));{body}return 0;}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(_lib.HEADER_PATH), "-o", str(exe), str(src)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(CGeometry)
    assert vals[1:] == [getattr(CGeometry, f).offset for f in fields]
    g = labdct_geometry(a=2.8665, lattice="I").to_c()
    assert g.detysize == 1024 and g.dety0 == 512.5 and abs(g.minEucDis - 0.5 * 6.72e-3) < 1e-18


def test_legacy_param_vector_layout():
    geo = labdct_geometry(a=2.8665, lattice="I", beamstop=((10.0, 20.0), (30.0, 40.0)))
    p = geo.legacy_param(86, 181)
    assert p.shape == (20,)
    assert list(p[:2]) == [86, 181] and list(p[10:16]) == [1024, 1024, 10, 20, 30, 40]
    assert p[16] == geo.thetamax and p[19] == geo.RotAxisOffset


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    try:
        _lib.lib()
    except RuntimeError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("expected a loud failure")
