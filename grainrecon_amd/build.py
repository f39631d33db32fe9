"""Build the gfx950 engine library (libgrscore.so) in-tree with hipcc.

Flags that matter for bit-exactness:
  -ffp-contract=off   no fused multiply-add except the explicit fma() calls
                      (HIP device code defaults to contraction on)
  no fast-math        IEEE division, sqrt and NaN semantics
Packed FP32 generation is disabled (-packed-fp32-ops); the engine computes in
FP64 anyway.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgrscore.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("GRS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["grs_kernels.hip", "grs_prep.hip", "grs_fit.hip", "grs_cand.hip", "grs_index.hip", "grs_pair.hip", "grs_region.hip", "grs_gt.hip", "grs_com.hip", "grs_capi.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-Wl,-rpath,{ROCM}/lib",
]


def hipcc():
    exe = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build_engine(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgrscore.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_engine(verbose=True))
