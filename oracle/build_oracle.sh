#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# The reference GrainRecon has no compilable source for this path (its CUDA
# MEX sources are absent and its C MEX ships only as a binary), so there is
# no oracle/_ref build; the oracle is this project's C restatement.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math -std=c11 \
    -o "$HERE/liboracle.so.tmp" "$HERE/grs_oracle.c" "$HERE/grs_candidates.c" "$HERE/grs_geofit.c" -lm
mv "$HERE/liboracle.so.tmp" "$HERE/liboracle.so"
echo "$HERE/liboracle.so"
