/*
 * Plain C caller of the legacy cuda_forward_comp entry, the way a MEX gateway
 * calls it (ForwardCalc_OR_space_comp_cuda.m:17-33): every input is a single
 * row vector, matrices as reshape(X',1,9), q as reshape(OR.q',1,[]), the
 * stack as reshape(proj_bin_bw,1,[]), param the 20 values of :25-27; the
 * output is 4 floats per orientation.  Built and run by
 * tests/test_capi_harness.py; no HIP or torch headers, only grscore.h.
 *
 * usage: capi_harness IN OUT
 *   IN : 9 int64 element counts (pos q S B hkl proj rot RotDet param), then
 *        the 9 float32 arrays in that order
 *   OUT: the 4*n_q/4 float32 results; exit status 0, or 2 with the engine's
 *        error text on stderr
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "grscore.h"

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 1;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 1;
  int64_t n[9];
  if (fread(n, sizeof(int64_t), 9, f) != 9) return 1;
  float *a[9];
  for (int k = 0; k < 9; k++) {
    if (n[k] < 0) return 1;
    a[k] = (float *)malloc(sizeof(float) * (size_t)(n[k] > 0 ? n[k] : 1));
    if (!a[k] || fread(a[k], sizeof(float), (size_t)n[k], f) != (size_t)n[k]) return 1;
  }
  fclose(f);
  const int64_t n_out = n[1]; /* 4 values per orientation, as many as q holds */
  float *out = (float *)malloc(sizeof(float) * (size_t)(n_out > 0 ? n_out : 1));
  if (!out) return 1;
  /* pos, q, S, B, hkl, proj, rot, RotDet, param */
  grs_status s = grs_cuda_forward_comp(a[0], a[1], n[1], a[2], a[3], a[4], n[4], a[5], n[5], a[6], n[6], a[7], a[8],
                                       n[8], out);
  if (s != GRS_OK) {
    fprintf(stderr, "grs_cuda_forward_comp: %d %s\n", (int)s, grs_last_error(NULL));
    return 2;
  }
  f = fopen(argv[2], "wb");
  if (!f || fwrite(out, sizeof(float), (size_t)n_out, f) != (size_t)n_out) return 1;
  fclose(f);
  for (int k = 0; k < 9; k++) free(a[k]);
  free(out);
  return 0;
}
