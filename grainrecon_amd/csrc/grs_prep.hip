/*
 * grs_prep.hip -- device-side input preparation for the scoring engine.
 *
 * k_render_spots   forward simulation of binary spot masks with the
 *                  forward_comp.m geometry: every selected (omega, hkl) event
 *                  of every (voxel, grain orientation) sets its pixel.  Used to
 *                  build synthetic data sets whose spots are the projections of
 *                  the grains.
 * k_edt_cols/rows  exact Euclidean distance transform of each binary
 *                  projection, the bwdist of GrainMapping_fun.m:73-75: the
 *                  distance from every pixel to the nearest spot pixel, stored
 *                  as single(sqrt(d^2)).  Separable algorithm of P. Felzenszwalb
 *                  and D. Huttenlocher, "Distance Transforms of Sampled
 *                  Functions", Theory of Computing 8 (2012): a 1-D distance
 *                  along detz, then the lower envelope of parabolas along dety.
 *                  Envelope intersections are rationals with denominators
 *                  below 2*ny, evaluated in double, so every decision and every
 *                  d^2 is exact.  A projection without any spot pixel gives
 *                  +Inf everywhere, as bwdist does.
 * Layout of masks and stacks: MATLAB column-major (detz fastest, then dety,
 * then omega).
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

#define GRS_EDT_INF 0x3FFFFFFF

__global__ void k_render_spots(const grs_dgeo *g, const double *pos, const int32_t *grain, int64_t M,
                               const double *SUB, const double *hkl, const double *cs, uint8_t *mask) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t nrot = g->nrot, H = g->H;
  if (i >= M * nrot) return;
  int64_t v = i / nrot;
  int32_t om = (int32_t)(i - v * nrot);
  const double *sub = SUB + 9 * (int64_t)grain[v];
  grs_vw t;
  grs_vw_terms(g, pos + 3 * v, cs[2 * om], cs[2 * om + 1], &t);
  const int64_t nz = g->nz, nzny = g->nz * g->ny;
  for (int32_t h = 0; h < H; h++) {
    double gw[3], Gt[3], Glen;
    grs_Gw(sub, hkl + 3 * h, gw);
    grs_Gt(t.c, t.s, gw[0], gw[1], gw[2], Gt);
    double beta = grs_event_beta(&t, Gt, &Glen);
    int32_t dy, dz;
    if (grs_event_finish(g, &t, Gt, Glen, beta, &dy, &dz))
      mask[(int64_t)(dz - 1) + (int64_t)(dy - 1) * nz + (int64_t)om * nzny] = 1;
  }
}

/* Every selected event of voxel v (orientation SUB[grain[v]]): the omega and
 * hkl indices and the unrounded pixel coordinates -dety22/py + dety0,
 * -detz22/pz + detz0.  Slots come from an atomic counter; the host sorts. */
__global__ void k_sim_events(const grs_dgeo *g, const double *pos, const int32_t *grain, int64_t M,
                             const double *SUB, const double *hkl, const double *cs, int64_t cap,
                             unsigned long long *counter, int32_t *ev_vox, int32_t *ev_om, int32_t *ev_h,
                             double *ev_yz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t nrot = g->nrot, H = g->H;
  if (i >= M * nrot) return;
  int64_t v = i / nrot;
  int32_t om = (int32_t)(i - v * nrot);
  const double *sub = SUB + 9 * (int64_t)grain[v];
  grs_vw t;
  grs_vw_terms(g, pos + 3 * v, cs[2 * om], cs[2 * om + 1], &t);
  for (int32_t h = 0; h < H; h++) {
    double gw[3], Gt[3], Glen;
    grs_Gw(sub, hkl + 3 * h, gw);
    grs_Gt(t.c, t.s, gw[0], gw[1], gw[2], Gt);
    double beta = grs_event_beta(&t, Gt, &Glen);
    int32_t dy, dz;
    double yz[2];
    if (grs_event_finish_yz(g, &t, Gt, Glen, beta, &dy, &dz, yz)) {
      unsigned long long slot = atomicAdd(counter, 1ull);
      if (slot < (unsigned long long)cap) {
        ev_vox[slot] = (int32_t)v;
        ev_om[slot] = om;
        ev_h[slot] = h;
        ev_yz[2 * slot] = yz[0];
        ev_yz[2 * slot + 1] = yz[1];
      }
    }
  }
}

/* 1-D distance along detz within each (dety, omega) column. */
__global__ void k_edt_cols(const uint8_t *mask, int64_t nz, int64_t ncols, int32_t *g1) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const uint8_t *m = mask + c * nz;
  int32_t *o = g1 + c * nz;
  int32_t d = GRS_EDT_INF;
  for (int64_t z = 0; z < nz; z++) {
    if (m[z]) d = 0;
    else if (d < GRS_EDT_INF) d++;
    o[z] = d;
  }
  d = GRS_EDT_INF;
  for (int64_t z = nz - 1; z >= 0; z--) {
    if (m[z]) d = 0;
    else if (d < GRS_EDT_INF) d++;
    if (d < o[z]) o[z] = d;
  }
}

/* Lower envelope along dety for each (detz row, omega) of a chunk of slices;
 * v/zb are per-thread work arrays, element k at [k*T + tid]. */
__global__ void k_edt_rows(const int32_t *g1, int64_t nz, int64_t ny, int64_t slice0, int64_t nslices,
                           int32_t *v, double *zb, float *out) {
  int64_t T = nz * nslices;
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= T) return;
  int64_t iz = tid % nz;
  int64_t sl = slice0 + tid / nz;
  const int32_t *col = g1 + sl * nz * ny + iz; /* f(q) = col[q*nz]^2 */
  float *o = out + sl * nz * ny + iz;
  int64_t k = -1;
  for (int64_t q = 0; q < ny; q++) {
    int32_t gq = col[q * nz];
    if (gq >= GRS_EDT_INF) continue;
    double fq = (double)gq * (double)gq;
    double s = 0.0;
    while (k >= 0) {
      int32_t p = v[k * T + tid];
      int32_t gp = col[(int64_t)p * nz];
      double fp = (double)gp * (double)gp;
      s = ((fq + (double)q * (double)q) - (fp + (double)p * (double)p)) / (2.0 * (double)q - 2.0 * (double)p);
      if (s <= zb[k * T + tid]) k--;
      else break;
    }
    k++;
    v[k * T + tid] = (int32_t)q;
    zb[k * T + tid] = k == 0 ? -1.0e300 : s;
  }
  if (k < 0) {
    for (int64_t q = 0; q < ny; q++) o[q * nz] = __int_as_float(0x7F800000);
    return;
  }
  int64_t j = 0;
  for (int64_t q = 0; q < ny; q++) {
    while (j < k && zb[(j + 1) * T + tid] < (double)q) j++;
    int32_t p = v[j * T + tid];
    int32_t gp = col[(int64_t)p * nz];
    double d2 = ((double)q - (double)p) * ((double)q - (double)p) + (double)gp * (double)gp;
    o[q * nz] = (float)__builtin_sqrt(d2);
  }
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t grs_launch_render(const grs_dgeo *d_g, const double *d_pos, const int32_t *d_grain, int64_t M,
                             int32_t nrot, const double *d_SUB, const double *d_hkl, const double *d_cs,
                             uint8_t *d_mask, hipStream_t st) {
  int64_t n = M * nrot;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_render_spots, dim3(nblk(n, 128)), dim3(128), 0, st, d_g, d_pos, d_grain, M, d_SUB, d_hkl,
                     d_cs, d_mask);
  return hipGetLastError();
}

hipError_t grs_launch_sim_events(const grs_dgeo *d_g, const double *d_pos, const int32_t *d_grain, int64_t M,
                                 int32_t nrot, const double *d_SUB, const double *d_hkl, const double *d_cs, int64_t cap,
                                 unsigned long long *d_counter, int32_t *d_vox, int32_t *d_om, int32_t *d_h,
                                 double *d_yz, hipStream_t st) {
  int64_t n = M * nrot;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sim_events, dim3(nblk(n, 128)), dim3(128), 0, st, d_g, d_pos, d_grain, M, d_SUB, d_hkl, d_cs,
                     cap, d_counter, d_vox, d_om, d_h, d_yz);
  return hipGetLastError();
}

hipError_t grs_launch_edt_cols(const uint8_t *d_mask, int64_t nz, int64_t ncols, int32_t *d_g1, hipStream_t st) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_edt_cols, dim3(nblk(ncols, 256)), dim3(256), 0, st, d_mask, nz, ncols, d_g1);
  return hipGetLastError();
}

hipError_t grs_launch_edt_rows(const int32_t *d_g1, int64_t nz, int64_t ny, int64_t slice0, int64_t nslices,
                               int32_t *d_v, double *d_z, float *d_out, hipStream_t st) {
  int64_t T = nz * nslices;
  if (T <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_edt_rows, dim3(nblk(T, 256)), dim3(256), 0, st, d_g1, nz, ny, slice0, nslices, d_v, d_z,
                     d_out);
  return hipGetLastError();
}
