/*
 * grs_pair.hip -- spot pairing of forward_comp_pair_spots.m on the GPU.
 *
 * forward_comp_pair_spots.m:3-93 repeats forward_comp.m for one (U, pos) and,
 * for every selected event that hits the distance stack (dis <= minEucDis /
 * mean pixel size, :70), looks for the nearest experimental spot of the same
 * projection (Spots{i}(:,9:10), Euclidean distance in pixels to the rounded
 * event pixel, first minimum, :73-74); the pair is kept when that distance is
 * below min(sqrt(Spots{i}(id,11)), 15) (:75).  The kept rows are
 *   [reshape(U,1,9) pos hkl' Energy_hkl rot_angle Spots{i}(id,9:10)
 *    Spots{i}(id,11) dety detz dis]                       (:78-79)
 * in (projection, reflection) order, where pos already has RotAxisOffset
 * subtracted from its second component (:16) and the whole row is single
 * precision because dis is (MATLAB concatenation of double and single).
 *
 * k_pair_events   one thread per (orientation, projection, reflection): the
 *                 exact event pipeline of grs_geom.h, the hit test on the
 *                 ordered keys, the nearest-spot search over the projection's
 *                 spots (CSR by projection, resident on the device).
 * k_pair_scan     one block per orientation: exclusive scan of the pair flags
 *                 (the row order) and the pair count.
 * k_pair_rows     the 23 values of every pair, rounded to single precision.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

__global__ __launch_bounds__(256) void k_pair_events(const grs_dgeo *g, const double *hkl, const double *cs,
                                                     const uint32_t *keys, const double *pos, const double *SUB,
                                                     int64_t M, const int64_t *fptr, const double *fsp, int32_t *flag,
                                                     int64_t *sid, double *lam, int32_t *dyz, float *dis) {
  GRS_NO_CONTRACT
  const int32_t H = g->H, nrot = g->nrot;
  const int64_t E = (int64_t)nrot * H;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * E) return;
  const int64_t m = t / E, rem = t - m * E;
  const int32_t om = (int32_t)(rem / H), h = (int32_t)(rem - (int64_t)om * H);
  flag[t] = 0;
  grs_vw v;
  grs_vw_terms(g, pos + 3 * m, cs[2 * om], cs[2 * om + 1], &v);
  double gw[3], Gt[3], Glen;
  grs_Gw(SUB + 9 * m, hkl + 3 * h, gw);
  grs_Gt(v.c, v.s, gw[0], gw[1], gw[2], Gt);
  const double beta = grs_event_beta(&v, Gt, &Glen);
  int32_t dy, dz;
  if (!grs_event_finish(g, &v, Gt, Glen, beta, &dy, &dz)) return;
  const uint32_t key = keys[(int64_t)(dz - 1) + (int64_t)(dy - 1) * g->nz + (int64_t)om * g->nz * g->ny];
  if (!(g->has_hit && key <= g->hit_key)) return;
  const int64_t a = fptr[om], b = fptr[om + 1];
  double best = 0.0;
  int64_t bi = -1;
  for (int64_t k = a; k < b; k++) {
    const double ey = fsp[3 * k] - (double)dy, ez = fsp[3 * k + 1] - (double)dz;
    const double d = grs_sqrt(ey * ey + ez * ez);
    if (d == d && (bi < 0 || d < best)) {
      best = d;
      bi = k;
    }
  }
  if (bi < 0) return;
  const double sa = grs_sqrt(fsp[3 * bi + 2]);
  const double lim = (sa == sa && sa < 15.0) ? sa : 15.0;
  if (!(best < lim)) return;
  /* lambdahkl with the operations of forward_comp.m:33-37 */
  const double theta = beta - g->pio2;
  const double d = ((1.0 / Glen) * 2.0) * GRS_PI;
  flag[t] = 1;
  sid[t] = bi;
  lam[t] = (2.0 * d) * grs_sin(theta);
  dyz[2 * t] = dy;
  dyz[2 * t + 1] = dz;
  dis[t] = __uint_as_float(key);
}

__global__ __launch_bounds__(1024) void k_pair_scan(const int32_t *flag, int64_t E, int32_t *off, int64_t *count) {
  __shared__ int32_t part[1024];
  const int tid = threadIdx.x, bd = blockDim.x;
  const int64_t m = blockIdx.x;
  const int32_t *f = flag + m * E;
  const int64_t per = (E + bd - 1) / bd;
  const int64_t lo = (int64_t)tid * per, hi = lo + per < E ? lo + per : E;
  int32_t c = 0;
  for (int64_t i = lo; i < hi; i++) c += f[i];
  part[tid] = c;
  __syncthreads();
  for (int o = 1; o < bd; o <<= 1) {
    const int32_t v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int32_t run = part[tid] - c;
  for (int64_t i = lo; i < hi; i++) {
    off[m * E + i] = run;
    run += f[i];
  }
  if (tid == bd - 1) count[m] = part[tid];
}

__global__ __launch_bounds__(256) void k_pair_rows(const grs_dgeo *g, const double *hkl, const double *rot,
                                                   const double *pos, const double *U, int64_t M, const double *fsp,
                                                   const int32_t *flag, const int64_t *sid, const double *lam,
                                                   const int32_t *dyz, const float *dis, const int32_t *off,
                                                   const int64_t *row_ptr, double *rows) {
  GRS_NO_CONTRACT
  const int32_t H = g->H, nrot = g->nrot;
  const int64_t E = (int64_t)nrot * H;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * E || !flag[t]) return;
  const int64_t m = t / E, rem = t - m * E;
  const int32_t om = (int32_t)(rem / H), h = (int32_t)(rem - (int64_t)om * H);
  double *r = rows + (row_ptr[m] + off[t]) * 23;
  const double *u = U + 9 * m;
  /* reshape(U,1,9): column-major */
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) r[3 * c + k] = (double)(float)u[3 * k + c];
  r[9] = (double)(float)pos[3 * m];
  r[10] = (double)(float)(pos[3 * m + 1] - g->RAO);
  r[11] = (double)(float)pos[3 * m + 2];
  for (int k = 0; k < 3; k++) r[12 + k] = (double)(float)hkl[3 * h + k];
  r[15] = (double)(float)(12.398 / lam[t]);
  r[16] = (double)(float)rot[om];
  const int64_t s = sid[t];
  r[17] = (double)(float)fsp[3 * s];
  r[18] = (double)(float)fsp[3 * s + 1];
  r[19] = (double)(float)fsp[3 * s + 2];
  r[20] = (double)(float)dyz[2 * t];
  r[21] = (double)(float)dyz[2 * t + 1];
  r[22] = (double)dis[t];
}

static unsigned pair_grid(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t grs_launch_pair_events(const grs_dgeo *d_g, const double *d_hkl, const double *d_cs,
                                  const uint32_t *d_keys, const double *d_pos, const double *d_SUB, int64_t M,
                                  int64_t E, const int64_t *d_fptr, const double *d_fsp, int32_t *d_flag,
                                  int64_t *d_sid, double *d_lam, int32_t *d_dyz, float *d_dis, hipStream_t st) {
  if (M <= 0 || E <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_events, dim3(pair_grid(M * E, 256)), dim3(256), 0, st, d_g, d_hkl, d_cs, d_keys, d_pos,
                     d_SUB, M, d_fptr, d_fsp, d_flag, d_sid, d_lam, d_dyz, d_dis);
  return hipGetLastError();
}

hipError_t grs_launch_pair_scan(const int32_t *d_flag, int64_t M, int64_t E, int32_t *d_off, int64_t *d_count,
                                hipStream_t st) {
  if (M <= 0 || E <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_scan, dim3((unsigned)M), dim3(1024), 0, st, d_flag, E, d_off, d_count);
  return hipGetLastError();
}

hipError_t grs_launch_pair_rows(const grs_dgeo *d_g, const double *d_hkl, const double *d_rot, const double *d_pos,
                                const double *d_U, int64_t M, int64_t E, const double *d_fsp, const int32_t *d_flag,
                                const int64_t *d_sid, const double *d_lam, const int32_t *d_dyz, const float *d_dis,
                                const int32_t *d_off, const int64_t *d_row_ptr, double *d_rows, hipStream_t st) {
  if (M <= 0 || E <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_rows, dim3(pair_grid(M * E, 256)), dim3(256), 0, st, d_g, d_hkl, d_rot, d_pos, d_U, M,
                     d_fsp, d_flag, d_sid, d_lam, d_dyz, d_dis, d_off, d_row_ptr, d_rows);
  return hipGetLastError();
}
