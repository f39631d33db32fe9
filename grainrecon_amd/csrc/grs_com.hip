/*
 * grs_com.hip -- the objective of fit_grain_COM.m on the GPU.
 *
 * fit_grain_COM_fun (fit_grain_COM.m:32-89) for x = [position (mm), ZXZ
 * Euler angles (degrees)]: for every paired spot of the grain the event of
 * its (hkl, omega) is forward-simulated from pos = [x1, x2-RotAxisOffset,
 * x3] and U = euler2u(x(4:6)) through the forward_comp.m statements
 * (grs_event_pixel), its pixel rounded, and the distance to the measured
 * centroid (columns 19:20 of the pair row) taken; ErrMean is the mean.
 * k_com_nm runs fminsearchbnd's Nelder-Mead for each grain inside the
 * kernel (one wave per grain), so a fit costs no host round trip per
 * simplex step.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"
#include "grs_nm.h"

/* distance of one pair row r to its event simulated at x (SUB of x's U) */
__device__ static double com_pair_dist(const grs_dgeo *g, const double *r, const double *xp, const double *SUB) {
  GRS_NO_CONTRACT
  double c, s;
  grs_omega_cs(r[16], &c, &s);
  grs_vw v;
  grs_vw_terms(g, xp, c, s, &v);
  double gw[3], Gt[3], Glen;
  grs_Gw(SUB, r + 12, gw);
  grs_Gt(v.c, v.s, gw[0], gw[1], gw[2], Gt);
  const double beta = grs_event_beta(&v, Gt, &Glen);
  double ry, rz;
  grs_event_pixel(g, &v, Gt, Glen, beta, &ry, &rz);
  const double ey = grs_round(ry) - r[17], ez = grs_round(rz) - r[18];
  return grs_sqrt(ey * ey + ez * ez);
}

/* One wave per fit m: lane 0 runs the Nelder-Mead state machine (started on
 * the host) in LDS; for every requested point the lanes compute the pair
 * distances 64 at a time and every lane adds them up in pair order, so
 * ErrMean is the sequential sum over the count.  The state machine ends
 * after at most max_iter iterations, so every wave exits. */
__global__ __launch_bounds__(64) void k_com_nm(const grs_dgeo *g, const double *rows, const int64_t *rptr,
                                               grs_nm *nms) {
  GRS_NO_CONTRACT
  __shared__ __align__(16) unsigned char nm_buf[sizeof(grs_nm)];
  __shared__ double xs[GRS_NM_NMAX + 1][6];
  __shared__ double fs[GRS_NM_NMAX + 1];
  __shared__ int npend;
  grs_nm *nm = reinterpret_cast<grs_nm *>(nm_buf);
  const int lane = threadIdx.x;
  const int64_t m = blockIdx.x;
  const int64_t r0 = rptr[m], n = rptr[m + 1] - r0;
  if (lane == 0) *nm = nms[m];
  __syncthreads();
  for (;;) {
    if (lane == 0) {
      npend = nm->done() ? 0 : nm->npend;
      for (int p = 0; p < npend; p++) nm->xform(nm->pend[p], xs[p]);
    }
    __syncthreads();
    const int np = npend;
    if (np == 0) break;
    for (int p = 0; p < np; p++) {
      const double *xp = xs[p];
      double U[9], SUB[9];
      grs_U_from_euler_deg(xp[3], xp[4], xp[5], U);
      grs_SUB(g, U, SUB);
      double acc = 0.0;
      for (int64_t base = 0; base < n; base += 64) {
        const int64_t i = base + lane;
        const double d = i < n ? com_pair_dist(g, rows + 23 * (r0 + i), xp, SUB) : 0.0;
        const int cnt = n - base < 64 ? (int)(n - base) : 64;
        for (int j = 0; j < cnt; j++) acc = acc + __shfl(d, j, 64);
      }
      if (lane == 0) fs[p] = acc / (double)n;
    }
    __syncthreads();
    if (lane == 0) nm->feed(fs);
    __syncthreads();
  }
  if (lane == 0) nms[m] = *nm;
}

hipError_t grs_launch_com_nm(const grs_dgeo *d_g, const double *d_rows, const int64_t *d_rptr, int64_t M,
                             grs_nm *d_nm, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (M > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_com_nm, dim3((unsigned)M), dim3(64), 0, st, d_g, d_rows, d_rptr, d_nm);
  return hipGetLastError();
}
