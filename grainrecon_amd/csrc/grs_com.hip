/*
 * grs_com.hip -- the objective of fit_grain_COM.m on the GPU.
 *
 * fit_grain_COM_fun (fit_grain_COM.m:32-89) for x = [position (mm), ZXZ
 * Euler angles (degrees)]: for every paired spot of the grain the event of
 * its (hkl, omega) is forward-simulated from pos = [x1, x2-RotAxisOffset,
 * x3] and U = euler2u(x(4:6)) through the forward_comp.m statements
 * (grs_event_pixel), its pixel rounded, and the distance to the measured
 * centroid (columns 19:20 of the pair row) taken; ErrMean is the mean.
 * k_com_dist computes the distances of every (point, pair): one block row
 * per point, threads over the grain's pairs; the host sums each point's
 * distances in pair order.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

/* rows: 23 values per pair (grs_pair_spots layout); x: 6 per point;
 * pfit[p] = the fit of point p; rptr: pair rows of each fit; doff[p] = first
 * output slot of point p */
__global__ __launch_bounds__(256) void k_com_dist(const grs_dgeo *g, const double *rows, const int64_t *rptr,
                                                  const double *x, const int32_t *pfit, const int64_t *doff,
                                                  double *dis) {
  GRS_NO_CONTRACT
  const int64_t p = blockIdx.y;
  const int32_t m = pfit[p];
  const int64_t r0 = rptr[m], n = rptr[m + 1] - r0;
  const double *xp = x + 6 * p;
  double U[9], SUB[9];
  grs_U_from_euler_deg(xp[3], xp[4], xp[5], U);
  grs_SUB(g, U, SUB);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double *r = rows + 23 * (r0 + i);
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
    dis[doff[p] + i] = grs_sqrt(ey * ey + ez * ez);
  }
}

hipError_t grs_launch_com_dist(const grs_dgeo *d_g, const double *d_rows, const int64_t *d_rptr, const double *d_x,
                               const int32_t *d_pfit, const int64_t *d_doff, int64_t npoints, int64_t maxpairs,
                               double *d_dis, hipStream_t st) {
  if (npoints <= 0 || maxpairs <= 0) return hipSuccess;
  if (npoints > 65535) return hipErrorInvalidValue;
  const unsigned bx = (unsigned)((maxpairs + 255) / 256 < 64 ? (maxpairs + 255) / 256 : 64);
  hipLaunchKernelGGL(k_com_dist, dim3(bx, (unsigned)npoints), dim3(256), 0, st, d_g, d_rows, d_rptr, d_x, d_pfit,
                     d_doff, d_dis);
  return hipGetLastError();
}
