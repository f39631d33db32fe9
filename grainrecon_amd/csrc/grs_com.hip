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
 * kernel (one block per grain), so a fit costs no host round trip per
 * simplex step.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"
#include "grs_nm.h"

/* cos/sin of every pair row's omega (column 17), once per call */
__global__ __launch_bounds__(256) void k_com_prep(const double *rows, int64_t total, double *cs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) grs_omega_cs(rows[23 * i + 16], &cs[2 * i], &cs[2 * i + 1]);
}

/* distance of one pair row r to its event simulated at x (SUB of x's U,
 * cs the cos/sin of the row's omega) */
__device__ static double com_pair_dist(const grs_dgeo *g, const double *r, const double *xp, const double *SUB,
                                       const double *cs) {
  GRS_NO_CONTRACT
  grs_vw v;
  grs_vw_terms(g, xp, cs[0], cs[1], &v);
  double gw[3], Gt[3], Glen;
  grs_Gw(SUB, r + 12, gw);
  grs_Gt(v.c, v.s, gw[0], gw[1], gw[2], Gt);
  const double beta = grs_event_beta(&v, Gt, &Glen);
  double ry, rz;
  grs_event_pixel(g, &v, Gt, Glen, beta, &ry, &rz);
  const double ey = grs_round(ry) - r[17], ez = grs_round(rz) - r[18];
  return grs_sqrt(ey * ey + ez * ez);
}

#define COM_THREADS 512

/* One block per fit m: thread 0 steps the Nelder-Mead state machine
 * (started on the host) in LDS; for the requested points the first threads
 * build each point's U and SUB, all threads compute the (point, pair)
 * distances into dis (9 slots of the grain's pair count per grain), and
 * thread p adds point p's distances in pair order, so ErrMean is the
 * sequential sum over the count.  The point coordinates and the six
 * cosines and sines of each point's angles are computed one per thread.
 * The state machine ends after at most max_iter iterations, so every block
 * exits. */
__global__ __launch_bounds__(COM_THREADS) void k_com_nm(const grs_dgeo *g, const double *rows, const int64_t *rptr,
                                                        const double *cs, double *dis, grs_nm *nms) {
  GRS_NO_CONTRACT
  __shared__ __align__(16) unsigned char nm_buf[sizeof(grs_nm)];
  __shared__ double xs[GRS_NM_NMAX + 1][6];
  __shared__ double sub[GRS_NM_NMAX + 1][9];
  __shared__ double trig[GRS_NM_NMAX + 1][6];
  __shared__ double fs[GRS_NM_NMAX + 1];
  __shared__ int npend;
  grs_nm *nm = reinterpret_cast<grs_nm *>(nm_buf);
  const int t = threadIdx.x;
  const int64_t m = blockIdx.x;
  const int64_t r0 = rptr[m];
  const int n = (int)(rptr[m + 1] - r0);
  double *dm = dis + (GRS_NM_NMAX + 1) * r0;
  if (t == 0) *nm = nms[m];
  __syncthreads();
  for (;;) {
    if (t == 0) npend = nm->done() ? 0 : nm->npend;
    __syncthreads();
    const int np = npend;
    if (np == 0) break;
    if (t < 6 * np) nm->xform1(nm->pend[t / 6], t % 6, &xs[t / 6][t % 6]); /* one coordinate per thread */
    __syncthreads();
    if (t < 6 * np) { /* cos, sin of phi1, PHI, phi2: one value per thread */
      const int p = t / 6, j = t % 6;
      const double a = grs_deg2rad(xs[p][3 + j / 2]);
      trig[p][j] = (j & 1) ? grs_sin(a) : grs_cos(a);
    }
    __syncthreads();
    if (t < np) {
      double U[9];
      grs_U_from_trig(trig[t][0], trig[t][1], trig[t][2], trig[t][3], trig[t][4], trig[t][5], U);
      grs_SUB(g, U, sub[t]);
    }
    __syncthreads();
    for (int w = t; w < np * n; w += COM_THREADS) {
      const int p = w / n, i = w - p * n;
      dm[w] = com_pair_dist(g, rows + 23 * (r0 + i), xs[p], sub[p], cs + 2 * (r0 + i));
    }
    __syncthreads();
    if (t < np) {
      /* loads 8 ahead of the in-order additions */
      const double *d = dm + t * n;
      double acc = 0.0;
      int i = 0;
      for (; i + 8 <= n; i += 8) {
        double b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) b[k] = d[i + k];
#pragma unroll
        for (int k = 0; k < 8; k++) acc = acc + b[k];
      }
      for (; i < n; i++) acc = acc + d[i];
      fs[t] = acc / (double)n;
    }
    __syncthreads();
    if (t == 0) nm->feed(fs);
    __syncthreads();
  }
  if (t == 0) nms[m] = *nm;
}

hipError_t grs_launch_com_nm(const grs_dgeo *d_g, const double *d_rows, const int64_t *d_rptr, int64_t M,
                             int64_t total, double *d_cs, double *d_dis, grs_nm *d_nm, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  if (M > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_com_prep, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, d_rows, total, d_cs);
  hipLaunchKernelGGL(k_com_nm, dim3((unsigned)M), dim3(COM_THREADS), 0, st, d_g, d_rows, d_rptr, d_cs, d_dis, d_nm);
  return hipGetLastError();
}
