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
 * kernel (one block per grain, its first wave stepping the method with one
 * lane per coordinate), so a fit costs no host round trip per simplex step.
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

/* grs_nm::feed stepped by the 64 lanes of one wave.  Every lane takes the
 * method's decisions on the shared values (fv, fxr, counts, state; all lanes
 * write the same values) and lane k alone updates coordinate k of the
 * simplex, the requested points, xbar and xr, with grs_nm's statements for
 * that coordinate; the stopping test's largest vertex distance and largest
 * coordinate are reduced across the lanes (a maximum is exact).  The state
 * afterwards is grs_nm::feed's, bit for bit: the sequential form on one lane
 * spent most of a fit's time on these loops. */
struct nm_lanes {
  grs_nm *s;
  int k;
  bool on;

  __device__ static double wave_max(double x) {
    for (int off = 32; off > 0; off >>= 1) x = grs_nm_max(x, __shfl_xor(x, off, 64));
    return x;
  }

  __device__ void sort() {
    const int n = s->n;
    for (int i = 1; i <= n; i++) {
      const double f = s->fv[i];
      const double row = on ? s->v[i][k] : 0.0;
      int j = i - 1;
      while (j >= 0 && (f == f) && (s->fv[j] != s->fv[j] || s->fv[j] > f)) {
        s->fv[j + 1] = s->fv[j];
        if (on) s->v[j + 1][k] = s->v[j][k];
        j--;
      }
      s->fv[j + 1] = f;
      if (on) s->v[j + 1][k] = row;
    }
  }

  __device__ void begin() {
    GRS_NO_CONTRACT
    const int n = s->n;
    if (!(s->nf < s->max_fe && s->it < s->max_it)) {
      s->st = grs_nm::DONE;
      s->npend = 0;
      return;
    }
    double dfv = __builtin_nan(""), dx = __builtin_nan(""), vmax = __builtin_nan("");
    for (int j = 1; j <= n; j++) dfv = grs_nm_max(dfv, fabs(s->fv[0] - s->fv[j]));
    if (on) {
      const double v0 = s->v[0][k];
      for (int j = 1; j <= n; j++) dx = grs_nm_max(dx, fabs(s->v[j][k] - v0));
      vmax = v0;
    }
    dx = wave_max(dx);
    vmax = wave_max(vmax);
    if (dfv <= grs_nm_max(s->tol_f, 10.0 * grs_nm_eps(s->fv[0])) &&
        dx <= grs_nm_max(s->tol_x, 10.0 * grs_nm_eps(vmax))) {
      s->st = grs_nm::DONE;
      s->npend = 0;
      return;
    }
    if (on) {
      double acc = 0.0;
      for (int j = 0; j < n; j++) acc += s->v[j][k];
      const double xb = acc / (double)n;
      s->xbar[k] = xb;
      const double xr = (1.0 + 1.0) * xb - 1.0 * s->v[n][k];
      s->xr[k] = xr;
      s->pend[0][k] = xr;
    }
    s->npend = 1;
    s->st = grs_nm::REFLECT;
  }

  __device__ void end_iter() {
    sort();
    s->it = s->it + 1;
    begin();
  }

  __device__ void shrink() {
    GRS_NO_CONTRACT
    const int n = s->n;
    if (on) {
      const double v0 = s->v[0][k];
      for (int j = 1; j <= n; j++) {
        const double w = v0 + 0.5 * (s->v[j][k] - v0);
        s->v[j][k] = w;
        s->pend[j - 1][k] = w;
      }
    }
    s->npend = n;
    s->st = grs_nm::SHRINK;
  }

  /* coordinate k of the next request: a xbar - b v_n */
  __device__ void request(double a, double b, grs_nm::State st) {
    GRS_NO_CONTRACT
    if (on) s->pend[0][k] = a * s->xbar[k] - b * s->v[s->n][k];
    s->npend = 1;
    s->st = st;
  }

  __device__ void feed(const double *f) {
    GRS_NO_CONTRACT
    const int n = s->n;
    switch (s->st) {
      case grs_nm::INIT:
        for (int j = 0; j <= n; j++) {
          if (on) s->v[j][k] = s->pend[j][k];
          s->fv[j] = f[j];
        }
        s->nf = n + 1;
        sort();
        s->it = 1;
        begin();
        return;
      case grs_nm::REFLECT: {
        const double fxr = f[0];
        s->fxr = fxr;
        s->nf = s->nf + 1;
        if (fxr < s->fv[0]) {
          request(1.0 + 2.0, 2.0, grs_nm::EXPAND);
          return;
        }
        if (fxr < s->fv[n - 1]) {
          if (on) s->v[n][k] = s->xr[k];
          s->fv[n] = fxr;
          end_iter();
          return;
        }
        if (fxr < s->fv[n]) {
          request(1.0 + 0.5, 0.5, grs_nm::CONTRACT_OUT);
        } else {
          /* (1 - 0.5) xbar + 0.5 v_n */
          if (on) s->pend[0][k] = (1.0 - 0.5) * s->xbar[k] + 0.5 * s->v[n][k];
          s->npend = 1;
          s->st = grs_nm::CONTRACT_IN;
        }
        return;
      }
      case grs_nm::EXPAND:
        s->nf = s->nf + 1;
        if (f[0] < s->fxr) {
          if (on) s->v[n][k] = s->pend[0][k];
          s->fv[n] = f[0];
        } else {
          if (on) s->v[n][k] = s->xr[k];
          s->fv[n] = s->fxr;
        }
        end_iter();
        return;
      case grs_nm::CONTRACT_OUT:
        s->nf = s->nf + 1;
        if (f[0] <= s->fxr) {
          if (on) s->v[n][k] = s->pend[0][k];
          s->fv[n] = f[0];
          end_iter();
        } else {
          shrink();
        }
        return;
      case grs_nm::CONTRACT_IN:
        s->nf = s->nf + 1;
        if (f[0] < s->fv[n]) {
          if (on) s->v[n][k] = s->pend[0][k];
          s->fv[n] = f[0];
          end_iter();
        } else {
          shrink();
        }
        return;
      case grs_nm::SHRINK:
        for (int j = 1; j <= n; j++) s->fv[j] = f[j - 1];
        s->nf = s->nf + n;
        end_iter();
        return;
      case grs_nm::DONE:
        return;
    }
  }
};

#define COM_THREADS 512

/* One block per fit m: the first wave steps the Nelder-Mead state machine
 * (started on the host) in LDS (nm_lanes); for the requested points the first threads
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
    if (t < 64) { /* the first wave steps the method */
      nm_lanes L{nm, t, t < nm->n};
      L.feed(fs);
    }
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
