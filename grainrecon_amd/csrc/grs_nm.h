/*
 * grs_nm.h -- bounded Nelder-Mead as a resumable state machine (host code).
 *
 * fminsearchbnd (fminsearchbnd.m:152-165 start point, :263-268 sine
 * transform for two-sided bounds) over MATLAB's fminsearch.  The method
 * (Lagarias, Reeds, Wright, Wright, SIAM J. Optim. 9 (1998) 112-147, with
 * fminsearch's documented constants) is restated here:
 *   simplex x0 and x0 with coordinate j scaled by 1.05 (0.00025 when 0),
 *   sorted by f (stable, NaN last), iteration count 1;
 *   loop while evaluations < MaxFunEvals and iterations < MaxIter; stop when
 *   max|f1-fj| <= max(TolFun, 10 eps(f1)) and max|vj-v1| <= max(TolX,
 *   10 eps(max(v1))) (max ignores NaN);
 *   xbar = (sum of the n best) / n; reflect 2 xbar - v_end; expand 3 xbar -
 *   2 v_end if the reflection beats the best; keep the reflection if it beats
 *   the second worst; else contract outside 1.5 xbar - 0.5 v_end (reflection
 *   beats the worst; kept if <= the reflection) or inside 0.5 xbar +
 *   0.5 v_end (kept if it beats the worst); otherwise shrink towards the best
 *   by 1/2; sort again.
 * The caller evaluates the points of each request in any order (typically
 * many fits in one device launch) and feeds the values back, so the
 * evaluations of independent fits are batched while each fit follows the
 * sequential method exactly.  sin/asin are the engine's FP64 functions.
 * Everything is host and device code: a fit can also run the whole method
 * inside one kernel, with the same operations and so the same results.
 */
#ifndef GRS_NM_H
#define GRS_NM_H

#include <hip/hip_runtime.h>

#include <cmath>

#include "grs_math.h"

#define GRS_NM_NMAX 8
#define GRS_NM_FN __host__ __device__

/* eps(x) of MATLAB: the spacing of doubles at |x| */
GRS_NM_FN static inline double grs_nm_eps(double x) {
  if (x != x) return x;
  x = fabs(x);
  if (__builtin_isinf(x)) return __builtin_nan("");
  if (x < 2.2250738585072014e-308) return 4.9406564584124654e-324;
  int e;
  frexp(x, &e);
  return ldexp(1.0, e - 53);
}

/* The loops of the per-iteration steps run over the fixed bounds with the
 * live entries selected inside, so that on the device they unroll into
 * independent loads and stores (a kernel steps the method from one lane);
 * the entries and the order of the arithmetic are those of plain loops. */
#define GRS_NM_UNROLL _Pragma("unroll")

GRS_NM_FN static inline void grs_nm_copy(double *d, const double *s, int n) {
  GRS_NM_UNROLL
  for (int k = 0; k < GRS_NM_NMAX; k++)
    if (k < n) d[k] = s[k];
}

/* MATLAB max of two values: NaN ignored unless both are NaN */
GRS_NM_FN static inline double grs_nm_max(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return a > b ? a : b;
}

struct grs_nm {
  enum State { INIT, REFLECT, EXPAND, CONTRACT_OUT, CONTRACT_IN, SHRINK, DONE };
  int n = 0;
  double lb[GRS_NM_NMAX], ub[GRS_NM_NMAX];
  int max_fe = 2000, max_it = 500;
  double tol_f = 1e-10, tol_x = 1e-4;
  double v[GRS_NM_NMAX + 1][GRS_NM_NMAX], fv[GRS_NM_NMAX + 1];
  double xbar[GRS_NM_NMAX], xr[GRS_NM_NMAX], fxr = 0.0;
  int nf = 0, it = 0;
  State st = DONE;
  int npend = 0;
  double pend[GRS_NM_NMAX + 1][GRS_NM_NMAX]; /* requested points, transformed (u) space */

  /* coordinate k of u -> x (fminsearchbnd.m:263-268) */
  GRS_NM_FN void xform1(const double *u, int k, double *xk) const {
    GRS_NO_CONTRACT
    double t = (grs_sin(u[k]) + 1.0) / 2.0;
    t = t * (ub[k] - lb[k]) + lb[k];
    *xk = fmax(lb[k], fmin(ub[k], t));
  }

  GRS_NM_FN void xform(const double *u, double *x) const {
    for (int k = 0; k < n; k++) xform1(u, k, &x[k]);
  }

  GRS_NM_FN void start(int dims, const double *x0, const double *lo, const double *hi) {
    GRS_NO_CONTRACT
    n = dims;
    grs_nm_copy(lb, lo, n);
    grs_nm_copy(ub, hi, n);
    double u0[GRS_NM_NMAX];
    for (int k = 0; k < n; k++) {
      if (x0[k] <= lb[k]) {
        u0[k] = -GRS_PI / 2.0;
      } else if (x0[k] >= ub[k]) {
        u0[k] = GRS_PI / 2.0;
      } else {
        double z = 2.0 * (x0[k] - lb[k]) / (ub[k] - lb[k]) - 1.0;
        z = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
        u0[k] = 2.0 * GRS_PI + grs_asin(z);
      }
    }
    grs_nm_copy(pend[0], u0, n);
    for (int j = 0; j < n; j++) {
      grs_nm_copy(pend[j + 1], u0, n);
      pend[j + 1][j] = u0[j] != 0.0 ? (1.0 + 0.05) * u0[j] : 0.00025;
    }
    npend = n + 1;
    st = INIT;
  }

  GRS_NM_FN bool done() const { return st == DONE; }

  GRS_NM_FN void sort_simplex() {
    /* stable insertion sort, NaN last */
    for (int i = 1; i <= n; i++) {
      double f = fv[i], row[GRS_NM_NMAX];
      grs_nm_copy(row, v[i], n);
      int j = i - 1;
      while (j >= 0 && (f == f) && (fv[j] != fv[j] || fv[j] > f)) {
        fv[j + 1] = fv[j];
        grs_nm_copy(v[j + 1], v[j], n);
        j--;
      }
      fv[j + 1] = f;
      grs_nm_copy(v[j + 1], row, n);
    }
  }

  GRS_NM_FN void begin_iteration() {
    GRS_NO_CONTRACT
    if (!(nf < max_fe && it < max_it)) {
      st = DONE;
      npend = 0;
      return;
    }
    double dfv = __builtin_nan(""), dx = __builtin_nan(""), vmax = __builtin_nan("");
    GRS_NM_UNROLL
    for (int j = 1; j <= GRS_NM_NMAX; j++)
      if (j <= n) dfv = grs_nm_max(dfv, fabs(fv[0] - fv[j]));
    GRS_NM_UNROLL
    for (int j = 1; j <= GRS_NM_NMAX; j++) {
      GRS_NM_UNROLL
      for (int k = 0; k < GRS_NM_NMAX; k++)
        if (j <= n && k < n) dx = grs_nm_max(dx, fabs(v[j][k] - v[0][k]));
    }
    GRS_NM_UNROLL
    for (int k = 0; k < GRS_NM_NMAX; k++)
      if (k < n) vmax = grs_nm_max(vmax, v[0][k]);
    if (dfv <= grs_nm_max(tol_f, 10.0 * grs_nm_eps(fv[0])) && dx <= grs_nm_max(tol_x, 10.0 * grs_nm_eps(vmax))) {
      st = DONE;
      npend = 0;
      return;
    }
    GRS_NM_UNROLL
    for (int k = 0; k < GRS_NM_NMAX; k++) {
      double acc = 0.0;
      GRS_NM_UNROLL
      for (int j = 0; j < GRS_NM_NMAX; j++)
        if (j < n) acc += v[j][k];
      if (k < n) xbar[k] = acc / (double)n;
    }
    GRS_NM_UNROLL
    for (int k = 0; k < GRS_NM_NMAX; k++)
      if (k < n) xr[k] = (1.0 + 1.0) * xbar[k] - 1.0 * v[n][k];
    grs_nm_copy(pend[0], xr, n);
    npend = 1;
    st = REFLECT;
  }

  /* While a reflection is requested: the three points its value can lead to
   * next (expansion, outside and inside contraction), computed with the
   * statements feed() uses, so that a caller can evaluate them in the same
   * batch; returns how many (0 in any other state).  The method itself is
   * unchanged: feed() still requests each of them when the method needs it. */
  GRS_NM_FN int follow_ups(double (*u)[GRS_NM_NMAX]) const {
    GRS_NO_CONTRACT
    if (st != REFLECT) return 0;
    for (int k = 0; k < n; k++) {
      u[0][k] = (1.0 + 2.0) * xbar[k] - 2.0 * v[n][k];
      u[1][k] = (1.0 + 0.5) * xbar[k] - 0.5 * v[n][k];
      u[2][k] = (1.0 - 0.5) * xbar[k] + 0.5 * v[n][k];
    }
    return 3;
  }

  GRS_NM_FN void end_iteration() {
    sort_simplex();
    it += 1;
    begin_iteration();
  }

  GRS_NM_FN void request_shrink() {
    GRS_NO_CONTRACT
    for (int j = 1; j <= n; j++) {
      for (int k = 0; k < n; k++) v[j][k] = v[0][k] + 0.5 * (v[j][k] - v[0][k]);
      grs_nm_copy(pend[j - 1], v[j], n);
    }
    npend = n;
    st = SHRINK;
  }

  /* values of the npend requested points, in request order */
  GRS_NM_FN void feed(const double *f) {
    GRS_NO_CONTRACT
    switch (st) {
      case INIT:
        for (int j = 0; j <= n; j++) {
          grs_nm_copy(v[j], pend[j], n);
          fv[j] = f[j];
        }
        nf = n + 1;
        sort_simplex();
        it = 1;
        begin_iteration();
        return;
      case REFLECT:
        fxr = f[0];
        nf += 1;
        if (fxr < fv[0]) {
          GRS_NM_UNROLL
          for (int k = 0; k < GRS_NM_NMAX; k++)
            if (k < n) pend[0][k] = (1.0 + 2.0) * xbar[k] - 2.0 * v[n][k];
          npend = 1;
          st = EXPAND;
          return;
        }
        if (fxr < fv[n - 1]) {
          grs_nm_copy(v[n], xr, n);
          fv[n] = fxr;
          end_iteration();
          return;
        }
        if (fxr < fv[n]) {
          GRS_NM_UNROLL
          for (int k = 0; k < GRS_NM_NMAX; k++)
            if (k < n) pend[0][k] = (1.0 + 0.5) * xbar[k] - 0.5 * v[n][k];
          npend = 1;
          st = CONTRACT_OUT;
        } else {
          GRS_NM_UNROLL
          for (int k = 0; k < GRS_NM_NMAX; k++)
            if (k < n) pend[0][k] = (1.0 - 0.5) * xbar[k] + 0.5 * v[n][k];
          npend = 1;
          st = CONTRACT_IN;
        }
        return;
      case EXPAND:
        nf += 1;
        if (f[0] < fxr) {
          grs_nm_copy(v[n], pend[0], n);
          fv[n] = f[0];
        } else {
          grs_nm_copy(v[n], xr, n);
          fv[n] = fxr;
        }
        end_iteration();
        return;
      case CONTRACT_OUT:
        nf += 1;
        if (f[0] <= fxr) {
          grs_nm_copy(v[n], pend[0], n);
          fv[n] = f[0];
          end_iteration();
        } else {
          request_shrink();
        }
        return;
      case CONTRACT_IN:
        nf += 1;
        if (f[0] < fv[n]) {
          grs_nm_copy(v[n], pend[0], n);
          fv[n] = f[0];
          end_iteration();
        } else {
          request_shrink();
        }
        return;
      case SHRINK:
        for (int j = 1; j <= n; j++) fv[j] = f[j - 1];
        nf += n;
        end_iteration();
        return;
      case DONE:
        return;
    }
  }

  /* result in the original variables */
  GRS_NM_FN void result(double *x, double *fval) const {
    xform(v[0], x);
    *fval = fv[0];
  }
};

#endif /* GRS_NM_H */
