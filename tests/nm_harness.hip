/*
 * Host-side driver of the engine's resumable Nelder-Mead (grs_nm.h) for the
 * CPU tests: the state machine is run on the host with an objective given
 * by the caller (a ctypes callback), exactly as the engine's batched fits
 * feed it device-computed values.  Built by tests/test_nm_kat.py with hipcc.
 */
#include <cstring>

#include "grs_nm.h"

typedef double (*nm_objective)(const double *u, int n);

extern "C" {

/* fminsearch's iteration from an explicit starting simplex sim0 ((n+1) x n,
 * row-major), objective in the transformed variables. */
int nm_run_simplex(int n, const double *sim0, int max_fe, int max_it, double tol_f, double tol_x, nm_objective f,
                   double *sim_out, double *fv_out, int *nf, int *it) {
  if (n < 1 || n > GRS_NM_NMAX) return -1;
  grs_nm s;
  s.n = n;
  for (int k = 0; k < n; k++) {
    s.lb[k] = -1.0;
    s.ub[k] = 1.0;
  }
  s.max_fe = max_fe;
  s.max_it = max_it;
  s.tol_f = tol_f;
  s.tol_x = tol_x;
  for (int j = 0; j <= n; j++)
    for (int k = 0; k < n; k++) s.pend[j][k] = sim0[j * n + k];
  s.npend = n + 1;
  s.st = grs_nm::INIT;
  double vals[GRS_NM_NMAX + 1];
  while (!s.done()) {
    for (int j = 0; j < s.npend; j++) vals[j] = f(s.pend[j], n);
    s.feed(vals);
  }
  for (int j = 0; j <= n; j++) {
    fv_out[j] = s.fv[j];
    for (int k = 0; k < n; k++) sim_out[j * n + k] = s.v[j][k];
  }
  *nf = s.nf;
  *it = s.it;
  return 0;
}

/* The same iteration driven as the engine's batched fits drive it: every
 * requested reflection is evaluated together with its three follow-ups
 * (grs_nm::follow_ups), and an expansion or contraction the method then
 * requests is fed from those values when it equals one bit for bit.  fed
 * receives the points fed to the method, in order (up to fed_cap of them);
 * hits[0..2] count the expansions, outside and inside contractions fed from
 * the follow-ups. */
int nm_run_follow_ups(int n, const double *sim0, int max_fe, int max_it, double tol_f, double tol_x, nm_objective f,
                      double *sim_out, double *fv_out, int *nf, int *it, double *fed, int fed_cap, int *n_fed,
                      int *hits) {
  if (n < 1 || n > GRS_NM_NMAX) return -1;
  grs_nm s;
  s.n = n;
  for (int k = 0; k < n; k++) {
    s.lb[k] = -1.0;
    s.ub[k] = 1.0;
  }
  s.max_fe = max_fe;
  s.max_it = max_it;
  s.tol_f = tol_f;
  s.tol_x = tol_x;
  for (int j = 0; j <= n; j++)
    for (int k = 0; k < n; k++) s.pend[j][k] = sim0[j * n + k];
  s.npend = n + 1;
  s.st = grs_nm::INIT;
  double vals[GRS_NM_NMAX + 1], fu[3][GRS_NM_NMAX], fuv[3];
  *n_fed = 0;
  hits[0] = hits[1] = hits[2] = 0;
  auto log_fed = [&](const double *u) {
    if (*n_fed < fed_cap)
      for (int k = 0; k < n; k++) fed[*n_fed * n + k] = u[k];
    *n_fed += 1;
  };
  while (!s.done()) {
    const int nfu = s.follow_ups(fu);
    for (int j = 0; j < s.npend; j++) {
      vals[j] = f(s.pend[j], n);
      log_fed(s.pend[j]);
    }
    for (int j = 0; j < nfu; j++) fuv[j] = f(fu[j], n);
    s.feed(vals);
    while (nfu > 0 && !s.done() && s.npend == 1 &&
           (s.st == grs_nm::EXPAND || s.st == grs_nm::CONTRACT_OUT || s.st == grs_nm::CONTRACT_IN)) {
      int hit = -1;
      for (int j = 0; j < nfu && hit < 0; j++)
        if (memcmp(s.pend[0], fu[j], sizeof(double) * (size_t)n) == 0) hit = j;
      if (hit < 0) break;
      hits[s.st == grs_nm::EXPAND ? 0 : (s.st == grs_nm::CONTRACT_OUT ? 1 : 2)] += 1;
      log_fed(s.pend[0]);
      s.feed(&fuv[hit]);
    }
  }
  for (int j = 0; j <= n; j++) {
    fv_out[j] = s.fv[j];
    for (int k = 0; k < n; k++) sim_out[j * n + k] = s.v[j][k];
  }
  *nf = s.nf;
  *it = s.it;
  return 0;
}

/* The starting simplex fminsearchbnd builds for x0 within [lb, ub]. */
int nm_start(int n, const double *x0, const double *lb, const double *ub, double *sim_out) {
  if (n < 1 || n > GRS_NM_NMAX) return -1;
  grs_nm s;
  s.start(n, x0, lb, ub);
  for (int j = 0; j <= n; j++)
    for (int k = 0; k < n; k++) sim_out[j * n + k] = s.pend[j][k];
  return 0;
}

/* Whole fminsearchbnd with the objective in the original variables. */
int nm_fminsearchbnd(int n, const double *x0, const double *lb, const double *ub, int max_fe, int max_it,
                     double tol_f, double tol_x, nm_objective f, double *x, double *fval, int *nf, int *it) {
  if (n < 1 || n > GRS_NM_NMAX) return -1;
  grs_nm s;
  s.max_fe = max_fe;
  s.max_it = max_it;
  s.tol_f = tol_f;
  s.tol_x = tol_x;
  s.start(n, x0, lb, ub);
  double vals[GRS_NM_NMAX + 1], xx[GRS_NM_NMAX];
  while (!s.done()) {
    for (int j = 0; j < s.npend; j++) {
      s.xform(s.pend[j], xx);
      vals[j] = f(xx, n);
    }
    s.feed(vals);
  }
  s.result(x, fval);
  *nf = s.nf;
  *it = s.it;
  return 0;
}

}  // extern "C"
