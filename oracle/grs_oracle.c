/*
 * grs_oracle.c -- CPU FP64 restatement of GrainRecon's forward-completeness
 * scorer.  TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP
 * engine.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * It follows the MATLAB code statement by statement (file:line below are in
 * the GrainRecon repository):
 *   forward_comp.m:4-90            -> oracle_forward_comp()
 *   ForwardCalc_OR_space_comp.m:22-53 (q0 recompute, quaternion -> U)
 *                                  -> oracle_or_space_comp()
 *   grow_indexed_region_parallel.m:61-124 + ind2pos.m:6-10 (growth rule)
 *                                  -> oracle_grow()
 *   revise_all_unindexed_voxel.m:138-153 + euler2u.m (argmax over candidates)
 *                                  -> oracle_compete()
 * Like the vectorised MATLAB, every quantity is computed for every (omega,
 * hkl) event before the selection mask is applied; the median is taken by
 * sorting the list of selected distances (forward_comp.m:79-90).
 *
 * Arithmetic conventions (the MATLAB built-ins are not available to run; the
 * choices below define the semantics and are shared with the engine):
 *  - matrix products and dot products are sequential sums in index order
 *    without fused multiply-add; the structural zeros and the 1 of the
 *    omega rotation matrix are not multiplied out;
 *  - x.^2 is x*x, x.^0.5, x.^(1/2) and norm() are sqrt of the sequential sum
 *    of squares;
 *  - acos, sin, cos, atan are the project's FP64 implementations
 *    (grainrecon_amd/csrc/grs_math.h, < 1 ulp), or the C library's when
 *    use_libm is set (used only to measure decision sensitivity);
 *  - round() is halfway-away-from-zero (MATLAB round);
 *  - a distance is a hit when (double)dis <= minEucDis/mean([py pz]);
 *  - median of an even count is a + (b - a)/2 in single precision.
 *
 * Build: oracle/build_oracle.sh (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../grainrecon_amd/csrc/grs_math.h"

static int g_use_libm = 0;

void oracle_set_libm(int on) { g_use_libm = on; }

/* lm != 0: the C library's functions, else the project's */
static double m_acos(int lm, double x) { return lm ? acos(x) : grs_acos(x); }
static double m_sin(int lm, double x) { return lm ? sin(x) : grs_sin(x); }
static double m_cos(int lm, double x) { return lm ? cos(x) : grs_cos(x); }
static double m_atan(int lm, double x) { return lm ? atan(x) : grs_atan(x); }
static double o_sin(double x) { return m_sin(g_use_libm, x); }
static double o_cos(double x) { return m_cos(g_use_libm, x); }

/* Geometry and selection parameters, the argument list of forward_comp.m:4-6. */
typedef struct {
  double S[9], B[9], RotDet[9]; /* row-major 3x3 */
  double thetamax, lambda_min, lambda_max;
  double Lsam2sou, Lsam2det, minEucDis;
  double dety00, detz00, P0y, P0z, RotAxisOffset;
  double pixelysize, pixelzsize, dety0, detz0, detysize, detzsize;
  double BeamStopY[2], BeamStopZ[2];
} oracle_geo;

/* The distance stack proj_bin_bw(detz, dety, i), MATLAB column-major. */
typedef struct {
  const float *proj;
  int64_t nz, ny, nrot;
} oracle_proj;

static void matmul3(const double *A, const double *Bm, double *C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[i * 3 + j] = (A[i * 3 + 0] * Bm[0 * 3 + j] + A[i * 3 + 1] * Bm[1 * 3 + j]) + A[i * 3 + 2] * Bm[2 * 3 + j];
}

static int cmp_float(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}

/* Gw = S*U*B*hkl (forward_comp.m:15-16), evaluated ((S*U)*B)*hkl; Gw[h*3 + r] */
static void oracle_Gw(const oracle_geo *g, const double *U, const double *hkl, int64_t H, double *Gw) {
  double SU[9], SUB[9];
  matmul3(g->S, U, SU);
  matmul3(SU, g->B, SUB);
  for (int64_t h = 0; h < H; h++) {
    const double *x = hkl + 3 * h;
    for (int r = 0; r < 3; r++) Gw[h * 3 + r] = (SUB[r * 3 + 0] * x[0] + SUB[r * 3 + 1] * x[1]) + SUB[r * 3 + 2] * x[2];
  }
}

/* Every event of one rotation angle (forward_comp.m:19-59) for the shifted
 * position pos (:17), with the math of mode lm: beta, lambdahkl, the
 * detector coordinates dety22/detz22 (mm) and the unrounded pixel
 * coordinates -dety22/pixelysize + dety0 (and z), H each. */
static void omega_events(const oracle_geo *g, const double *pos, double rot_deg, const double *Gw, int64_t H, int lm,
                         double *beta, double *lam, double *d22y, double *d22z, double *yraw, double *zraw) {
  const double pi = GRS_PI;
  const double L = g->Lsam2sou + g->Lsam2det;
  const double Lss = g->Lsam2sou, Lsd = g->Lsam2det, RAO = g->RotAxisOffset;
  const double *R = g->RotDet;
  const double pio2 = pi / 2.0;
  double omega = (rot_deg * pi) / 180.0; /* :19 */
  double c = m_cos(lm, omega), s = m_sin(lm, omega);
  /* Omega = [c -s 0; s c 0; 0 0 1] (:20); Omega*S without the structural zeros */
  double OS[9];
  for (int j = 0; j < 3; j++) {
    OS[0 * 3 + j] = c * g->S[0 * 3 + j] - s * g->S[1 * 3 + j];
    OS[1 * 3 + j] = s * g->S[0 * 3 + j] + c * g->S[1 * 3 + j];
    OS[2 * 3 + j] = g->S[2 * 3 + j];
  }
  double SW[3];
  for (int r = 0; r < 3; r++) SW[r] = (OS[r * 3 + 0] * pos[0] + OS[r * 3 + 1] * pos[1]) + OS[r * 3 + 2] * pos[2]; /* :22 */
  double gp1 = Lss + SW[0];
  double gp2 = (SW[1] - g->P0y) + RAO;
  double gp3 = SW[2] - g->P0z;
  double center[3] = {L, (gp2 * L) / gp1, (gp3 * L) / gp1};    /* :23-24 */
  double alpha = m_atan(lm, sqrt(gp2 * gp2 + gp3 * gp3) / gp1); /* :25 */
  double ng = sqrt((gp1 * gp1 + gp2 * gp2) + gp3 * gp3);        /* norm(grainpos) :31 */
  double u[3] = {gp1 / ng, gp2 / ng, gp3 / ng};
  double nc = sqrt((center[0] * center[0] + center[1] * center[1]) + center[2] * center[2]);
  double cu[3] = {center[0] / nc, center[1] / nc, center[2] / nc};
  double L2 = (Lsd - SW[0]) / m_cos(lm, alpha); /* :42 */
  double tnum = (R[0] * (Lsd - SW[0]) + R[3] * ((g->dety00 - RAO) - SW[1])) + R[6] * (g->detz00 - SW[2]); /* :51 */
  double ov[3] = {SW[0] - Lsd, (SW[1] - g->dety00) + RAO, SW[2] - g->detz00};                            /* :54 */

  for (int64_t h = 0; h < H; h++) {
    const double *gw = Gw + 3 * h;
    double Gt1 = c * gw[0] - s * gw[1]; /* Gt = Omega*Gw (:28) */
    double Gt2 = s * gw[0] + c * gw[1];
    double Gt3 = gw[2];
    double Glen = sqrt((Gt1 * Gt1 + Gt2 * Gt2) + Gt3 * Gt3);                     /* :30 */
    double cb = (u[0] * (Gt1 / Glen) + u[1] * (Gt2 / Glen)) + u[2] * (Gt3 / Glen); /* :31 */
    double b = m_acos(lm, cb);
    double theta = b - pio2;                     /* :33 */
    double sintth = m_sin(lm, 2.0 * theta);      /* :34 */
    double d = ((1.0 / Glen) * 2.0) * pi;        /* :36 */
    double lambdahkl = (2.0 * d) * m_sin(lm, theta); /* :37 */
    double nv = sqrt(Gt2 * Gt2 + Gt3 * Gt3);      /* |v1|, v1 = [0 Gt2 Gt3] (:29,40) */
    double phix = m_acos(lm, (Gt2 / nv) * cu[1] + (Gt3 / nv) * cu[2]); /* :40 */
    double phiy = phix - 2.0 * theta;            /* :41 */
    double diffvec = (L2 * sintth) / m_sin(lm, phiy); /* :43 */
    double konst = sqrt(Gt2 * Gt2 + Gt3 * Gt3);   /* :44 */
    double y22 = center[1] + ((diffvec * Gt2) / konst); /* :46 */
    double z22 = center[2] + ((diffvec * Gt3) / konst); /* :47 */
    double k0 = Lsd - SW[0], k1 = y22 - SW[1], k2 = z22 - SW[2]; /* :49 */
    double nk = sqrt((k0 * k0 + k1 * k1) + k2 * k2);             /* :50 */
    double K0 = k0 / nk, K1 = k1 / nk, K2 = k2 / nk;
    double t = tnum / ((R[0] * K0 + R[3] * K1) + R[6] * K2);     /* :51-52 */
    double w0 = t * K0 + ov[0], w1 = t * K1 + ov[1], w2 = t * K2 + ov[2];
    double Y22 = (R[1] * w0 + R[4] * w1) + R[7] * w2; /* :54 */
    double Z22 = (R[2] * w0 + R[5] * w1) + R[8] * w2; /* :55 */
    yraw[h] = (-Y22) / g->pixelysize + g->dety0; /* :58, before round */
    zraw[h] = (-Z22) / g->pixelzsize + g->detz0; /* :59 */
    beta[h] = b;
    lam[h] = lambdahkl;
    d22y[h] = Y22;
    d22z[h] = Z22;
  }
}

/* forward_comp.m:4-90 for one (U, pos).  Returns the counts and the median;
 * optionally records every selected event (omega index, hkl index, dety,
 * detz, dety22, detz22, dis, lambdahkl) into ev (capacity ev_cap, 8 doubles each). */
int64_t oracle_forward_comp(const oracle_geo *g, const oracle_proj *pj, const double *U, const double *pos_in,
                            const double *rot_angles, int64_t nrot, const double *hkl, int64_t nrhkl,
                            int64_t *Nr_simu_out, int64_t *Nr_intersect_out, float *dis_median_out,
                            double *ev, int64_t ev_cap) {
  const double pi = GRS_PI;
  const double RAO = g->RotAxisOffset;
  int64_t Nr_simu = 0, Nr_intersect = 0, nev = 0;
  int64_t H = nrhkl;
  double *buf = (double *)malloc(sizeof(double) * 3 * H * 8);
  double *Gw = buf;           /* 3*H, [h*3 + r] */
  double *dety = buf + 3 * H; /* H */
  double *detz = dety + H;
  double *beta = detz + H;
  double *lam = beta + H;
  double *d22y = lam + H;
  double *d22z = d22y + H;
  float *dis_list = (float *)malloc(sizeof(float) * (nrot * H + 1));
  int64_t ndis = 0;

  oracle_Gw(g, U, hkl, H, Gw);
  double pos[3] = {pos_in[0], pos_in[1] - RAO, pos_in[2]}; /* :17 */
  const double thr = g->minEucDis / ((g->pixelysize + g->pixelzsize) / 2.0);
  const double beta_max = ((90.0 + g->thetamax * 4.0) / 180.0) * pi;
  const double pio2 = pi / 2.0;

  for (int64_t i = 0; i < nrot; i++) {
    omega_events(g, pos, rot_angles[i], Gw, H, g_use_libm, beta, lam, d22y, d22z, dety, detz);
    for (int64_t h = 0; h < H; h++) {
      dety[h] = grs_round(dety[h]); /* :58 */
      detz[h] = grs_round(detz[h]); /* :59 */
    }
    /* select3 (:61-65) and the hit loop (:66-82) */
    for (int64_t h = 0; h < H; h++) {
      double dy = dety[h], dz = detz[h];
      int sel = beta[h] > pio2 && beta[h] < beta_max && lam[h] > g->lambda_min && lam[h] < g->lambda_max &&
                dy >= 1.0 && dy <= g->detysize && dz >= 1.0 && dz <= g->detzsize &&
                !(dy >= g->BeamStopY[0] && dy <= g->BeamStopY[1] && dz >= g->BeamStopZ[0] && dz <= g->BeamStopZ[1]);
      if (!sel) continue;
      Nr_simu++;
      int64_t iz = (int64_t)dz - 1, iy = (int64_t)dy - 1;
      float dis;
      if (iz < pj->nz && iy < pj->ny && i < pj->nrot)
        dis = pj->proj[iz + iy * pj->nz + i * pj->nz * pj->ny];
      else
        dis = NAN; /* index beyond the stack: MATLAB would error */
      if ((double)dis <= thr) Nr_intersect++;
      dis_list[ndis++] = dis;
      if (ev && nev < ev_cap) {
        double *e = ev + 8 * nev;
        e[0] = (double)i; e[1] = (double)h; e[2] = dy; e[3] = dz; e[4] = d22y[h]; e[5] = d22z[h]; e[6] = dis;
        e[7] = lam[h];
      }
      nev++;
    }
  }
  float med = 2000.0f;
  if (ndis > 0) {
    qsort(dis_list, (size_t)ndis, sizeof(float), cmp_float);
    if (ndis & 1) {
      med = dis_list[ndis / 2];
    } else {
      float a = dis_list[ndis / 2 - 1], b = dis_list[ndis / 2];
      med = a + (b - a) / 2.0f;
    }
  }
  *Nr_simu_out = Nr_simu;
  *Nr_intersect_out = Nr_intersect;
  *dis_median_out = med;
  free(dis_list);
  free(buf);
  return nev;
}

/* quaternion2U.m:3-28 / ForwardCalc_OR_space_comp.m:33-49 */
static void quat_to_U(double q0, double q1, double q2, double q3, double *U) {
  U[0] = 2.0 * (q0 * q0 + q1 * q1) - 1.0;
  U[1] = 2.0 * (q1 * q2 - q0 * q3);
  U[2] = 2.0 * (q1 * q3 + q0 * q2);
  U[3] = 2.0 * (q1 * q2 + q0 * q3);
  U[4] = 2.0 * (q0 * q0 + q2 * q2) - 1.0;
  U[5] = 2.0 * (q2 * q3 - q0 * q1);
  U[6] = 2.0 * (q1 * q3 - q0 * q2);
  U[7] = 2.0 * (q2 * q3 + q0 * q1);
  U[8] = 2.0 * (q0 * q0 + q3 * q3) - 1.0;
}

/* euler2u.m:5-16 with angles in degrees converted as (deg*pi)/180. */
void oracle_euler2u_deg(double p1d, double Pd, double p2d, double *U) {
  const double pi = GRS_PI;
  double phi1 = (p1d * pi) / 180.0, PHI = (Pd * pi) / 180.0, phi2 = (p2d * pi) / 180.0;
  double c1 = o_cos(phi1), s1 = o_sin(phi1), c2 = o_cos(phi2), s2 = o_sin(phi2), cP = o_cos(PHI), sP = o_sin(PHI);
  U[0] = c1 * c2 - (s1 * s2) * cP;
  U[3] = s1 * c2 + (c1 * s2) * cP;
  U[6] = s2 * sP;
  U[1] = -c1 * s2 - (s1 * c2) * cP;
  U[4] = -s1 * s2 + (c1 * c2) * cP;
  U[7] = c2 * sP;
  U[2] = s1 * sP;
  U[5] = -c1 * sP;
  U[8] = cP;
}

/* ForwardCalc_OR_space_comp.m:22-53: rows of [q0 q1 q2 q3]; q0 is recomputed
 * from the vector part and rows with |q_vec|^2 > 1 are left as zeros.
 * out: N x 4 doubles [Nr_simu Nr_intersect C dis_median]. */
void oracle_or_space_comp(const oracle_geo *g, const oracle_proj *pj, const double *pos, const double *q, int64_t N,
                          const double *rot_angles, int64_t nrot, const double *hkl, int64_t nrhkl, double *out,
                          int nthreads) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t i = 0; i < N; i++) {
    double q1 = q[4 * i + 1], q2 = q[4 * i + 2], q3 = q[4 * i + 3];
    double ss = (q1 * q1 + q2 * q2) + q3 * q3;
    double *o = out + 4 * i;
    o[0] = o[1] = o[2] = o[3] = 0.0;
    if (!(ss <= 1.0)) continue;
    double q0 = sqrt(1.0 - ss);
    double U[9];
    quat_to_U(q0, q1, q2, q3, U);
    int64_t ns, ni;
    float med;
    oracle_forward_comp(g, pj, U, pos, rot_angles, nrot, hkl, nrhkl, &ns, &ni, &med, NULL, 0);
    o[0] = (double)ns;
    o[1] = (double)ni;
    o[2] = (double)ni / (double)ns;
    o[3] = (double)med;
  }
}

static float median_sorted(float *v, int64_t n) {
  if (n <= 0) return 2000.0f;
  qsort(v, (size_t)n, sizeof(float), cmp_float);
  if (n & 1) return v[n / 2];
  float a = v[n / 2 - 1], b = v[n / 2];
  return a + (b - a) / 2.0f;
}

static double dmin(double a, double b) { return b < a ? b : a; }
static double dmax(double a, double b) { return b > a ? b : a; }
static double ulp_of(double x) { x = fabs(x); return nextafter(x, INFINITY) - x; }
/* margin / max(|difference|, ulp(x)) */
static double sratio(double margin, double diff, double x) { return margin / dmax(fabs(diff), ulp_of(x)); }

/* Decision-margin audit of forward_comp.m's continuous decisions for one
 * voxel and N orientation rows (as oracle_or_space_comp): every event is
 * computed with the project's math (p) and with the C library's (c).
 * out[0]  events examined
 * out[1]  events whose selection or pixel differs between p and c
 * out[2]  orientations whose Nr_simu, Nr_intersect or median differ
 * out[3]  max |beta_p - beta_c| (rad)
 * out[4]  max |lam_p - lam_c| / lam_p
 * out[5]  max |raw pixel coordinate p - c| (px), events passing the windows
 * out[6]  min |beta - pi/2| (rad) over events whose other tests pass
 * out[7]  min |beta - beta_max| (rad), same condition
 * out[8]  min |lam - lambda_min| / lambda_min, same condition
 * out[9]  min |lam - lambda_max| / lambda_max, same condition
 * out[10] min distance of a raw pixel coordinate from a half-integer (px)
 *         over events passing the windows with both raw coordinates within
 *         1.5 px of the detector area (a changed rounding could change the
 *         pixel, the bounds or the beamstop test)
 * out[11] events passing the window tests (p)
 * out[12..16] per-event ratios margin / max(|p - c|, 1 ulp) of the same five
 *         decisions (beta low, beta high, lambda_min, lambda_max, rounding):
 *         how many times the difference two < 1-ulp libraries make at that
 *         event fits into its own margin (the smallest over the events) */
void oracle_margin_audit(const oracle_geo *g, const oracle_proj *pj, const double *pos_in, const double *q, int64_t N,
                         const double *rot_angles, int64_t nrot, const double *hkl, int64_t H, double *out,
                         int nthreads) {
  const double pi = GRS_PI;
  const double pio2 = pi / 2.0;
  const double beta_max = ((90.0 + g->thetamax * 4.0) / 180.0) * pi;
  const double thr = g->minEucDis / ((g->pixelysize + g->pixelzsize) / 2.0);
  const double pos[3] = {pos_in[0], pos_in[1] - g->RotAxisOffset, pos_in[2]};
  double n_ev = 0, n_ev_diff = 0, n_or_diff = 0, n_win = 0;
  double mx_b = 0, mx_l = 0, mx_px = 0;
  double mn_b0 = INFINITY, mn_b1 = INFINITY, mn_l0 = INFINITY, mn_l1 = INFINITY, mn_rnd = INFINITY;
  double rr_b0 = INFINITY, rr_b1 = INFINITY, rr_l0 = INFINITY, rr_l1 = INFINITY, rr_rnd = INFINITY;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) \
    reduction(+ : n_ev, n_ev_diff, n_or_diff, n_win) reduction(max : mx_b, mx_l, mx_px) \
    reduction(min : mn_b0, mn_b1, mn_l0, mn_l1, mn_rnd, rr_b0, rr_b1, rr_l0, rr_l1, rr_rnd)
  for (int64_t i = 0; i < N; i++) {
    double q1 = q[4 * i + 1], q2 = q[4 * i + 2], q3 = q[4 * i + 3];
    double ss = (q1 * q1 + q2 * q2) + q3 * q3;
    if (!(ss <= 1.0)) continue;
    double U[9];
    quat_to_U(sqrt(1.0 - ss), q1, q2, q3, U);
    double *buf = (double *)malloc(sizeof(double) * H * 15);
    double *Gw = buf, *a = buf + 3 * H; /* a: 2 modes x 6 arrays of H */
    float *dl[2];
    dl[0] = (float *)malloc(sizeof(float) * (nrot * H + 1));
    dl[1] = (float *)malloc(sizeof(float) * (nrot * H + 1));
    int64_t ns[2] = {0, 0}, ni[2] = {0, 0};
    oracle_Gw(g, U, hkl, H, Gw);
    for (int64_t om = 0; om < nrot; om++) {
      for (int lm = 0; lm < 2; lm++) {
        double *b = a + lm * 6 * H;
        omega_events(g, pos, rot_angles[om], Gw, H, lm, b, b + H, b + 2 * H, b + 3 * H, b + 4 * H, b + 5 * H);
      }
      for (int64_t h = 0; h < H; h++) {
        int sel[2], win[2];
        double dy[2], dz[2];
        for (int lm = 0; lm < 2; lm++) {
          const double *b = a + lm * 6 * H;
          double be = b[h], la = b[H + h];
          dy[lm] = grs_round(b[4 * H + h]);
          dz[lm] = grs_round(b[5 * H + h]);
          win[lm] = be > pio2 && be < beta_max && la > g->lambda_min && la < g->lambda_max;
          int ondet = dy[lm] >= 1.0 && dy[lm] <= g->detysize && dz[lm] >= 1.0 && dz[lm] <= g->detzsize &&
                      !(dy[lm] >= g->BeamStopY[0] && dy[lm] <= g->BeamStopY[1] && dz[lm] >= g->BeamStopZ[0] &&
                        dz[lm] <= g->BeamStopZ[1]);
          sel[lm] = win[lm] && ondet;
          if (sel[lm]) {
            ns[lm]++;
            int64_t iz = (int64_t)dz[lm] - 1, iy = (int64_t)dy[lm] - 1;
            float dis = (iz < pj->nz && iy < pj->ny && om < pj->nrot) ? pj->proj[iz + iy * pj->nz + om * pj->nz * pj->ny]
                                                                      : NAN;
            if ((double)dis <= thr) ni[lm]++;
            dl[lm][ns[lm] - 1] = dis;
          }
        }
        const double *p = a, *c = a + 6 * H;
        n_ev += 1;
        if (sel[0] != sel[1] || (sel[0] && (dy[0] != dy[1] || dz[0] != dz[1]))) n_ev_diff += 1;
        mx_b = dmax(mx_b, fabs(p[h] - c[h]));
        if (p[H + h] == p[H + h] && p[H + h] != 0.0) mx_l = dmax(mx_l, fabs(p[H + h] - c[H + h]) / fabs(p[H + h]));
        double be = p[h], la = p[H + h];
        int ondet = dy[0] >= 1.0 && dy[0] <= g->detysize && dz[0] >= 1.0 && dz[0] <= g->detzsize &&
                    !(dy[0] >= g->BeamStopY[0] && dy[0] <= g->BeamStopY[1] && dz[0] >= g->BeamStopZ[0] &&
                      dz[0] <= g->BeamStopZ[1]);
        int t0 = be > pio2, t1 = be < beta_max, t2 = la > g->lambda_min, t3 = la < g->lambda_max;
        const double db = p[h] - c[h], dl_ = p[H + h] - c[H + h];
        if (t1 && t2 && t3 && ondet) {
          mn_b0 = dmin(mn_b0, fabs(be - pio2));
          rr_b0 = dmin(rr_b0, sratio(fabs(be - pio2), db, be));
        }
        if (t0 && t2 && t3 && ondet) {
          mn_b1 = dmin(mn_b1, fabs(be - beta_max));
          rr_b1 = dmin(rr_b1, sratio(fabs(be - beta_max), db, be));
        }
        if (t0 && t1 && t3 && ondet) {
          mn_l0 = dmin(mn_l0, fabs(la - g->lambda_min) / g->lambda_min);
          rr_l0 = dmin(rr_l0, sratio(fabs(la - g->lambda_min), dl_, la));
        }
        if (t0 && t1 && t2 && ondet) {
          mn_l1 = dmin(mn_l1, fabs(la - g->lambda_max) / g->lambda_max);
          rr_l1 = dmin(rr_l1, sratio(fabs(la - g->lambda_max), dl_, la));
        }
        if (win[0]) {
          n_win += 1;
          double y = p[4 * H + h], z = p[5 * H + h];
          mx_px = dmax(mx_px, dmax(fabs(y - c[4 * H + h]), fabs(z - c[5 * H + h])));
          if (y >= -1.0 && y <= g->detysize + 2.0 && z >= -1.0 && z <= g->detzsize + 2.0) {
            double ry = fabs(y - (floor(y) + 0.5)), rz = fabs(z - (floor(z) + 0.5));
            mn_rnd = dmin(mn_rnd, dmin(ry, rz));
            rr_rnd = dmin(rr_rnd, dmin(sratio(ry, y - c[4 * H + h], y), sratio(rz, z - c[5 * H + h], z)));
          }
        }
      }
    }
    float m0 = median_sorted(dl[0], ns[0]), m1 = median_sorted(dl[1], ns[1]);
    if (ns[0] != ns[1] || ni[0] != ni[1] || !(m0 == m1 || (m0 != m0 && m1 != m1))) n_or_diff += 1;
    free(dl[0]);
    free(dl[1]);
    free(buf);
  }
  out[0] = n_ev;
  out[1] = n_ev_diff;
  out[2] = n_or_diff;
  out[3] = mx_b;
  out[4] = mx_l;
  out[5] = mx_px;
  out[6] = mn_b0;
  out[7] = mn_b1;
  out[8] = mn_l0;
  out[9] = mn_l1;
  out[10] = mn_rnd;
  out[11] = n_win;
  out[12] = rr_b0;
  out[13] = rr_b1;
  out[14] = rr_l0;
  out[15] = rr_l1;
  out[16] = rr_rnd;
}

/* Voxel index -> lab position (ind2pos.m:6-10, grow_indexed_region_parallel.m:70-74).
 * p1 = [RecVolumePixel(:,1)' (3) simap VoxSize Center(3) Dimension(3)] */
static void voxel_pos(const double *v, const double *p1, double *pos) {
  for (int k = 0; k < 3; k++) pos[k] = ((v[k] + p1[k] - 1.0) - p1[8 + k] / 2.0) * p1[4] + p1[5 + k];
  if (p1[3] == 1.0) {
    pos[0] = -pos[0];
    pos[1] = -pos[1];
  }
}

/* Growth scoring over a voxel list (grow_indexed_region_parallel.m:61-124 with
 * the engine's 6-column input): vox M x 6 [x y z C mask DisMedian],
 * p1 = 14 values [RecVolumePixel(:,1)' simap VoxSize Center' Dimension seed_comp drop_off maxDmedian].
 * out M x 7 [x y z Nr_simu Nr_intersect dis_median growth_flag]. */
void oracle_grow(const oracle_geo *g, const oracle_proj *pj, const double *vox, int64_t M, const double *U,
                 const double *rot_angles, int64_t nrot, const double *hkl, int64_t nrhkl, const double *p1,
                 double *out, int nthreads) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t i = 0; i < M; i++) {
    const double *v = vox + 6 * i;
    double *o = out + 7 * i;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
    o[3] = o[4] = o[5] = o[6] = 0.0;
    if (!(v[4] > 0.0)) continue;
    double pos[3];
    voxel_pos(v, p1, pos);
    int64_t ns, ni;
    float med;
    oracle_forward_comp(g, pj, U, pos, rot_angles, nrot, hkl, nrhkl, &ns, &ni, &med, NULL, 0);
    double C = (double)ni / (double)ns;
    double dm = (double)med;
    double lim = v[5] < p1[13] ? v[5] : p1[13];
    int flag = (C > v[3] && dm <= lim) && (C >= p1[11] * (1.0 - p1[12]));
    o[3] = (double)ns;
    o[4] = (double)ni;
    o[5] = dm;
    o[6] = flag ? 1.0 : 0.0;
  }
}

/* Compete (revise_all_unindexed_voxel.m:138-153): for each voxel row
 * [x y z start end] (1-based inclusive into id_all), score every candidate
 * grain's orientation (Euler, degrees, 3 per grain id) and keep the first
 * maximum of C (NaN never wins).  out M x 8 [i phi1 PHI phi2 gid C Nr_intersect dis_median]. */
void oracle_compete(const oracle_geo *g, const oracle_proj *pj, const double *vox, int64_t M, const double *id_all,
                    const double *euler, const double *rot_angles, int64_t nrot, const double *hkl, int64_t nrhkl,
                    const double *p1, double *out, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t i = 0; i < M; i++) {
    const double *v = vox + 5 * i;
    double *o = out + 8 * i;
    double pos[3];
    voxel_pos(v, p1, pos);
    int64_t a = (int64_t)v[3], b = (int64_t)v[4];
    int64_t best = -1;
    double bestC = 0.0, best_ni = 0.0, best_med = 0.0;
    double firstC = 0.0, first_ni = 0.0, first_med = 0.0;
    for (int64_t j = a; j <= b; j++) {
      int64_t gid = (int64_t)id_all[j - 1];
      const double *e = euler + 3 * (gid - 1);
      double U[9];
      oracle_euler2u_deg(e[0], e[1], e[2], U);
      int64_t ns, ni;
      float med;
      oracle_forward_comp(g, pj, U, pos, rot_angles, nrot, hkl, nrhkl, &ns, &ni, &med, NULL, 0);
      double C = (double)ni / (double)ns;
      if (j == a) {
        firstC = C; first_ni = (double)ni; first_med = (double)med;
      }
      /* MATLAB max: NaN is skipped, the first maximum wins */
      if (C == C && (best < 0 || C > bestC)) {
        best = j;
        bestC = C;
        best_ni = (double)ni;
        best_med = (double)med;
      }
    }
    if (best < 0 && b >= a) { /* every candidate NaN: max returns NaN at index 1 */
      best = a;
      bestC = firstC;
      best_ni = first_ni;
      best_med = first_med;
    }
    o[0] = (double)(i + 1);
    if (best < 0) {
      o[1] = o[2] = o[3] = o[4] = o[5] = o[6] = o[7] = 0.0;
      continue;
    }
    int64_t gid = (int64_t)id_all[best - 1];
    o[1] = euler[3 * (gid - 1) + 0];
    o[2] = euler[3 * (gid - 1) + 1];
    o[3] = euler[3 * (gid - 1) + 2];
    o[4] = (double)gid;
    o[5] = bestC;
    o[6] = best_ni;
    o[7] = best_med;
  }
}

/* Elementwise evaluation of the project math functions (for tests):
 * fn 0 sin, 1 cos, 2 acos, 3 atan, 4 round. */
void oracle_math_eval(int fn, const double *x, double *y, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    switch (fn) {
      case 0: y[i] = grs_sin(x[i]); break;
      case 1: y[i] = grs_cos(x[i]); break;
      case 2: y[i] = grs_acos(x[i]); break;
      case 3: y[i] = grs_atan(x[i]); break;
      case 7: y[i] = grs_asin(x[i]); break;
      case 8: y[i] = grs_filt_acos(x[i]); break;
      case 9: y[i] = grs_filt_atan2(grs_sin(x[i]), grs_cos(x[i])); break;
      default: y[i] = grs_round(x[i]); break;
    }
  }
}
