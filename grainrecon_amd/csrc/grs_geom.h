/*
 * grs_geom.h -- the FP64 forward-completeness geometry of GrainRecon's
 * forward_comp.m, split into the pieces the gfx950 kernels schedule
 * separately:
 *
 *   grs_U_from_q / grs_U_from_euler_deg   orientation -> U
 *       (ForwardCalc_OR_space_comp.m:27-49, quaternion2U.m:3-28, euler2u.m:5-16)
 *   grs_SUB                               S*U*B        (forward_comp.m:16)
 *   grs_voxel_pos                         voxel index -> lab mm
 *       (ind2pos.m:6-10, grow_indexed_region_parallel.m:70-74)
 *   grs_vw_terms                          per (voxel, omega) terms shared by
 *       every orientation and reflection  (forward_comp.m:19-26, 42, 51, 54)
 *   grs_event_beta / grs_event_finish     per (omega, hkl) event
 *       (forward_comp.m:28-65)
 *
 * Every operation is written in the order the MATLAB statement evaluates it,
 * with no fused multiply-add (the translation units are compiled with
 * -ffp-contract=off) and with the transcendentals of grs_math.h, so the
 * device and the host produce the same doubles.  Matrix products are
 * sequential sums in index order; the structural zeros and the 1 of the
 * omega rotation are not multiplied out.
 */
#ifndef GRS_GEOM_H
#define GRS_GEOM_H

#include <stdint.h>

#include "grs_math.h"

/* Geometry constants in the form the kernels consume. */
typedef struct grs_dgeo {
  double S[9], B[9], R[9]; /* row-major */
  double Lss, Lsd, L, P0y, P0z, RAO, dety00, detz00;
  double py, pz, dety0, detz0, detysize, detzsize;
  double bsy0, bsy1, bsz0, bsz1;
  double beta_max, lambda_min, lambda_max, pio2;
  double thr;        /* minEucDis / mean([py pz]) */
  double cos_beta_max; /* cos(beta_max), used only by the conservative filters */
  double ipy, ipz;     /* 1/py, 1/pz, used only by the conservative filters */
  /* omega solve of the candidate filter (used only when om_solve != 0):
   * rotation angles om0 + i*dom (radians, ascending, span <= 2 pi + 2 dom),
   * detector corners (x, y, z) in the SamposW frame, 2 pixels outside the
   * pixel area, and the largest Bragg angle of the window */
  double om0, dom, om_last;
  double corner[12];
  double theta_win;
  int32_t om_solve;
  int32_t pad_;
  uint32_t hit_key;  /* largest key k with (double)as_float(k) <= thr */
  int32_t has_hit;   /* 0 when no non-negative distance can be a hit */
  int32_t H, nrot;
  int64_t nz, ny;    /* distance stack rows (detz) and columns (dety) */
} grs_dgeo;

/* Per (voxel, omega) terms.  ov0 = SW0 - Lsd equals -k0 exactly. */
typedef struct grs_vw {
  double c, s;          /* cos(omega), sin(omega) */
  double u0, u1, u2;    /* grainpos / norm(grainpos) */
  double ctr1, ctr2;    /* center(2), center(3) */
  double cu1, cu2;      /* center / norm(center), components 2 and 3 */
  double L2;            /* (Lsd - SW1) / cos(alpha) */
  double k0;            /* Lsd - SW(1) */
  double SW1, SW2;      /* SamposW(2), SamposW(3) */
  double tnum;          /* numerator of t */
  double ov1, ov2;      /* [.. SW(2)-dety00+RAO, SW(3)-detz00] */
} grs_vw;

#define GRS_VW_DOUBLES 16

GRS_HD void grs_matmul3(const double *A, const double *Bm, double *C) {
  GRS_NO_CONTRACT
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[i * 3 + j] = (A[i * 3 + 0] * Bm[0 * 3 + j] + A[i * 3 + 1] * Bm[1 * 3 + j]) + A[i * 3 + 2] * Bm[2 * 3 + j];
}

/* Quaternion row [q0 q1 q2 q3]: q0 is recomputed from the vector part as in
 * ForwardCalc_OR_space_comp.m:27-28.  Returns 0 (and leaves U) when
 * q1^2+q2^2+q3^2 > 1, a row the MATLAB loop skips. */
GRS_HD int grs_U_from_q(double q1, double q2, double q3, double *U) {
  GRS_NO_CONTRACT
  double ss = (q1 * q1 + q2 * q2) + q3 * q3;
  if (!(ss <= 1.0)) return 0;
  double q0 = grs_sqrt(1.0 - ss);
  U[0] = 2.0 * (q0 * q0 + q1 * q1) - 1.0;
  U[1] = 2.0 * (q1 * q2 - q0 * q3);
  U[2] = 2.0 * (q1 * q3 + q0 * q2);
  U[3] = 2.0 * (q1 * q2 + q0 * q3);
  U[4] = 2.0 * (q0 * q0 + q2 * q2) - 1.0;
  U[5] = 2.0 * (q2 * q3 - q0 * q1);
  U[6] = 2.0 * (q1 * q3 - q0 * q2);
  U[7] = 2.0 * (q2 * q3 + q0 * q1);
  U[8] = 2.0 * (q0 * q0 + q3 * q3) - 1.0;
  return 1;
}

/* euler2u.m:5-16 from the cosines and sines of phi1, PHI and phi2 */
GRS_HD void grs_U_from_trig(double c1, double s1, double cP, double sP, double c2, double s2, double *U) {
  GRS_NO_CONTRACT
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

/* euler2u.m:5-16 with the angles given in degrees and converted as
 * (deg*pi)/180 (grow_unindexed_compete_comp_cuda.m:84) */
GRS_HD double grs_deg2rad(double d) {
  GRS_NO_CONTRACT
  return (d * GRS_PI) / 180.0;
}

GRS_HD void grs_U_from_euler_deg(double p1d, double Pd, double p2d, double *U) {
  const double phi1 = grs_deg2rad(p1d), PHI = grs_deg2rad(Pd), phi2 = grs_deg2rad(p2d);
  grs_U_from_trig(grs_cos(phi1), grs_sin(phi1), grs_cos(PHI), grs_sin(PHI), grs_cos(phi2), grs_sin(phi2), U);
}

/* SUB = (S*U)*B; Gw(:,h) = SUB*hkl(:,h). */
GRS_HD void grs_SUB(const grs_dgeo *g, const double *U, double *SUB) {
  double SU[9];
  grs_matmul3(g->S, U, SU);
  grs_matmul3(SU, g->B, SUB);
}

GRS_HD void grs_Gw(const double *SUB, const double *hkl, double *gw) {
  GRS_NO_CONTRACT
  for (int r = 0; r < 3; r++) gw[r] = (SUB[r * 3 + 0] * hkl[0] + SUB[r * 3 + 1] * hkl[1]) + SUB[r * 3 + 2] * hkl[2];
}

/* p1 = [RecVolumePixel(:,1)' simap VoxSize Center(3) Dimension(3) ...] */
GRS_HD void grs_voxel_pos(double x, double y, double z, const double *p1, double *pos) {
  GRS_NO_CONTRACT
  double v[3] = {x, y, z};
  for (int k = 0; k < 3; k++) pos[k] = ((v[k] + p1[k] - 1.0) - p1[8 + k] / 2.0) * p1[4] + p1[5 + k];
  if (p1[3] == 1.0) {
    pos[0] = -pos[0];
    pos[1] = -pos[1];
  }
}

/* cos/sin of rot_angles(i)*pi/180 (forward_comp.m:19-20). */
GRS_HD void grs_omega_cs(double rot_deg, double *c, double *s) {
  GRS_NO_CONTRACT
  double omega = (rot_deg * GRS_PI) / 180.0;
  *c = grs_cos(omega);
  *s = grs_sin(omega);
}

/* forward_comp.m:17,22-26,42,51,54 for one voxel position (unshifted, mm) and
 * one omega given by its cos and sin. */
GRS_HD void grs_vw_terms(const grs_dgeo *g, const double *pos_in, double c, double s, grs_vw *v) {
  GRS_NO_CONTRACT
  double pos0 = pos_in[0], pos1 = pos_in[1] - g->RAO, pos2 = pos_in[2];
  const double *S = g->S;
  double OS[9];
  for (int j = 0; j < 3; j++) {
    OS[0 * 3 + j] = c * S[0 * 3 + j] - s * S[1 * 3 + j];
    OS[1 * 3 + j] = s * S[0 * 3 + j] + c * S[1 * 3 + j];
    OS[2 * 3 + j] = S[2 * 3 + j];
  }
  double SW0 = (OS[0] * pos0 + OS[1] * pos1) + OS[2] * pos2;
  double SW1 = (OS[3] * pos0 + OS[4] * pos1) + OS[5] * pos2;
  double SW2 = (OS[6] * pos0 + OS[7] * pos1) + OS[8] * pos2;
  double gp1 = g->Lss + SW0;
  double gp2 = (SW1 - g->P0y) + g->RAO;
  double gp3 = SW2 - g->P0z;
  double L = g->L;
  double ctr1 = (gp2 * L) / gp1, ctr2 = (gp3 * L) / gp1;
  double alpha = grs_atan(grs_sqrt(gp2 * gp2 + gp3 * gp3) / gp1);
  double ng = grs_sqrt((gp1 * gp1 + gp2 * gp2) + gp3 * gp3);
  double nc = grs_sqrt((L * L + ctr1 * ctr1) + ctr2 * ctr2);
  v->c = c;
  v->s = s;
  v->u0 = gp1 / ng;
  v->u1 = gp2 / ng;
  v->u2 = gp3 / ng;
  v->ctr1 = ctr1;
  v->ctr2 = ctr2;
  v->cu1 = ctr1 / nc;
  v->cu2 = ctr2 / nc;
  v->L2 = (g->Lsd - SW0) / grs_cos(alpha);
  v->k0 = g->Lsd - SW0;
  v->SW1 = SW1;
  v->SW2 = SW2;
  const double *R = g->R;
  v->tnum = (R[0] * (g->Lsd - SW0) + R[3] * ((g->dety00 - g->RAO) - SW1)) + R[6] * (g->detz00 - SW2);
  v->ov1 = (SW1 - g->dety00) + g->RAO;
  v->ov2 = SW2 - g->detz00;
}

/* Gt = Omega*Gw (forward_comp.m:28) */
GRS_HD void grs_Gt(double c, double s, double gw0, double gw1, double gw2, double *Gt) {
  GRS_NO_CONTRACT
  Gt[0] = c * gw0 - s * gw1;
  Gt[1] = s * gw0 + c * gw1;
  Gt[2] = gw2;
}

/* Glen and beta (forward_comp.m:30-31). */
GRS_HD double grs_event_beta(const grs_vw *v, const double *Gt, double *Glen_out) {
  GRS_NO_CONTRACT
  double Glen = grs_sqrt((Gt[0] * Gt[0] + Gt[1] * Gt[1]) + Gt[2] * Gt[2]);
  double cb = (v->u0 * (Gt[0] / Glen) + v->u1 * (Gt[1] / Glen)) + v->u2 * (Gt[2] / Glen);
  *Glen_out = Glen;
  return grs_acos(cb);
}

/* Remainder of the event given beta: wavelength window, detector spot,
 * tilted-detector intersection, pixel rounding and the bounds / beamstop
 * tests (forward_comp.m:33-65).  Returns 1 when the event is selected and
 * writes the 1-based pixel indices. */
GRS_HD int grs_event_finish_yz(const grs_dgeo *g, const grs_vw *v, const double *Gt, double Glen, double beta,
                               int32_t *dety_out, int32_t *detz_out, double *yz_out) {
  GRS_NO_CONTRACT
  const double pi = GRS_PI;
  if (!(beta > g->pio2 && beta < g->beta_max)) return 0;
  double theta = beta - g->pio2;
  double d = ((1.0 / Glen) * 2.0) * pi;
  double lambdahkl = (2.0 * d) * grs_sin(theta);
  if (!(lambdahkl > g->lambda_min && lambdahkl < g->lambda_max)) return 0;
  double sintth = grs_sin(2.0 * theta);
  double Gt2 = Gt[1], Gt3 = Gt[2];
  double nv = grs_sqrt(Gt2 * Gt2 + Gt3 * Gt3);
  double phix = grs_acos((Gt2 / nv) * v->cu1 + (Gt3 / nv) * v->cu2);
  double phiy = phix - 2.0 * theta;
  double diffvec = (v->L2 * sintth) / grs_sin(phiy);
  double konst = nv; /* sqrt(Gt(2,:).^2+Gt(3,:).^2), the same operations */
  double y22 = v->ctr1 + ((diffvec * Gt2) / konst);
  double z22 = v->ctr2 + ((diffvec * Gt3) / konst);
  double k0 = v->k0, k1 = y22 - v->SW1, k2 = z22 - v->SW2;
  double nk = grs_sqrt((k0 * k0 + k1 * k1) + k2 * k2);
  double K0 = k0 / nk, K1 = k1 / nk, K2 = k2 / nk;
  const double *R = g->R;
  double t = v->tnum / ((R[0] * K0 + R[3] * K1) + R[6] * K2);
  double w0 = t * K0 + (-k0), w1 = t * K1 + v->ov1, w2 = t * K2 + v->ov2;
  double Y22 = (R[1] * w0 + R[4] * w1) + R[7] * w2;
  double Z22 = (R[2] * w0 + R[5] * w1) + R[8] * w2;
  double ry = (-Y22) / g->py + g->dety0;
  double rz = (-Z22) / g->pz + g->detz0;
  double dety = grs_round(ry);
  double detz = grs_round(rz);
  if (!(dety >= 1.0 && dety <= g->detysize && detz >= 1.0 && detz <= g->detzsize)) return 0;
  if (dety >= g->bsy0 && dety <= g->bsy1 && detz >= g->bsz0 && detz <= g->bsz1) return 0;
  *dety_out = (int32_t)dety;
  *detz_out = (int32_t)detz;
  if (yz_out) {
    yz_out[0] = ry;
    yz_out[1] = rz;
  }
  return 1;
}

GRS_HD int grs_event_finish(const grs_dgeo *g, const grs_vw *v, const double *Gt, double Glen, double beta,
                            int32_t *dety_out, int32_t *detz_out) {
  return grs_event_finish_yz(g, v, Gt, Glen, beta, dety_out, detz_out, (double *)0);
}

/* The spot pixel of an event without the selection tests
 * (fit_grain_COM.m:60-83, the same statements as forward_comp.m:33-59):
 * the unrounded pixel coordinates -dety22/py + dety0 and -detz22/pz + detz0. */
GRS_HD void grs_event_pixel(const grs_dgeo *g, const grs_vw *v, const double *Gt, double Glen, double beta,
                            double *ry, double *rz) {
  GRS_NO_CONTRACT
  (void)Glen;
  double theta = beta - g->pio2;
  double sintth = grs_sin(2.0 * theta);
  double Gt2 = Gt[1], Gt3 = Gt[2];
  double nv = grs_sqrt(Gt2 * Gt2 + Gt3 * Gt3);
  double phix = grs_acos((Gt2 / nv) * v->cu1 + (Gt3 / nv) * v->cu2);
  double phiy = phix - 2.0 * theta;
  double diffvec = (v->L2 * sintth) / grs_sin(phiy);
  double y22 = v->ctr1 + ((diffvec * Gt2) / nv);
  double z22 = v->ctr2 + ((diffvec * Gt3) / nv);
  double k0 = v->k0, k1 = y22 - v->SW1, k2 = z22 - v->SW2;
  double nk = grs_sqrt((k0 * k0 + k1 * k1) + k2 * k2);
  double K0 = k0 / nk, K1 = k1 / nk, K2 = k2 / nk;
  const double *R = g->R;
  double t = v->tnum / ((R[0] * K0 + R[3] * K1) + R[6] * K2);
  double w0 = t * K0 + (-k0), w1 = t * K1 + v->ov1, w2 = t * K2 + v->ov2;
  double Y22 = (R[1] * w0 + R[4] * w1) + R[7] * w2;
  double Z22 = (R[2] * w0 + R[5] * w1) + R[8] * w2;
  *ry = (-Y22) / g->py + g->dety0;
  *rz = (-Z22) / g->pz + g->detz0;
}

/* MATLAB a:d:b, following MathWorks' published description of the colon
 * operator: n = round((b-a)/d) intervals, one fewer when a+n*d overshoots b
 * by more than tol = 2*eps*max(|a|,|b|); the end point snapped to b within
 * tol; values a+k*d from the left, last-k*d from the right, (a+last)/2 in the
 * middle of an even count.  Returns the count, -1 if above cap or not finite. */
GRS_HD int64_t grs_colon(double a, double d, double b, double *v, int64_t cap) {
  GRS_NO_CONTRACT
  if (!(a - a == 0.0) || !(d - d == 0.0) || !(b - b == 0.0)) return -1; /* NaN or infinite */
  if (d == 0.0 || (a < b && d < 0.0) || (b < a && d > 0.0)) return 0;
  const double fa = grs_fabs(a), fb = grs_fabs(b);
  const double tol = 2.0 * 2.220446049250313e-16 * (fa > fb ? fa : fb);
  const double sig = d > 0.0 ? 1.0 : -1.0;
  double n;
  if (a == __builtin_floor(a) && d == 1.0) {
    n = __builtin_floor(b) - a;
  } else if (a == __builtin_floor(a) && d == __builtin_floor(d)) {
    const double q = __builtin_floor(a / d);
    const double r = a - q * d;
    n = __builtin_floor((b - r) / d) - q;
  } else {
    n = __builtin_round((b - a) / d);
    if (sig * ((a + n * d) - b) > tol) n = n - 1.0;
  }
  if (!(n + 1.0 <= (double)cap)) return -1;
  double last = a + n * d;
  if (sig * (last - b) > -tol) last = b;
  const int64_t N = (int64_t)n;
  for (int64_t k = 0; k <= N / 2; k++) v[k] = a + (double)k * d;
  for (int64_t k = 0; k <= N / 2; k++) v[N - k] = last - (double)k * d;
  if (N % 2 == 0) v[N / 2] = (a + last) / 2.0;
  return N + 1;
}

#endif /* GRS_GEOM_H */
