/*
 * grs_orient.h -- the orientation bookkeeping after an orientation fit.
 *
 * fit_OR_constrain_with_local_comp_cuda.m:133-148 turns the fitted vector
 * part x = (q1, q2, q3) into the reported Euler angles, Rodrigues vector,
 * quaternion and the U used to verify the fit:
 *   [yaw,pitch,roll] = rod2angle(quat2rod([sqrt(1-x1^2-x2^2-x3^2) x]),'ZXZ')
 *   in degrees, each +360 when negative; rod = angle2rod(angles,'ZXZ');
 *   quat = rod2quat(rod); U = quaternion2U(quat).
 * The steps follow quat2rod.m:4-14 (with quatnormalize.m, quatmod.m,
 * quatnorm.m), rod2angle.m ('zxz' case and twoaxisrot), angle2rod.m ('zxz'
 * case and the final division), rod2quat.m and quaternion2U.m, with each
 * statement evaluated in MATLAB's left-to-right order and the engine's FP64
 * functions (sum(q.^2) and norm() as sequential sums of squares).
 * rod2angle's gimbal-lock branch (it calls dcm2angle, which is not part of
 * the reference) and arguments for which MATLAB would return complex values
 * make the conversion fail: the function returns 0 and the outputs are NaN.
 */
#ifndef GRS_ORIENT_H
#define GRS_ORIENT_H

#include "grs_math.h"

/* angle2rod(a1*pi/180, a2*pi/180, a3*pi/180, 'ZXZ') (angle2rod.m 'zxz' case
 * and :120-126) for angles in degrees; 0 when acos((tr-1)/2) would be
 * complex. */
GRS_HD int grs_angle2rod_zxz_deg(const double *ang, double *rod) {
  GRS_NO_CONTRACT
  const double b1 = (ang[0] * GRS_PI) / 180.0, b2 = (ang[1] * GRS_PI) / 180.0, b3 = (ang[2] * GRS_PI) / 180.0;
  const double cps = grs_cos(b1), sps = grs_sin(b1);
  const double cth = grs_cos(b2), sth = grs_sin(b2);
  const double cph = grs_cos(b3), sph = grs_sin(b3);
  const double tr = (((cph * cps - (sph * cth) * sps) - sph * sps) + (cph * cth) * cps) + cth;
  const double e31 = sth * sps - sph * sth;
  const double e23 = cph * sth + sth * cps;
  const double e12 = ((cph * sps + (sph * cth) * cps) + sph * cps) + (cph * cth) * sps;
  const double c = (tr - 1.0) / 2.0;
  rod[0] = rod[1] = rod[2] = 0.0;
  if (!(c >= -1.0 && c <= 1.0)) return 0;
  const double th = grs_acos(c);
  if (th != 0.0) {
    const double d = 2.0 * grs_sin(th);
    const double sy = e31 / d, sx = e23 / d, sz = e12 / d;
    const double t = grs_tan(th / 2.0);
    rod[0] = t * sx;
    rod[1] = t * sy;
    rod[2] = t * sz;
  }
  return 1;
}

/* rod2quat.m: the zero quaternion for a zero vector. */
GRS_HD void grs_rod2quat(const double *rod, double *qo) {
  GRS_NO_CONTRACT
  qo[0] = qo[1] = qo[2] = qo[3] = 0.0;
  const double nr = grs_sqrt((rod[0] * rod[0] + rod[1] * rod[1]) + rod[2] * rod[2]);
  if (nr != 0.0) {
    const double th2 = grs_atan(nr);
    const double s2 = grs_sin(th2);
    qo[0] = grs_cos(th2);
    for (int k = 0; k < 3; k++) qo[k + 1] = (rod[k] * s2) / nr;
  }
}

/* out10 = [yaw pitch roll (degrees), rod(3), quat(4)], U row-major 3x3. */
GRS_HD int grs_fit_orientation(const double *x, double *out10, double *U) {
  GRS_NO_CONTRACT
  const double nan = __builtin_nan("");
  for (int k = 0; k < 10; k++) out10[k] = nan;
  for (int k = 0; k < 9; k++) U[k] = nan;
  /* sqrt(1-x(1)^2-x(2)^2-x(3)^2) */
  const double r0 = ((1.0 - x[0] * x[0]) - x[1] * x[1]) - x[2] * x[2];
  if (!(r0 >= 0.0)) return 0;
  const double q[4] = {grs_sqrt(r0), x[0], x[1], x[2]};
  /* quat2rod: quatnormalize, then R = q(2:4)/cos(acos(q(1))) */
  const double qm = grs_sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  double qn[4];
  for (int k = 0; k < 4; k++) qn[k] = q[k] / qm;
  double R[3] = {0.0, 0.0, 0.0};
  const double thalf = grs_acos(qn[0]);
  if (thalf != thalf) return 0;
  if (thalf != 0.0) {
    const double ct = grs_cos(thalf);
    for (int k = 0; k < 3; k++) R[k] = qn[k + 1] / ct;
  }
  /* rod2angle(R, 'ZXZ') */
  double r1 = 0.0, r2 = 0.0, r3 = 0.0;
  const double n = grs_sqrt((R[0] * R[0] + R[1] * R[1]) + R[2] * R[2]);
  if (n != 0.0) {
    const double th = 2.0 * grs_atan(n);
    const double sx = R[0] / n, sy = R[1] / n, sz = R[2] / n;
    const double cth = grs_cos(th), sth = grs_sin(th);
    const double a11 = (sx * sz) * (1.0 - cth) + sy * sth;
    const double a12 = sx * sth - (sy * sz) * (1.0 - cth);
    double a21 = sz * sz + (1.0 - sz * sz) * cth;
    const double a31 = (sx * sz) * (1.0 - cth) - sy * sth;
    const double a32 = (sy * sz) * (1.0 - cth) + sx * sth;
    if ((a11 == 0.0 && a12 == 0.0) || (a31 == 0.0 && a32 == 0.0)) return 0;
    r1 = grs_atan2(a11, a12);
    if (a21 < -1.0) a21 = -1.0;
    if (a21 > 1.0) a21 = 1.0;
    r2 = grs_acos(a21);
    r3 = grs_atan2(a31, a32);
  }
  /* degrees, shifted into [0, 360) */
  double ang[3] = {(r1 * 180.0) / GRS_PI, (r2 * 180.0) / GRS_PI, (r3 * 180.0) / GRS_PI};
  for (int k = 0; k < 3; k++)
    if (ang[k] < 0.0) ang[k] = ang[k] + 360.0;
  double rod[3], qo[4];
  if (!grs_angle2rod_zxz_deg(ang, rod)) return 0;
  grs_rod2quat(rod, qo);
  /* quaternion2U */
  const double q0 = qo[0], q1 = qo[1], q2 = qo[2], q3 = qo[3];
  U[0] = 2.0 * (q0 * q0 + q1 * q1) - 1.0;
  U[1] = 2.0 * (q1 * q2 - q0 * q3);
  U[2] = 2.0 * (q1 * q3 + q0 * q2);
  U[3] = 2.0 * (q1 * q2 + q0 * q3);
  U[4] = 2.0 * (q0 * q0 + q2 * q2) - 1.0;
  U[5] = 2.0 * (q2 * q3 - q0 * q1);
  U[6] = 2.0 * (q1 * q3 - q0 * q2);
  U[7] = 2.0 * (q2 * q3 + q0 * q1);
  U[8] = 2.0 * (q0 * q0 + q3 * q3) - 1.0;
  for (int k = 0; k < 3; k++) out10[k] = ang[k];
  for (int k = 0; k < 3; k++) out10[3 + k] = rod[k];
  for (int k = 0; k < 4; k++) out10[6 + k] = qo[k];
  return 1;
}

/* euler2u.m:5-16 with the angles in radians, U row-major. */
GRS_HD void grs_euler2u_rad(double phi1, double PHI, double phi2, double *U) {
  GRS_NO_CONTRACT
  const double c1 = grs_cos(phi1), s1 = grs_sin(phi1), c2 = grs_cos(phi2), s2 = grs_sin(phi2);
  const double cP = grs_cos(PHI), sP = grs_sin(PHI);
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

/* u2euler_corr.m:1-52 for U given column-major (as reshape(row,3,3) of a
 * HittedSpots row): the ZXZ angles in degrees of the branch whose euler2u is
 * closest to U in the sum of absolute differences (first minimum). */
GRS_HD void grs_u2euler_corr(const double *Ucm, double *out3) {
  GRS_NO_CONTRACT
#define GRS_UCM(r, c) Ucm[((c) - 1) * 3 + ((r) - 1)]
  const double ph = grs_acos(GRS_UCM(3, 3));
  if (ph < 0.0001) {
    const double p2 = grs_atan(GRS_UCM(2, 1) / GRS_UCM(1, 1)) / 2.0;
    out3[0] = (p2 * 180.0) / GRS_PI;
    out3[1] = (ph * 180.0) / GRS_PI;
    out3[2] = (p2 * 180.0) / GRS_PI;
    return;
  }
  double a1 = grs_atan2(-GRS_UCM(1, 3), GRS_UCM(2, 3)), a2;
  double b1 = grs_atan2(GRS_UCM(3, 1), GRS_UCM(3, 2)), b2;
  if (a1 < 0.0) {
    a1 = a1 + 2.0 * GRS_PI;
    a2 = a1 - GRS_PI;
  } else {
    a2 = a1 + GRS_PI;
  }
  if (b1 < 0.0) {
    b1 = b1 + 2.0 * GRS_PI;
    b2 = b1 - GRS_PI;
  } else {
    b2 = b1 + GRS_PI;
  }
  const double p1s[4] = {a1, a2, a1, a2}, p2s[4] = {b1, b1, b2, b2};
  double best = 0.0;
  int bi = 0;
  for (int k = 0; k < 4; k++) {
    double Uk[9];
    grs_euler2u_rad(p1s[k], ph, p2s[k], Uk);
    double tot = 0.0;
    for (int c = 1; c <= 3; c++) {
      double col = 0.0;
      for (int r = 1; r <= 3; r++) col = col + grs_fabs(Uk[(r - 1) * 3 + (c - 1)] - GRS_UCM(r, c));
      tot = tot + col;
    }
    if (k == 0 || tot < best) {
      best = tot;
      bi = k;
    }
  }
  out3[0] = (p1s[bi] * 180.0) / GRS_PI;
  out3[1] = (ph * 180.0) / GRS_PI;
  out3[2] = (p2s[bi] * 180.0) / GRS_PI;
#undef GRS_UCM
}

#endif /* GRS_ORIENT_H */
