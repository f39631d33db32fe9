/*
 * grs_fit.hip -- spot-centroid residual of the geometry / orientation fits and
 * its hand-written adjoint (reverse mode).
 *
 * Residual of one matched spot (pixels, unrounded):
 *   r = (-dety22/py + dety0 - y_meas, -detz22/pz + detz0 - z_meas)
 * with dety22/detz22 from the forward_comp geometry (forward_comp.m:15-55, the
 * convention of dis_calc_spotspair.m:35-71 and fit_grain_COM.m:49-83).
 * Parameters: the 7 geometry values [Lsam2sou Lsam2det dety00 detz00 tilt_x
 * tilt_y tilt_z] (geo_fitting.m:8-17) and, per grain, the Rodrigues vector
 * (U = quaternion2U(rod2quat(r))) and the centre of mass in mm
 * (fit_grain_COM.m:116-118).  The reference minimises the mean of rounded
 * distances with fminsearchbnd (geo_fitting.m:70-78); the smooth residual
 * here is what a gradient method needs.  cos(alpha) of forward_comp.m:42 is
 * used in its algebraic form gp1/|gp|, smooth on the rotation axis.
 *
 * One thread per spot.  The forward chain is recomputed in registers and
 * swept backwards once per seed; nothing is stored between the sweeps.
 * Outputs: the residual, and either the vector-Jacobian product with an
 * upstream gradient (13 values per spot, reduced per grain and globally in a
 * fixed order, so results are reproducible) or the 2 x 13 Jacobian rows.
 * The tilt derivatives use dR/dtilt matrices precomputed on the host.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

typedef struct grs_fitc {
  double S[9], B[9];
  double P0y, P0z, RAO, py, pz, dety0, detz0;
  double Lss, Lsd, dy00, dz00;
  double R[9];      /* RotDet at the current tilts */
  double dR[3][9];  /* dRotDet / dtilt_{x,y,z} (per degree) */
} grs_fitc;

/* forward intermediates of one spot */
struct SpotFwd {
  double s0, q0, q1, q2, q3, bh[3], U[9], uv[3], Gw[3], Gt[3], sp[3], SW[3], gp[3], L, ctr1, ctr2, ng, u[3], nc,
      cu1, cu2, L2, Glen, dot, cb, tt, st2, nv, ap, phix, phiy, sphy, diffvec, y22, z22, k[3], nk, K[3], tnum,
      tden, t, w[3], Y, Z;
};

__device__ __forceinline__ void spot_forward(const grs_fitc &f, const double *rod, const double *P, const double *h,
                                             double c, double s, SpotFwd &v) {
  double n2 = rod[0] * rod[0] + rod[1] * rod[1] + rod[2] * rod[2];
  v.s0 = 1.0 / __builtin_sqrt(1.0 + n2);
  v.q0 = v.s0;
  v.q1 = rod[0] * v.s0;
  v.q2 = rod[1] * v.s0;
  v.q3 = rod[2] * v.s0;
  double q0 = v.q0, q1 = v.q1, q2 = v.q2, q3 = v.q3;
  v.U[0] = 2 * (q0 * q0 + q1 * q1) - 1;
  v.U[1] = 2 * (q1 * q2 - q0 * q3);
  v.U[2] = 2 * (q1 * q3 + q0 * q2);
  v.U[3] = 2 * (q1 * q2 + q0 * q3);
  v.U[4] = 2 * (q0 * q0 + q2 * q2) - 1;
  v.U[5] = 2 * (q2 * q3 - q0 * q1);
  v.U[6] = 2 * (q1 * q3 - q0 * q2);
  v.U[7] = 2 * (q2 * q3 + q0 * q1);
  v.U[8] = 2 * (q0 * q0 + q3 * q3) - 1;
  for (int i = 0; i < 3; i++) v.bh[i] = f.B[3 * i] * h[0] + f.B[3 * i + 1] * h[1] + f.B[3 * i + 2] * h[2];
  for (int i = 0; i < 3; i++) v.uv[i] = v.U[3 * i] * v.bh[0] + v.U[3 * i + 1] * v.bh[1] + v.U[3 * i + 2] * v.bh[2];
  for (int i = 0; i < 3; i++) v.Gw[i] = f.S[3 * i] * v.uv[0] + f.S[3 * i + 1] * v.uv[1] + f.S[3 * i + 2] * v.uv[2];
  v.Gt[0] = c * v.Gw[0] - s * v.Gw[1];
  v.Gt[1] = s * v.Gw[0] + c * v.Gw[1];
  v.Gt[2] = v.Gw[2];
  double pos[3] = {P[0], P[1] - f.RAO, P[2]};
  for (int i = 0; i < 3; i++) v.sp[i] = f.S[3 * i] * pos[0] + f.S[3 * i + 1] * pos[1] + f.S[3 * i + 2] * pos[2];
  v.SW[0] = c * v.sp[0] - s * v.sp[1];
  v.SW[1] = s * v.sp[0] + c * v.sp[1];
  v.SW[2] = v.sp[2];
  v.gp[0] = f.Lss + v.SW[0];
  v.gp[1] = v.SW[1] - f.P0y + f.RAO;
  v.gp[2] = v.SW[2] - f.P0z;
  v.L = f.Lss + f.Lsd;
  v.ctr1 = v.gp[1] * v.L / v.gp[0];
  v.ctr2 = v.gp[2] * v.L / v.gp[0];
  v.ng = __builtin_sqrt(v.gp[0] * v.gp[0] + v.gp[1] * v.gp[1] + v.gp[2] * v.gp[2]);
  for (int i = 0; i < 3; i++) v.u[i] = v.gp[i] / v.ng;
  v.nc = __builtin_sqrt(v.L * v.L + v.ctr1 * v.ctr1 + v.ctr2 * v.ctr2);
  v.cu1 = v.ctr1 / v.nc;
  v.cu2 = v.ctr2 / v.nc;
  v.L2 = (f.Lsd - v.SW[0]) * v.ng / v.gp[0];
  v.Glen = __builtin_sqrt(v.Gt[0] * v.Gt[0] + v.Gt[1] * v.Gt[1] + v.Gt[2] * v.Gt[2]);
  v.dot = v.u[0] * v.Gt[0] + v.u[1] * v.Gt[1] + v.u[2] * v.Gt[2];
  v.cb = v.dot / v.Glen;
  double beta = grs_acos(v.cb);
  v.tt = 2.0 * (beta - GRS_PI / 2.0);
  v.st2 = grs_sin(v.tt);
  v.nv = __builtin_sqrt(v.Gt[1] * v.Gt[1] + v.Gt[2] * v.Gt[2]);
  v.ap = (v.Gt[1] * v.cu1 + v.Gt[2] * v.cu2) / v.nv;
  v.phix = grs_acos(v.ap);
  v.phiy = v.phix - v.tt;
  v.sphy = grs_sin(v.phiy);
  v.diffvec = v.L2 * v.st2 / v.sphy;
  v.y22 = v.ctr1 + v.diffvec * v.Gt[1] / v.nv;
  v.z22 = v.ctr2 + v.diffvec * v.Gt[2] / v.nv;
  v.k[0] = f.Lsd - v.SW[0];
  v.k[1] = v.y22 - v.SW[1];
  v.k[2] = v.z22 - v.SW[2];
  v.nk = __builtin_sqrt(v.k[0] * v.k[0] + v.k[1] * v.k[1] + v.k[2] * v.k[2]);
  for (int i = 0; i < 3; i++) v.K[i] = v.k[i] / v.nk;
  const double *R = f.R;
  v.tnum = R[0] * v.k[0] + R[3] * (f.dy00 - f.RAO - v.SW[1]) + R[6] * (f.dz00 - v.SW[2]);
  v.tden = R[0] * v.K[0] + R[3] * v.K[1] + R[6] * v.K[2];
  v.t = v.tnum / v.tden;
  v.w[0] = v.t * v.K[0] - v.k[0];
  v.w[1] = v.t * v.K[1] + (v.SW[1] - f.dy00 + f.RAO);
  v.w[2] = v.t * v.K[2] + (v.SW[2] - f.dz00);
  v.Y = R[1] * v.w[0] + R[4] * v.w[1] + R[7] * v.w[2];
  v.Z = R[2] * v.w[0] + R[5] * v.w[1] + R[8] * v.w[2];
}

/* Reverse sweep for seeds (gy, gz) on (ry, rz).  g[0..6] geometry, g[7..9]
 * Rodrigues, g[10..12] centre of mass. */
__device__ __forceinline__ void spot_reverse(const grs_fitc &f, const double *rod, const double *h, double c, double s,
                                             const SpotFwd &v, double gy, double gz, double *g) {
  const double *R = f.R;
  double Rb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double Lssb = 0, Lsdb = 0, dy00b = 0, dz00b = 0;
  double SWb[3] = {0, 0, 0};
  /* ry = -Y/py + dety0 - ym */
  double Yb = -gy / f.py, Zb = -gz / f.pz;
  double wb[3];
  for (int j = 0; j < 3; j++) {
    wb[j] = Yb * R[3 * j + 1] + Zb * R[3 * j + 2];
    Rb[3 * j + 1] += Yb * v.w[j];
    Rb[3 * j + 2] += Zb * v.w[j];
  }
  /* w = t*K + offsets */
  double tb = wb[0] * v.K[0] + wb[1] * v.K[1] + wb[2] * v.K[2];
  double Kb[3] = {wb[0] * v.t, wb[1] * v.t, wb[2] * v.t};
  double kb[3] = {-wb[0], 0.0, 0.0};
  SWb[1] += wb[1];
  dy00b -= wb[1];
  SWb[2] += wb[2];
  dz00b -= wb[2];
  /* t = tnum/tden */
  double tnumb = tb / v.tden, tdenb = -tb * v.t / v.tden;
  Rb[0] += tdenb * v.K[0];
  Rb[3] += tdenb * v.K[1];
  Rb[6] += tdenb * v.K[2];
  Kb[0] += tdenb * R[0];
  Kb[1] += tdenb * R[3];
  Kb[2] += tdenb * R[6];
  Rb[0] += tnumb * v.k[0];
  kb[0] += tnumb * R[0];
  Rb[3] += tnumb * (f.dy00 - f.RAO - v.SW[1]);
  dy00b += tnumb * R[3];
  SWb[1] -= tnumb * R[3];
  Rb[6] += tnumb * (f.dz00 - v.SW[2]);
  dz00b += tnumb * R[6];
  SWb[2] -= tnumb * R[6];
  /* K = k/nk */
  double KbK = Kb[0] * v.K[0] + Kb[1] * v.K[1] + Kb[2] * v.K[2];
  for (int i = 0; i < 3; i++) kb[i] += (Kb[i] - KbK * v.K[i]) / v.nk;
  Lsdb += kb[0];
  SWb[0] -= kb[0];
  double y22b = kb[1], z22b = kb[2];
  SWb[1] -= kb[1];
  SWb[2] -= kb[2];
  /* y22 = ctr1 + diffvec*Gt1/nv, z22 = ctr2 + diffvec*Gt2/nv */
  double ctr1b = y22b, ctr2b = z22b;
  double diffb = (y22b * v.Gt[1] + z22b * v.Gt[2]) / v.nv;
  double Gtb[3] = {0.0, y22b * v.diffvec / v.nv, z22b * v.diffvec / v.nv};
  double nvb = -(y22b * v.Gt[1] + z22b * v.Gt[2]) * v.diffvec / (v.nv * v.nv);
  /* diffvec = L2*st2/sin(phiy) */
  double L2b = diffb * v.st2 / v.sphy;
  double st2b = diffb * v.L2 / v.sphy;
  double sphyb = -diffb * v.diffvec / v.sphy;
  double phiyb = sphyb * grs_cos(v.phiy);
  double phixb = phiyb, ttb = -phiyb;
  /* phix = acos(ap), ap = (Gt1 cu1 + Gt2 cu2)/nv */
  double apb = -phixb / __builtin_sqrt(1.0 - v.ap * v.ap);
  Gtb[1] += apb * v.cu1 / v.nv;
  Gtb[2] += apb * v.cu2 / v.nv;
  double cu1b = apb * v.Gt[1] / v.nv, cu2b = apb * v.Gt[2] / v.nv;
  nvb -= apb * v.ap / v.nv;
  Gtb[1] += nvb * v.Gt[1] / v.nv;
  Gtb[2] += nvb * v.Gt[2] / v.nv;
  /* st2 = sin(tt), tt = 2*(acos(cb) - pi/2) */
  ttb += st2b * grs_cos(v.tt);
  double cbb = -2.0 * ttb / __builtin_sqrt(1.0 - v.cb * v.cb);
  /* cb = dot/Glen */
  double dotb = cbb / v.Glen, Glenb = -cbb * v.cb / v.Glen;
  double ub[3];
  for (int i = 0; i < 3; i++) {
    ub[i] = dotb * v.Gt[i];
    Gtb[i] += dotb * v.u[i] + Glenb * v.Gt[i] / v.Glen;
  }
  /* L2 = (Lsd - SW0)*ng/gp0 */
  double a = f.Lsd - v.SW[0];
  double ab = L2b * v.ng / v.gp[0];
  Lsdb += ab;
  SWb[0] -= ab;
  double ngb = L2b * a / v.gp[0];
  double gpb[3] = {-L2b * v.L2 / v.gp[0], 0.0, 0.0};
  /* cu = (L, ctr1, ctr2)/nc, only components 1 and 2 are used */
  double cdot = cu1b * v.cu1 + cu2b * v.cu2;
  double Lb = (0.0 - cdot * (v.L / v.nc)) / v.nc;
  ctr1b += (cu1b - cdot * v.cu1) / v.nc;
  ctr2b += (cu2b - cdot * v.cu2) / v.nc;
  /* u = gp/ng (and ng in L2) */
  double udot = ub[0] * v.u[0] + ub[1] * v.u[1] + ub[2] * v.u[2];
  for (int i = 0; i < 3; i++) gpb[i] += (ub[i] - udot * v.u[i]) / v.ng + ngb * v.u[i];
  /* ctr1 = gp1*L/gp0, ctr2 = gp2*L/gp0 */
  gpb[1] += ctr1b * v.L / v.gp[0];
  gpb[2] += ctr2b * v.L / v.gp[0];
  Lb += (ctr1b * v.gp[1] + ctr2b * v.gp[2]) / v.gp[0];
  gpb[0] -= (ctr1b * v.ctr1 + ctr2b * v.ctr2) / v.gp[0];
  /* gp = (Lss + SW0, SW1 - P0y + RAO, SW2 - P0z), L = Lss + Lsd */
  Lssb += gpb[0] + Lb;
  Lsdb += Lb;
  SWb[0] += gpb[0];
  SWb[1] += gpb[1];
  SWb[2] += gpb[2];
  /* SW = Omega*(S*pos) */
  double spb[3] = {c * SWb[0] + s * SWb[1], -s * SWb[0] + c * SWb[1], SWb[2]};
  for (int j = 0; j < 3; j++) g[10 + j] = f.S[j] * spb[0] + f.S[3 + j] * spb[1] + f.S[6 + j] * spb[2];
  /* Gt = Omega*Gw, Gw = S*uv, uv = U*bh */
  double Gwb[3] = {c * Gtb[0] + s * Gtb[1], -s * Gtb[0] + c * Gtb[1], Gtb[2]};
  double uvb[3];
  for (int j = 0; j < 3; j++) uvb[j] = f.S[j] * Gwb[0] + f.S[3 + j] * Gwb[1] + f.S[6 + j] * Gwb[2];
  double Ub[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Ub[3 * i + j] = uvb[i] * v.bh[j];
  /* U(q) */
  double q0 = v.q0, q1 = v.q1, q2 = v.q2, q3 = v.q3;
  double q0b = 4 * q0 * (Ub[0] + Ub[4] + Ub[8]) + 2 * (-q3 * Ub[1] + q2 * Ub[2] + q3 * Ub[3] - q1 * Ub[5] -
                                                        q2 * Ub[6] + q1 * Ub[7]);
  double q1b = 4 * q1 * Ub[0] + 2 * (q2 * Ub[1] + q3 * Ub[2] + q2 * Ub[3] - q0 * Ub[5] + q3 * Ub[6] + q0 * Ub[7]);
  double q2b = 4 * q2 * Ub[4] + 2 * (q1 * Ub[1] + q0 * Ub[2] + q1 * Ub[3] + q3 * Ub[5] - q0 * Ub[6] + q3 * Ub[7]);
  double q3b = 4 * q3 * Ub[8] + 2 * (-q0 * Ub[1] + q1 * Ub[2] + q0 * Ub[3] + q2 * Ub[5] + q1 * Ub[6] + q2 * Ub[7]);
  /* q = (1, r)/sqrt(1+|r|^2) */
  double s0b = q0b + q1b * rod[0] + q2b * rod[1] + q3b * rod[2];
  double n2b = -0.5 * v.s0 * v.s0 * v.s0 * s0b;
  g[7] = q1b * v.s0 + 2 * rod[0] * n2b;
  g[8] = q2b * v.s0 + 2 * rod[1] * n2b;
  g[9] = q3b * v.s0 + 2 * rod[2] * n2b;
  g[0] = Lssb;
  g[1] = Lsdb;
  g[2] = dy00b;
  g[3] = dz00b;
  for (int a3 = 0; a3 < 3; a3++) {
    double acc = 0.0;
    for (int k = 0; k < 9; k++) acc += Rb[k] * f.dR[a3][k];
    g[4 + a3] = acc;
  }
  (void)h;
}

__global__ void k_spot_residual(grs_fitc f, int64_t n, const double *hkl, const double *omega, const double *meas,
                                const int32_t *grain, const double *gr6, const double *up, double *res,
                                double *contrib, double *jac) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *p = gr6 + 6 * (int64_t)grain[i];
  double om = omega[i] * (GRS_PI / 180.0);
  double c = grs_cos(om), s = grs_sin(om);
  SpotFwd v;
  spot_forward(f, p, p + 3, hkl + 3 * i, c, s, v);
  double ry = -v.Y / f.py + f.dety0 - meas[2 * i];
  double rz = -v.Z / f.pz + f.detz0 - meas[2 * i + 1];
  res[2 * i] = ry;
  res[2 * i + 1] = rz;
  if (contrib) {
    double gy = up ? up[2 * i] : ry, gz = up ? up[2 * i + 1] : rz;
    spot_reverse(f, p, hkl + 3 * i, c, s, v, gy, gz, contrib + 13 * i);
  }
  if (jac) {
    spot_reverse(f, p, hkl + 3 * i, c, s, v, 1.0, 0.0, jac + 26 * i);
    spot_reverse(f, p, hkl + 3 * i, c, s, v, 0.0, 1.0, jac + 26 * i + 13);
  }
}

/* per-grain sums of the 6 grain components (spots sorted by grain) */
__global__ void k_grain_reduce(const double *contrib, const int64_t *goff, int64_t G, double *out6) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G * 6) return;
  int64_t gidx = t / 6;
  int k = (int)(t - gidx * 6);
  double acc = 0.0;
  for (int64_t i = goff[gidx]; i < goff[gidx + 1]; i++) acc += contrib[13 * i + 7 + k];
  out6[t] = acc;
}

/* the 7 geometry components, fixed-order reduction in one block */
__global__ __launch_bounds__(256) void k_geo_reduce(const double *contrib, int64_t n, double *out7) {
  __shared__ double part[7][256];
  int tid = threadIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = tid; i < n; i += 256)
    for (int k = 0; k < 7; k++) acc[k] += contrib[13 * i + k];
  for (int k = 0; k < 7; k++) part[k][tid] = acc[k];
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w)
      for (int k = 0; k < 7; k++) part[k][tid] += part[k][tid + w];
    __syncthreads();
  }
  if (tid < 7) out7[tid] = part[tid][0];
}

/* Gauss-Newton normal equations of the block-arrow system, one block per grain
 * (spots sorted by grain).  With the extended rows E = [J(13) r] of a spot's
 * two residual components, scaled by sqrt(w), every output is
 * sum over the grain's spots of E_y[p] E_y[q] + E_z[p] E_z[q] for a column pair
 * (p, q): the grain block Akk (upper triangle, 21), the coupling Agk (7 x 6),
 * bk (6), and the grain's share of Agg (upper triangle, 28) and bg (7).
 * w = 1 / sqrt(|r|^2 + delta^2) and rho = sqrt(|r|^2 + delta^2) - delta
 * (Charbonnier) when delta > 0, else w = 1 and rho = |r|^2 / 2.  The NE_T
 * threads compute NE_T spots' rows into LDS, then thread o adds output o over
 * them in spot order, so the sums are reproducible. */
#define GRS_NE_T 128
#define GRS_NE_PAIRS 104
#define GRS_NE_ROW 30 /* sqrt(w) E_y (14), sqrt(w) E_z (14), rho, |r|^2 */

__device__ __forceinline__ void ne_pair(int o, int *p, int *q) {
  int k = 0;
  for (int a = 0; a < 6; a++)
    for (int b = a; b < 6; b++, k++)
      if (k == o) { *p = 7 + a; *q = 7 + b; return; }
  for (int a = 0; a < 7; a++)
    for (int b = 0; b < 6; b++, k++)
      if (k == o) { *p = a; *q = 7 + b; return; }
  for (int b = 0; b < 6; b++, k++)
    if (k == o) { *p = 7 + b; *q = 13; return; }
  for (int a = 0; a < 7; a++)
    for (int b = a; b < 7; b++, k++)
      if (k == o) { *p = a; *q = b; return; }
  for (int a = 0; a < 7; a++, k++)
    if (k == o) { *p = a; *q = 13; return; }
  *p = 0;
  *q = 0;
}

__global__ __launch_bounds__(GRS_NE_T) void k_spot_normal(grs_fitc f, const double *hkl, const double *omega,
                                                          const double *meas, const int64_t *goff, const double *gr6,
                                                          double delta, double *pergrain, double *partial) {
  __shared__ double rows[GRS_NE_T][GRS_NE_ROW];
  const int64_t g = blockIdx.x;
  const int t = threadIdx.x;
  int p = 0, q = 0;
  if (t < GRS_NE_PAIRS) ne_pair(t, &p, &q);
  const double *par = gr6 + 6 * g;
  double acc = 0.0;
  for (int64_t base = goff[g]; base < goff[g + 1]; base += GRS_NE_T) {
    const int64_t i = base + t;
    const int cnt = goff[g + 1] - base < GRS_NE_T ? (int)(goff[g + 1] - base) : GRS_NE_T;
    if (i < goff[g + 1]) {
      const double om = omega[i] * (GRS_PI / 180.0);
      const double c = grs_cos(om), s = grs_sin(om);
      SpotFwd v;
      spot_forward(f, par, par + 3, hkl + 3 * i, c, s, v);
      const double ry = -v.Y / f.py + f.dety0 - meas[2 * i];
      const double rz = -v.Z / f.pz + f.detz0 - meas[2 * i + 1];
      double jy[13], jz[13];
      spot_reverse(f, par, hkl + 3 * i, c, s, v, 1.0, 0.0, jy);
      spot_reverse(f, par, hkl + 3 * i, c, s, v, 0.0, 1.0, jz);
      const double r2 = ry * ry + rz * rz;
      double sw = 1.0, rho = 0.5 * r2;
      if (delta > 0.0) {
        const double h = __builtin_sqrt(r2 + delta * delta);
        sw = 1.0 / __builtin_sqrt(h);
        rho = h - delta;
      }
      double *row = rows[t];
      for (int k = 0; k < 13; k++) {
        row[k] = sw * jy[k];
        row[14 + k] = sw * jz[k];
      }
      row[13] = sw * ry;
      row[27] = sw * rz;
      row[28] = rho;
      row[29] = r2;
    }
    __syncthreads();
    if (t < GRS_NE_PAIRS) {
      for (int j = 0; j < cnt; j++) acc += rows[j][p] * rows[j][q] + rows[j][14 + p] * rows[j][14 + q];
    } else if (t < GRS_NE_PAIRS + 2) {
      const int col = t == GRS_NE_PAIRS ? 28 : 29;
      for (int j = 0; j < cnt; j++) acc += rows[j][col];
    }
    __syncthreads();
  }
  /* outputs 0..68: Akk, Agk, bk of the grain; 69..103 and the two sums: its share of Agg, bg, rho, |r|^2 */
  if (t < 69)
    pergrain[69 * g + t] = acc;
  else if (t < GRS_NE_PAIRS + 2)
    partial[37 * g + (t - 69)] = acc;
}

/* the grains' shares of Agg, bg and the two sums added in grain order */
__global__ void k_normal_sum(const double *partial, int64_t G, double *out37) {
  const int j = threadIdx.x;
  if (j >= 37) return;
  double acc = 0.0;
  for (int64_t g = 0; g < G; g++) acc += partial[37 * g + j];
  out37[j] = acc;
}

hipError_t grs_launch_spot_normal(const void *fitc, int64_t G, const double *hkl, const double *omega,
                                  const double *meas, const int64_t *goff, const double *gr6, double delta,
                                  double *pergrain, double *partial, double *out37, hipStream_t st) {
  if (G <= 0) return hipSuccess;
  const grs_fitc &f = *static_cast<const grs_fitc *>(fitc);
  hipLaunchKernelGGL(k_spot_normal, dim3((unsigned)G), dim3(GRS_NE_T), 0, st, f, hkl, omega, meas, goff, gr6, delta,
                     pergrain, partial);
  hipLaunchKernelGGL(k_normal_sum, dim3(1), dim3(64), 0, st, partial, G, out37);
  return hipGetLastError();
}

hipError_t grs_launch_spot_residual(const void *fitc, int64_t n, const double *hkl, const double *omega,
                                    const double *meas, const int32_t *grain, const double *gr6, const double *up,
                                    double *res, double *contrib, double *jac, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const grs_fitc &f = *static_cast<const grs_fitc *>(fitc);
  hipLaunchKernelGGL(k_spot_residual, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, f, n, hkl, omega, meas,
                     grain, gr6, up, res, contrib, jac);
  return hipGetLastError();
}

hipError_t grs_launch_grad_reduce(const double *contrib, int64_t n, const int64_t *goff, int64_t G, double *out7,
                                  double *out6, hipStream_t st) {
  hipLaunchKernelGGL(k_geo_reduce, dim3(1), dim3(256), 0, st, contrib, n, out7);
  if (G > 0)
    hipLaunchKernelGGL(k_grain_reduce, dim3((unsigned)((G * 6 + 127) / 128)), dim3(128), 0, st, contrib, goff, G,
                       out6);
  return hipGetLastError();
}

size_t grs_fitc_size(void) { return sizeof(grs_fitc); }
