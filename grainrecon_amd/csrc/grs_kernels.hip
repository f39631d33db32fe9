/*
 * grs_kernels.hip -- gfx950 kernels of the completeness-scoring engine.
 *
 * Hot path: for every evaluation (voxel, orientation) and every event
 * (omega, hkl) decide whether the reflection lands on the detector
 * (forward_comp.m:28-65), look the pixel up in the distance stack
 * (forward_comp.m:69-72) and reduce the integer counts and the median
 * distance (forward_comp.m:67, 86-90).
 *
 * k_fast  -- one 64-wide wavefront per evaluation.
 *   phase A: lanes walk the flattened (omega, hkl) event space 64 events at a
 *            time and apply a conservative Bragg/wavelength filter that only
 *            needs 4 multiplies and 3 FMAs per event; it rejects an event only
 *            when the exact test is proven to fail (margins 1e-12 on cos(beta),
 *            1e-9 relative on lambda, far above the FP64 rounding of the exact
 *            path).  Survivors are compacted into a 128-slot LDS ring with
 *            ballot + mbcnt.
 *   phase B1: whenever 64 survivors are queued the wave recomputes their
 *            spot positions through exact identities and refined hardware
 *            reciprocals (b1_classify) and decides every event whose pixel
 *            and window tests are unambiguous by a wide margin; the rest
 *            (about 1e-4 of them) are queued for
 *   phase B2: the exact FP64 event pipeline of forward_comp.m, 64 at a time.
 *   gather:  selected events are queued as stack indices; 64 at a time the
 *            wave gathers their distance keys (or hit bits), consumed one
 *            batch later, and counts hits with ballot/popcount.
 *   median : selected keys go to a per-wave list in HBM; the median is found
 *            by a 32-step radix bisection over the keys held in registers.
 *   Decisions are those of the exact pipeline, so the counts are bit-exact.
 *
 * k_plain -- the straightforward port: one thread per evaluation, the whole
 *   event pipeline for every event, per-thread key list and quickselect.
 *   It is the denominator of the batched-vs-plain speedup.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

#define GRS_WPB 8      /* largest number of waves per block of k_fast (the launch chooses 1, 2, 4 or 8) */
#ifndef GRS_FAST_MIN_WAVES
#define GRS_FAST_MIN_WAVES 4 /* resident waves per SIMD k_fast is compiled for (128 VGPRs) */
#endif
#define GRS_QCAP 128   /* LDS ring slots per wave */
#define GRS_GWT 5      /* doubles per hkl row in LDS: G (3), 1/|G|, |G|^2 */
#define GRS_RING_DOUBLES (3 * GRS_QCAP * 4 / 8) /* three u32 rings per wave */
#define GRS_RANGE_DOUBLES(H) (4 * (int64_t)(H) + 1) /* 8H + 2 u32: omega-solve ranges */
#define GRS_REG_KEYS 16 /* keys per lane kept in registers for the median */
#define GRS_BT 12       /* doubles per omega of the beam table: the omega_terms (10), cos, sin */

/* LDS doubles of one wave's region of k_fast: the G table, the rings, the
 * omega-solve ranges, S*pos of the voxel (4) and S*U*B (9) */
__host__ __device__ inline int64_t fast_wave_doubles(int32_t H) {
  return (int64_t)H * GRS_GWT + GRS_RING_DOUBLES + GRS_RANGE_DOUBLES(H) + 4 + 9;
}

/* LDS doubles of the one-voxel beam table behind the wave regions: the
 * table is kept when the block's dynamic LDS stays within 160 KB (one CU's
 * LDS; with 8 waves per block two blocks, 16 waves, share a CU) */
__host__ __device__ inline int64_t fast_tab_doubles(int32_t map, int32_t nrot, int32_t H, int wpb) {
  if (map != GRS_MAP_ONE_VOXEL || nrot <= 0) return 0;
  const int64_t t = (int64_t)GRS_BT * nrot;
  return ((int64_t)wpb * fast_wave_doubles(H) + t) * (int64_t)sizeof(double) <= 163840 ? t : 0;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lane_prefix(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  for (int off = 32; off >= 1; off >>= 1) {
    uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ float key_to_float(uint32_t k) { return __uint_as_float(k); }

/* ---------------------------------------------------------------------- */
/* small set-up kernels                                                    */

__global__ void k_omega_cs(const double *rot, int32_t nrot, double *cs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrot) return;
  double c, s;
  grs_omega_cs(rot[i], &c, &s);
  cs[2 * i] = c;
  cs[2 * i + 1] = s;
}

__global__ void k_voxel_pos(const double *rows, int64_t M, int32_t stride, const double *p1, double *pos) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double *r = rows + i * stride;
  grs_voxel_pos(r[0], r[1], r[2], p1, pos + 3 * i);
}

/* P = S * [pos0, pos1 - RotAxisOffset, pos2] of a voxel: SamposW at omega is
 * Omega * P (forward_comp.m:17,22 up to the order of the products, which
 * only the exact pipeline needs). */
__device__ __forceinline__ void voxel_P(const grs_dgeo *g, const double *pos, double *P) {
  const double x = pos[0], y = pos[1] - g->RAO, z = pos[2];
  for (int i = 0; i < 3; i++) P[i] = __builtin_fma(g->S[3 * i], x, __builtin_fma(g->S[3 * i + 1], y, g->S[3 * i + 2] * z));
}

/* SamposW and the unit beam direction u = grainpos / |grainpos| at
 * (cos omega, sin omega), to a few 1e-15 relative (filters only). */
__device__ __forceinline__ void beam_terms(const grs_dgeo *g, const double *P, double c, double s, double *sw,
                                           double *u) {
  sw[0] = c * P[0] - s * P[1];
  sw[1] = s * P[0] + c * P[1];
  sw[2] = P[2];
  const double gp0 = g->Lss + sw[0], gp1 = (sw[1] - g->P0y) + g->RAO, gp2 = sw[2] - g->P0z;
  const double n2 = __builtin_fma(gp0, gp0, __builtin_fma(gp1, gp1, gp2 * gp2));
  double r = __builtin_amdgcn_rsq(n2);
  r = r * __builtin_fma(-0.5 * n2, r * r, 1.5);
  u[0] = gp0 * r;
  u[1] = gp1 * r;
  u[2] = gp2 * r;
}

/* Per-voxel bounds of the omega-solve candidate filter (GRS_VB doubles per
 * voxel).  With u(omega) the unit beam direction at the voxel
 * (grainpos / norm(grainpos), forward_comp.m:26) and P = S*pos' (RotAxisOffset
 * taken off pos(2)), the first-order beam direction has
 * u1 ~ al sin(omega) + be cos(omega) + ga with al = P0/Lss, be = P1/Lss,
 * ga = (RotAxisOffset - P0y)/Lss; u0 ~ 1 and u2 varies only through the norm.
 *   vb[0] W  >= max |u0 - 1| + max |u1 - (al sin + be cos + ga)| + hypot(al, be)/2
 *   vb[1], vb[2]  min and max of u2 (widened by 1e-12)
 *   vb[3] K  >= sin^2(theta) = cos^2(beta) of every event that can reach the
 *            detector; 1 (no bound) when a detector corner is 90 degrees or
 *            more off the beam or when the window admits
 *            2 theta + angle(u, x) >= pi/2 - 1e-3
 *   vb[4..6] al, be, ga
 *   vb[7], vb[8]  Zt >= and Zb <= the detector rows term of every event
 *            (below); +-1e300 (no bound) where K is 1
 *   vb[9], vb[10], vb[11] sin(eta), cos(eta), tan(eta): eta >= asin |e . u|
 *   vb[12] Ya >= |Yb|, |Yt|, the detector columns term of every event
 *            (below); 1e300 (no bound) where K is 1
 *   vb[13] delta >= |u - (1, 0, 0)|
 * All extremes are over the rotation angles the scorer visits, so they hold
 * at every event.  Corners are the detector corners 2 pixels outside the
 * pixel area (g->corner, frame of SamposW).
 *
 * Spot geometry (forward_comp.m:40-50).  center - grainpos = L2 u
 * (L2 = (Lsd - SW1)/u0) and dety22/detz22 add diffvec e, e = (0, Gt2, Gt3)
 * / konst, so in the frame of grainpos the triangle (voxel, center, spot)
 * has the angle 2 theta at the voxel: phix = acos(e . u) and
 * phiy = phix - 2 theta are its other angles, and diffvec =
 * L2 sin(2 theta) / sin(phiy).  K_out (:49) takes the voxel as SamposW and
 * the spot as (Lsd, dety22, detz22), which differ from that frame by
 * o = (0, RotAxisOffset - P0y, -P0z):
 *   K_out ~ L2 u + diffvec e + o.
 * The tilted-detector spot (:51-55) lies on the ray SamposW + t K_out.
 * Cone: |L2 u + diffvec e| >= L2 sqrt(1 - a^2), a = e . u, |a| <=
 * |(u1, u2)|, so K_out is within eps = asin(|o| / (L2 sqrt(1 - a^2))) of
 * the triangle's side and 2 theta <= gamma + eps, gamma the largest angle
 * between u and the corner rays (the detector is convex, so the corners
 * bound it).  K = max sin^2((gamma + eps) / 2).
 * Rows: the z/x slope of K_out is u2/u0 + (diffvec/L2) e_z / u0 + o_z /
 * (L2 u0), and a spot on the detector has a slope between the smallest and
 * largest corner slope (all corners in front of the voxel).  So
 * (diffvec/L2) e_z lies in [s_min u0 - u2 - o_z/L2, s_max u0 - u2 - o_z/L2],
 * whose extremes over omega are Zb, Zt.
 * Columns: likewise with the y/x slopes, (diffvec/L2) e_y lies in
 * [t_min u0 - u1 - o_y/L2, t_max u0 - u1 - o_y/L2] (t the corner y/x slopes),
 * whose extremes over omega are Yb, Yt. */
__device__ __forceinline__ double wave_max_d(double v) {
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

__device__ __forceinline__ double wave_min_d(double v) {
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

/* The running extremes of the voxel bounds over the rotation angles. */
struct vb_acc {
  double d2max, kmax, e0, e1, u2lo, u2hi, ayz, zt, zb, yt, yb;
  bool front, rows;
};

__device__ __forceinline__ void vb_start(vb_acc &A) {
  A.d2max = 0.0, A.kmax = 0.0, A.e0 = 0.0, A.e1 = 0.0, A.u2lo = 2.0, A.u2hi = -2.0, A.ayz = 0.0;
  A.zt = -1e300, A.zb = 1e300, A.yt = -1e300, A.yb = 1e300;
  A.front = true, A.rows = true;
}

/* One rotation angle's terms of the voxel bounds (P = S*pos of the voxel,
 * al, be, ga, oy, oz, on as in the kernels below). */
__device__ __forceinline__ void vb_omega(const grs_dgeo *g, const double *P, double c, double s, double al, double be,
                                         double ga, double oy, double oz, double on, vb_acc &A) {
  double sw[3], u[3];
  beam_terms(g, P, c, s, sw, u);
  const double u0 = u[0], u1 = u[1], u2 = u[2];
  const double sw0 = sw[0], sw1 = sw[1], sw2 = sw[2];
  const double d2 = (1.0 - u0) * (1.0 - u0) + u1 * u1 + u2 * u2;
  A.d2max = d2 > A.d2max ? d2 : A.d2max;
  const double a0 = __builtin_fabs(u0 - 1.0), a1 = __builtin_fabs(u1 - ((al * s + be * c) + ga));
  A.e0 = a0 > A.e0 ? a0 : A.e0;
  A.e1 = a1 > A.e1 ? a1 : A.e1;
  A.u2lo = u2 < A.u2lo ? u2 : A.u2lo;
  A.u2hi = u2 > A.u2hi ? u2 : A.u2hi;
  const double ay = __builtin_sqrt(u1 * u1 + u2 * u2) * (1.0 + 1e-9) + 1e-12;
  A.ayz = ay > A.ayz ? ay : A.ayz;
  double cmin = 1.0, smin = 1e300, smax = -1e300, tmin = 1e300, tmax = -1e300;
  for (int j = 0; j < 4; j++) {
    const double dx = g->corner[3 * j] - sw0, dy = g->corner[3 * j + 1] - sw1, dz = g->corner[3 * j + 2] - sw2;
    const double cj = (u0 * dx + u1 * dy + u2 * dz) / __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    A.front = A.front && (cj > 1e-3);
    cmin = cj < cmin ? cj : cmin;
    A.rows = A.rows && (dx > 0.0);
    const double sl = dz / dx, ty = dy / dx;
    smin = sl < smin ? sl : smin;
    smax = sl > smax ? sl : smax;
    tmin = ty < tmin ? ty : tmin;
    tmax = ty > tmax ? ty : tmax;
  }
  const double L2 = (g->Lsd - sw0) / u0;
  A.front = A.front && (L2 > 0.0) && (ay < 1.0);
  /* sin and cos of gamma + eps, gamma = acos(cmin) */
  double se = on / (L2 * __builtin_sqrt(1.0 - (ay < 1.0 ? ay * ay : 0.0)));
  se = se * (1.0 + 1e-9) + 1e-12;
  A.front = A.front && (se < 1.0);
  se = se < 1.0 ? se : 1.0;
  const double ce = __builtin_sqrt(1.0 - se * se);
  const double cm = cmin < 1.0 ? cmin : 1.0;
  const double sg = __builtin_sqrt(1.0 - cm * cm);
  const double k = (1.0 - (cm * ce - sg * se)) * 0.5;
  A.kmax = k > A.kmax ? k : A.kmax;
  const double ozl = oz / L2;
  const double t1 = (smax * u0 - u2) - ozl, t0 = (smin * u0 - u2) - ozl;
  A.zt = t1 > A.zt ? t1 : A.zt;
  A.zb = t0 < A.zb ? t0 : A.zb;
  const double oyl = oy / L2;
  const double y1 = (tmax * u0 - u1) - oyl, y0 = (tmin * u0 - u1) - oyl;
  A.yt = y1 > A.yt ? y1 : A.yt;
  A.yb = y0 < A.yb ? y0 : A.yb;
}

/* a NaN anywhere turns the bound off: the comparisons above skip NaNs */
__device__ __forceinline__ bool vb_nan(const vb_acc &A) {
  return !(A.d2max == A.d2max && A.kmax == A.kmax && A.zt == A.zt && A.zb == A.zb && A.yt == A.yt && A.yb == A.yb);
}

/* The GRS_VB values of a voxel from the extremes over every rotation angle. */
__device__ __forceinline__ void vb_write(const grs_dgeo *g, double al, double be, double ga, const vb_acc &A,
                                         bool nan_any, double *o) {
  const double delta = __builtin_sqrt(A.d2max) * (1.0 + 1e-9) + 1e-12;
  const double alpha = 2.0 * grs_asin(delta * 0.5 < 1.0 ? delta * 0.5 : 1.0) + 1e-9;
  double K = 1.0;
  const bool ok = A.front && !nan_any && 2.0 * g->theta_win + alpha < g->pio2 - 1e-3;
  if (ok) {
    K = A.kmax * (1.0 + 1e-6) + 1e-9;
    K = K < 1.0 ? K : 1.0;
  }
  const bool use_rows = ok && A.rows && A.ayz < 1.0;
  o[0] = ((__builtin_sqrt(al * al + be * be) * 0.5 + A.e0) + A.e1) * (1.0 + 1e-9) + 1e-12;
  o[1] = A.u2lo - 1e-12;
  o[2] = A.u2hi + 1e-12;
  o[3] = K;
  o[4] = al;
  o[5] = be;
  o[6] = ga;
  o[7] = use_rows ? A.zt + (__builtin_fabs(A.zt) * 1e-9 + 1e-12) : 1e300;
  o[8] = use_rows ? A.zb - (__builtin_fabs(A.zb) * 1e-9 + 1e-12) : -1e300;
  const double sn = A.ayz < 1.0 ? A.ayz : 1.0;
  o[9] = sn;
  o[10] = __builtin_sqrt(1.0 - sn * sn);
  o[11] = o[10] > 0.0 ? sn / o[10] : 1e300;
  const double ya = __builtin_fabs(A.yt) > __builtin_fabs(A.yb) ? __builtin_fabs(A.yt) : __builtin_fabs(A.yb);
  o[12] = use_rows ? ya * (1.0 + 1e-9) + 1e-12 : 1e300;
  o[13] = delta;
}

/* One 256-thread block per voxel: thread t takes the rotation angles t,
 * t + 256, ... (one each for up to 256 angles, which keeps the latency of a
 * one-voxel call short), and the extremes are reduced across each wave and
 * then across the four waves; maxima, minima and flags are exact, so the
 * result does not depend on the partition. */
__global__ __launch_bounds__(256) void k_voxel_bounds(const grs_dgeo *g, const double *pos, int64_t V,
                                                      const double *cs, int32_t nrot, double *vb) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t v = blockIdx.x; /* one block per voxel, one omega per thread */
  double P[3];
  voxel_P(g, pos + 3 * v, P);
  const double al = P[0] / g->Lss, be = P[1] / g->Lss, ga = (g->RAO - g->P0y) / g->Lss;
  const double oy = g->RAO - g->P0y, oz = -g->P0z;
  const double on = __builtin_sqrt(oy * oy + oz * oz);
  vb_acc A;
  vb_start(A);
  for (int32_t om = tid; om < nrot; om += 256) vb_omega(g, P, cs[2 * om], cs[2 * om + 1], al, be, ga, oy, oz, on, A);
  const bool nan_seen = __ballot(vb_nan(A)) != 0;
  A.front = __ballot(!A.front) == 0;
  A.rows = __ballot(!A.rows) == 0;
  A.d2max = wave_max_d(A.d2max);
  A.kmax = wave_max_d(A.kmax);
  A.e0 = wave_max_d(A.e0);
  A.e1 = wave_max_d(A.e1);
  A.u2lo = wave_min_d(A.u2lo);
  A.u2hi = wave_max_d(A.u2hi);
  A.ayz = wave_max_d(A.ayz);
  A.zt = wave_max_d(A.zt);
  A.zb = wave_min_d(A.zb);
  A.yt = wave_max_d(A.yt);
  A.yb = wave_min_d(A.yb);
  /* the four waves' maxima, minima and flags (exact, so order-free) */
  __shared__ double red[4][14];
  if (lane == 0) {
    double *r = red[wv];
    r[0] = A.d2max, r[1] = A.kmax, r[2] = A.e0, r[3] = A.e1, r[4] = A.u2lo, r[5] = A.u2hi, r[6] = A.ayz;
    r[7] = A.zt, r[8] = A.zb, r[9] = A.front ? 1.0 : 0.0, r[10] = A.rows ? 1.0 : 0.0, r[11] = nan_seen ? 1.0 : 0.0;
    r[12] = A.yt, r[13] = A.yb;
  }
  __syncthreads();
  if (tid != 0) return;
  bool nan_any = false;
  for (int k = 0; k < 4; k++) {
    const double *r = red[k];
    A.d2max = r[0] > A.d2max ? r[0] : A.d2max;
    A.kmax = r[1] > A.kmax ? r[1] : A.kmax;
    A.e0 = r[2] > A.e0 ? r[2] : A.e0;
    A.e1 = r[3] > A.e1 ? r[3] : A.e1;
    A.u2lo = r[4] < A.u2lo ? r[4] : A.u2lo;
    A.u2hi = r[5] > A.u2hi ? r[5] : A.u2hi;
    A.ayz = r[6] > A.ayz ? r[6] : A.ayz;
    A.zt = r[7] > A.zt ? r[7] : A.zt;
    A.zb = r[8] < A.zb ? r[8] : A.zb;
    A.yt = r[12] > A.yt ? r[12] : A.yt;
    A.yb = r[13] < A.yb ? r[13] : A.yb;
    A.front = A.front && r[9] != 0.0;
    A.rows = A.rows && r[10] != 0.0;
    nan_any = nan_any || r[11] != 0.0;
  }
  vb_write(g, al, be, ga, A, nan_any, vb + GRS_VB * v);
}

/* The same bounds with one thread per voxel walking every rotation angle:
 * for thousands of voxels, where a block per voxel spends most of its time
 * in the reductions (3.2M voxels: about 34 ms as blocks).  Maxima, minima
 * and flags are exact, so both kernels write the same values. */
__global__ __launch_bounds__(256) void k_voxel_bounds_seq(const grs_dgeo *g, const double *pos, int64_t V,
                                                          const double *cs, int32_t nrot, double *vb) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  double P[3];
  voxel_P(g, pos + 3 * v, P);
  const double al = P[0] / g->Lss, be = P[1] / g->Lss, ga = (g->RAO - g->P0y) / g->Lss;
  const double oy = g->RAO - g->P0y, oz = -g->P0z;
  const double on = __builtin_sqrt(oy * oy + oz * oz);
  vb_acc A;
  vb_start(A);
  for (int32_t om = 0; om < nrot; om++) vb_omega(g, P, cs[2 * om], cs[2 * om + 1], al, be, ga, oy, oz, on, A);
  vb_write(g, al, be, ga, A, vb_nan(A), vb + GRS_VB * v);
}

/* Refined hardware reciprocal / reciprocal square root (one Newton step,
 * relative error below 5e-15): used only by the conservative filters, whose
 * margins are 1e-12 or wider. */
__device__ __forceinline__ double fast_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  return y * __builtin_fma(-x, y, 2.0);
}
__device__ __forceinline__ double fast_rsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  return y * __builtin_fma(-0.5 * x, y * y, 1.5);
}

/* Largest sin(theta) = -cos(beta) of an event that can reach the detector,
 * from the voxel bounds.  With F = diffvec / L2 = sin(2 theta) / sin(phiy)
 * and the unit vector e = (0, Gt2, Gt3) / konst, F e_z lies in [Zb, Zt]
 * (rows) and F e_y in [Yb, Yt] (columns), so F |e_y| <= Ya.
 *   Rows: for Gz > 0 (Gz < 0 likewise with -Zb) e_z >= Gz/|G| = gz, as
 *   konst <= |G|, so F <= T = Zt / gz.
 *   Columns: Gt1/|G| differs from cos(beta) = u . G/|G| by at most
 *   |u - (1, 0, 0)| <= delta, so |Gt1|/|G| <= s1 + delta for an event in the
 *   window (s1 >= -cos(beta), the caller's bound before this one), konst^2
 *   >= |G|^2 (1 - (s1 + delta)^2), e_z^2 <= gz^2 / (1 - (s1 + delta)^2) and
 *   |e_y| = sqrt(1 - e_z^2) gives F <= T = Ya / sqrt(1 - that bound).
 * With T the smaller of the two: phiy = pi/2 - asin(a) - 2 theta with
 * |asin a| <= eta gives F >= sin(2 theta) / cos(2 theta - eta) when
 * 2 theta >= eta, so sin(2 theta) (1 - T sin eta) <= T cos eta cos(2 theta),
 * i.e. tan(2 theta) <= T' = T cos(eta) / (1 - T sin(eta)), or 2 theta < eta.
 * Then sin(theta)^2 = T'^2 / (2 q (q + 1)), q = sqrt(1 + T'^2), with T'
 * raised to tan(eta).  Returns 1 (no bound) where neither bound applies. */
__device__ __forceinline__ double det_sin_theta(double gz, double s1, double zt, double zb, double ya, double dl,
                                                double sn, double cn, double te) {
  double T = 1e300;
  const double Z = gz > 0.0 ? zt : -zb;
  if (gz != 0.0 && Z > 0.0 && Z < 1e299) T = Z * fast_rcp(__builtin_fabs(gz));
  if (ya < 1e299) {
    const double gx = (s1 + 1e-12) + dl;
    if (gx < 0.99) {
      const double ez2 = (gz * gz) * fast_rcp(1.0 - gx * gx) * (1.0 + 1e-12);
      if (ez2 < 0.98) {
        const double Tc = ya * fast_rsq(1.0 - ez2);
        T = Tc < T ? Tc : T;
      }
    }
  }
  if (!(T < 1e299)) return 1.0;
  const double den = 1.0 - T * sn;
  if (!(den > 0.0) || !(cn > 0.0)) return 1.0;
  double Tp = T * cn * fast_rcp(den);
  Tp = Tp > te ? Tp : te;
  const double q2 = 1.0 + Tp * Tp;
  const double q = q2 * fast_rsq(q2);
  const double st = Tp * fast_rsq((2.0 * q) * (q + 1.0));
  return st * (1.0 + 1e-9) + 1e-12;
}

/* grs_filt_acos and grs_filt_atan2 (grs_math.h: the same polynomials and
 * reductions) with the square root and the quotient taken from the refined
 * hardware reciprocals, which adds about 1e-14 to their error bounds (5e-12
 * and 2e-12 rad); NaN outside [-1, 1] as there. */
__device__ __forceinline__ double acos_core(double x) {
  const double R[8] = GRS_FILT_ASIN_R;
  const double ax = __builtin_fabs(x);
  const bool small = ax <= 0.5;
  const double zb = (1.0 - ax) * 0.5;
  const double u = small ? x : (zb == 0.0 ? 0.0 : zb * fast_rsq(zb));
  const double z = small ? x * x : zb;
  double p = R[7];
  p = __builtin_fma(p, z, R[6]);
  p = __builtin_fma(p, z, R[5]);
  p = __builtin_fma(p, z, R[4]);
  p = __builtin_fma(p, z, R[3]);
  p = __builtin_fma(p, z, R[2]);
  p = __builtin_fma(p, z, R[1]);
  p = __builtin_fma(p, z, R[0]);
  const double as = __builtin_fma(u * z, p, u);
  if (small) return GRS_PIO2_HI - as;
  return x > 0.0 ? 2.0 * as : GRS_PI_HI - 2.0 * as;
}
__device__ __forceinline__ double atan2_core(double y, double x) {
  const double S[7] = GRS_FILT_ATAN_S;
  const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
  const double mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
  if (!(mx > 0.0)) return mx == 0.0 ? 0.0 : mx;
  const bool big = mn > GRS_TAN_PI_8 * mx;
  const double w = big ? (mn - mx) * fast_rcp(mn + mx) : mn * fast_rcp(mx);
  const double z = w * w;
  double p = S[6];
  p = __builtin_fma(p, z, S[5]);
  p = __builtin_fma(p, z, S[4]);
  p = __builtin_fma(p, z, S[3]);
  p = __builtin_fma(p, z, S[2]);
  p = __builtin_fma(p, z, S[1]);
  p = __builtin_fma(p, z, S[0]);
  double a = __builtin_fma(w * z, p, w);
  if (big) a = a + 0.5 * GRS_PIO2_HI;
  if (ay > ax) a = GRS_PIO2_HI - a;
  if (x < 0.0) a = GRS_PI_HI - a;
  return y < 0.0 ? -a : a;
}
/* the device math test's entries (fn 10, 11) */
__device__ __noinline__ double filt_acos(double x) { return acos_core(x); }
__device__ __noinline__ double filt_atan2(double y, double x) { return atan2_core(y, x); }

/* The omega solve's three angles in one out-of-line call: acos(a) (pi at
 * a <= -1), acos(b) (0 at b >= 1) and atan2(y, x), the two polynomial
 * chains of acos side by side.  Inlined, they raise the register pressure
 * of the whole kernel (spills in the candidate loop); three calls cost 2%
 * of the seeding launch more than one. */
struct arc_angles {
  double pa, pb, phi;
};
__device__ __noinline__ arc_angles arc_ends(double a, double b, double y, double x) {
  arc_angles r;
  r.pa = acos_core(a);
  r.pb = acos_core(b);
  r.phi = atan2_core(y, x);
  if (a <= -1.0) r.pa = GRS_PI;
  if (b >= 1.0) r.pb = 0.0;
  return r;
}

__global__ void k_sub_from_q(const grs_dgeo *g, const double *q, int64_t N, double *SUB, int32_t *valid) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double U[9];
  int ok = grs_U_from_q(q[4 * i + 1], q[4 * i + 2], q[4 * i + 3], U);
  valid[i] = ok;
  if (ok) grs_SUB(g, U, SUB + 9 * i);
  else
    for (int k = 0; k < 9; k++) SUB[9 * i + k] = 0.0;
}

__global__ void k_sub_from_U(const grs_dgeo *g, const double *U, int64_t N, double *SUB, int32_t *valid) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  grs_SUB(g, U + 9 * i, SUB + 9 * i);
  valid[i] = 1;
}

__global__ void k_sub_from_euler(const grs_dgeo *g, const double *eu, int64_t N, double *SUB, int32_t *valid) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double U[9];
  grs_U_from_euler_deg(eu[3 * i], eu[3 * i + 1], eu[3 * i + 2], U);
  grs_SUB(g, U, SUB + 9 * i);
  valid[i] = 1;
}

__global__ void k_math_eval(int fn, const double *x, double *y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = x[i], r;
  switch (fn) {
    case 0: r = grs_sin(v); break;
    case 1: r = grs_cos(v); break;
    case 2: r = grs_acos(v); break;
    case 3: r = grs_atan(v); break;
    case 4: r = grs_round(v); break;
    case 5: r = grs_sqrt(v); break;
    case 7: r = grs_asin(v); break;
    case 8: r = grs_filt_acos(v); break;
    case 9: r = grs_filt_atan2(grs_sin(v), grs_cos(v)); break;
    case 10: r = filt_acos(v); break;
    case 11: r = filt_atan2(grs_sin(v), grs_cos(v)); break;
    case 12: r = fast_rcp(v); break;
    case 13: r = fast_rsq(v); break;
    default: r = 1.0 / v; break;
  }
  y[i] = r;
}

/* Plain device copy (calibrates memory counters: reads and writes n words). */
__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

/* Flags keys that are not non-negative, non-NaN floats. */
__global__ void k_check_keys(const uint32_t *keys, int64_t n, uint32_t *bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint32_t b = 0;
  for (; i < n; i += stride) b |= (keys[i] > 0x7F800000u) ? 1u : 0u;
  if (b) atomicOr(bad, 1u);
}

/* ---------------------------------------------------------------------- */
/* mapping of evaluations                                                   */

__device__ __forceinline__ void eval_slots(const grs_eval_args &a, int64_t e, int64_t *vi, int64_t *ui) {
  if (a.map == GRS_MAP_ONE_VOXEL) {
    *vi = 0;
    *ui = e;
  } else if (a.map == GRS_MAP_ONE_U) {
    *vi = e;
    *ui = 0;
  } else if (a.map == GRS_MAP_GRID) {
    const int64_t v = e / a.n_u;
    *vi = v;
    *ui = (e - v * a.n_u) + v * a.u_stride;
  } else {
    *vi = a.eval_vox[e];
    *ui = a.eval_u[e];
  }
}

/* ---------------------------------------------------------------------- */
/* wave-level median over a key list in global memory                      */

/* number of keys < x among list[0..n) */
__device__ __forceinline__ int32_t wave_count_lt_global(const uint32_t *list, int32_t n, uint32_t x, int lane) {
  int32_t c = 0;
  for (int32_t b = 0; b < n; b += 64) {
    int32_t j = b + lane;
    bool p = j < n && list[j] < x;
    c += __popcll(__ballot(p));
  }
  return c;
}

__device__ uint32_t wave_select_global(const uint32_t *list, int32_t n, int32_t k, int lane) {
  uint32_t res = 0;
  for (int bit = 31; bit >= 0; --bit) {
    uint32_t cand = res | (1u << bit);
    if (wave_count_lt_global(list, n, cand, lane) <= k) res = cand;
  }
  return res;
}

/* Median of n <= 64*R keys held R per lane in registers: the k-th smallest
 * key by bisection on its bits (count of keys below a candidate by ballot). */
template <int R>
__device__ float wave_median_reg(const uint32_t *list, int32_t n, int lane) {
  const int32_t k1 = (n & 1) ? (n >> 1) : (n >> 1) - 1;
  uint32_t r[R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    int32_t idx = j * 64 + lane;
    r[j] = idx < n ? list[idx] : 0xFFFFFFFFu;
  }
  uint32_t res = 0;
  /* keys are non-negative floats: bit 31 is clear */
  for (int bit = 30; bit >= 0; --bit) {
    uint32_t cand = res | (1u << bit);
    int32_t c = 0;
#pragma unroll
    for (int j = 0; j < R; j++) c += __popcll(__ballot(r[j] < cand));
    if (c <= k1) res = cand;
  }
  const uint32_t v1 = res;
  if (n & 1) return key_to_float(v1);
  int32_t cle = 0;
  uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < R; j++) {
    cle += __popcll(__ballot(r[j] <= v1));
    if (r[j] > v1 && r[j] < mn) mn = r[j];
  }
  const uint32_t v2 = cle > k1 + 1 ? v1 : wave_min_u32(mn);
  float a = key_to_float(v1), b = key_to_float(v2);
  return a + (b - a) / 2.0f;
}

__device__ float wave_median(const uint32_t *list, int32_t n, int lane) {
  if (n <= 128) return wave_median_reg<2>(list, n, lane);
  if (n <= 256) return wave_median_reg<4>(list, n, lane);
  if (n <= 512) return wave_median_reg<8>(list, n, lane);
  if (n <= 64 * GRS_REG_KEYS) return wave_median_reg<GRS_REG_KEYS>(list, n, lane);
  const int32_t k1 = (n & 1) ? (n >> 1) : (n >> 1) - 1;
  uint32_t v1 = wave_select_global(list, n, k1, lane), v2;
  if (n & 1) return key_to_float(v1);
  int32_t cle = 0;
  uint32_t mn = 0xFFFFFFFFu;
  for (int32_t b = 0; b < n; b += 64) {
    int32_t j = b + lane;
    uint32_t kv = j < n ? list[j] : 0xFFFFFFFFu;
    cle += __popcll(__ballot(j < n && kv <= v1));
    if (j < n && kv > v1 && kv < mn) mn = kv;
  }
  v2 = cle > k1 + 1 ? v1 : wave_min_u32(mn);
  float a = key_to_float(v1), b = key_to_float(v2);
  return a + (b - a) / 2.0f;
}

/* ---------------------------------------------------------------------- */
/* orientation of an evaluation                                            */

/* Fills sub (S*U*B) for evaluation e and the voxel slot; returns 1 when the
 * orientation is valid, 0 when its row is skipped (|q_vec| > 1), -1 when the
 * evaluation is an empty CSR row (GRS_MAP_CSR_LIST).  row/jloc receive the CSR
 * row and the index inside the row's expanded list. */
template <int MAP>
__device__ __forceinline__ int eval_setup(const grs_eval_args &a, int64_t e, int64_t *vi, double *sub, int64_t *row,
                                          int64_t *jloc, int64_t *hint = nullptr) {
  if (MAP == GRS_MAP_CSR || MAP == GRS_MAP_CSR_LIST) {
    const grs_csr &c = a.csr;
    int64_t ge = e;
    if (MAP == GRS_MAP_CSR_LIST) {
      ge = c.elist[e];
      if (ge < 0) return -1;
    }
    /* the row of the previous evaluation of this wave, or the next one,
     * usually holds ge (a wave's evaluations ascend); else a binary search */
    int64_t lo = -1;
    if (hint && *hint >= 0) {
      const int64_t h = *hint;
      if (c.row_eoff[h] <= ge) {
        if (h + 1 >= c.nrows || c.row_eoff[h + 1] > ge) lo = h;
        else if (h + 2 >= c.nrows || c.row_eoff[h + 2] > ge) lo = h + 1;
      }
    }
    if (lo < 0) {
      int64_t hi = c.nrows;
      lo = 0;
      while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (c.row_eoff[mid] <= ge) lo = mid;
        else hi = mid;
      }
    }
    if (hint) *hint = lo;
    int64_t j = ge - c.row_eoff[lo];
    int64_t cand = c.row_ptr[lo] + j / c.S;
    int32_t smp = (int32_t)(j % c.S);
    int32_t n1 = c.nr + 1;
    int32_t k1 = smp / (n1 * n1), k2 = (smp / n1) % n1, k3 = smp % n1;
    const double *q = c.cand_q + 4 * cand;
    const double half = (c.gridsize * (double)c.nr) / 2.0;
    double q1 = (q[1] - half) + (double)k1 * c.gridsize;
    double q2 = (q[2] - half) + (double)k2 * c.gridsize;
    double q3 = (q[3] - half) + (double)k3 * c.gridsize;
    double U[9];
    *vi = lo;
    *row = lo;
    *jloc = j;
    if (!grs_U_from_q(q1, q2, q3, U)) return 0;
    grs_SUB(a.g, U, sub);
    return 1;
  } else {
    int64_t ui;
    eval_slots(a, e, vi, &ui);
    *row = e;
    *jloc = 0;
    if (!a.valid[ui]) return 0;
#pragma unroll
    for (int k = 0; k < 9; k++) sub[k] = a.SUB[ui * 9 + k];
    return 1;
  }
}

/* ---------------------------------------------------------------------- */
/* k_fast                                                                  */

/* Counts-only output of a grid evaluation: the counts and the key
 * (class << 32) | (2^32 - 1 - i) of orientation i of its voxel, where class
 * orders C = ni/ns as MATLAB's sortrows(...,'descend') does: NaN (ns = 0)
 * first as 2^31, else floor(C * 2^30) + 1, which separates every two distinct
 * ratios of counts up to GRS_KEY_MAX_EVENTS; a skipped orientation has C = 0
 * (class 1).  Equal classes order by ascending i (a stable sort). */
__device__ __forceinline__ void grid_counts_out(const grs_eval_args &a, int64_t e, int32_t ns, int32_t ni,
                                                bool valid) {
  if (a.out_ns) {
    a.out_ns[e] = ns;
    a.out_ni[e] = ni;
  }
  if (a.out_key) {
    uint64_t cls = 1;
    if (valid) {
      if (ns > 0) cls = (uint64_t)__builtin_floor(((double)ni / (double)ns) * 1073741824.0) + 1ull;
      else cls = 1ull << 31;
    }
    const uint64_t i = (uint64_t)(e % a.n_u);
    a.out_key[e] = (unsigned long long)((cls << 32) | (0xFFFFFFFFull - i));
  }
}

/* Phase B1: resolve an event without the exact pipeline where that is safe.
 *
 * The spot position is recomputed from the same quantities through exact
 * identities -- sin(theta) = -cos(beta), cos(theta) = sqrt(1 - cos^2 beta),
 * lambda = -4 pi cos(beta) / |G|, sin(phix) = sqrt(1 - cos^2 phix),
 * sin(phix - 2 theta) by the difference formula -- with refined hardware
 * reciprocals (relative error below 5e-15).  Guards send ill-conditioned
 * events to the exact pipeline: |sin(phiy)| < 1e-3, |t denominator| < 1e-3,
 * |cos(phix)| > 0.99, G nearly along the beam.  Under the guards both this
 * value and the exact pipeline's are within about 1e-8 pixel of the true spot
 * position, so when the unrounded coordinate is at least GRS_RES_TAU (1e-4
 * pixel) from every half-integer the two round to the same pixel; the window
 * tests are decided here when cos(beta) and lambda clear their bounds by the
 * phase-A margins (1e-12 on cos(beta), 1e-9 lambda_max on lambda), either
 * way.  Everything else (and every NaN) returns 2 and goes
 * through the exact pipeline, so the decisions are those of forward_comp.m.
 * Returns 0 (not selected), 1 (selected, pixel in *dy, *dz) or 2 (exact). */
#define GRS_RES_TAU 1e-4

struct b1_out {
  double ry, rz;
  int32_t dy, dz;
  int32_t win; /* cleared by the caller; 1 once the event enters the spot construction */
};

__device__ __forceinline__ bool res_round(double r, double *k) {
  const double f = r + 0.5;
  const double fl = __builtin_floor(f);
  const double fr = f - fl;
  *k = fl;
  return fr >= GRS_RES_TAU && fr <= 1.0 - GRS_RES_TAU;
}

/* the per (voxel, omega) terms of the spot construction (forward_comp.m:
 * 36-54): the beam direction u, k0 = Lsd - SamposW(1), d = center(2:3) -
 * SamposW(2:3) with center(2:3) = L u(2:3) / u0, L2 = (Lsd - SamposW(1)) /
 * u0 (cos(alpha) = u0), the t numerator and the offsets of the
 * tilted-detector intersection */
struct om_terms {
  double u0, u1, u2, k0, d1, d2, L2, tnum, ov1, ov2;
};

__device__ __forceinline__ void omega_terms(const grs_dgeo *g, const double *P, double c, double s, om_terms *w) {
  double swv[3], uv[3];
  beam_terms(g, P, c, s, swv, uv);
  const double iu0 = fast_rcp(uv[0]);
  w->u0 = uv[0];
  w->u1 = uv[1];
  w->u2 = uv[2];
  w->k0 = g->Lsd - swv[0];
  w->d1 = g->L * uv[1] * iu0 - swv[1];
  w->d2 = g->L * uv[2] * iu0 - swv[2];
  w->L2 = w->k0 * iu0;
  const double *R = g->R;
  w->tnum = __builtin_fma(R[0], w->k0, __builtin_fma(R[3], (g->dety00 - g->RAO) - swv[1], R[6] * (g->detz00 - swv[2])));
  w->ov1 = (swv[1] - g->dety00) + g->RAO;
  w->ov2 = swv[2] - g->detz00;
}

__device__ __forceinline__ int b1_classify(const grs_dgeo *g, double c, double s, const double *P, const double *bt,
                                           const double *gt, const double *thr, b1_out *o) {
  om_terms w;
  if (bt) {
    /* the beam table row of this omega: the same values as omega_terms */
    w.u0 = bt[0];
    w.u1 = bt[1];
    w.u2 = bt[2];
    w.k0 = bt[3];
    w.d1 = bt[4];
    w.d2 = bt[5];
    w.L2 = bt[6];
    w.tnum = bt[7];
    w.ov1 = bt[8];
    w.ov2 = bt[9];
  } else {
    omega_terms(g, P, c, s, &w);
  }
  const double gw0 = gt[0], gw1 = gt[1], gw2 = gt[2];
  const double u0 = w.u0, u1 = w.u1, u2 = w.u2;
  const double G0 = c * gw0 - s * gw1, G1 = s * gw0 + c * gw1, G2 = gw2;
  /* |G| is invariant under the rotation: 1/|G| and |G|^2 of the reflection */
  const double rg = gt[3], g2 = gt[4];
  const double cb = __builtin_fma(u0, G0, __builtin_fma(u1, G1, u2 * G2)) * rg;
  const double lam = (-4.0 * GRS_PI) * cb * rg;
  /* outside the window by more than the margins: not selected */
  if (!(cb < 1e-12 && cb > thr[0] && lam > thr[1] && lam < thr[2])) return cb == cb && lam == lam ? 0 : 2;
  /* inside by less than the margins: the exact pipeline decides */
  if (!(cb < -1e-12 && cb > thr[3] && lam > thr[4] && lam < thr[5])) return 2;
  o->win = 1;
  const double omc = __builtin_fma(-cb, cb, 1.0);
  const double cth = omc * fast_rsq(omc);
  const double s2t = -2.0 * cb * cth;
  const double c2t = __builtin_fma(-2.0 * cb, cb, 1.0);
  const double nv2 = __builtin_fma(G1, G1, G2 * G2);
  if (!(nv2 >= 1e-12 * g2)) return 2;
  /* the remaining per (voxel, omega) terms from u and SamposW: center is
   * grainpos * L / grainpos(1), so center/|center| = u, center(2:3) =
   * L u(2:3) / u0, L2 = (Lsd - SW1) / u0 (cos(alpha) = u0); K_out(2:3) =
   * (dety22, detz22) - SamposW(2:3) = d + diffvec e; and the t numerator
   * and offsets of forward_comp.m:51-54 */
  const double k0 = w.k0, d1 = w.d1, d2 = w.d2, L2 = w.L2;
  const double *R = g->R;
  const double tnum = w.tnum;
  const double ov1 = w.ov1, ov2 = w.ov2;
  const double inv = fast_rsq(nv2);
  const double ap = __builtin_fma(G1, u1, G2 * u2) * inv;
  if (!(__builtin_fabs(ap) <= 0.99)) return 2;
  const double oma = __builtin_fma(-ap, ap, 1.0);
  const double sph = oma * fast_rsq(oma);
  const double sphy = __builtin_fma(sph, c2t, -ap * s2t);
  if (!(__builtin_fabs(sphy) >= 1e-3)) return 2;
  const double diffvec = L2 * s2t * fast_rcp(sphy);
  /* K_out = (Lsd, y22, z22) - SamposW unnormalised: t K_out is the same
   * vector for K_out and K_out / |K_out|, and the guard |R(:,1) . K_out| >=
   * 1e-3 |K_out| is taken squared */
  const double k1 = __builtin_fma(diffvec * G1, inv, d1);
  const double k2 = __builtin_fma(diffvec * G2, inv, d2);
  const double kk = __builtin_fma(k0, k0, __builtin_fma(k1, k1, k2 * k2));
  const double tden = __builtin_fma(R[0], k0, __builtin_fma(R[3], k1, R[6] * k2));
  if (!(tden * tden >= 1e-6 * kk)) return 2;
  const double t = tnum * fast_rcp(tden);
  const double w0 = __builtin_fma(t, k0, -k0), w1 = __builtin_fma(t, k1, ov1), w2 = __builtin_fma(t, k2, ov2);
  /* thr[6..11]: -R(:,2) / pixelysize and -R(:,3) / pixelzsize */
  const double ry = __builtin_fma(thr[6], w0, __builtin_fma(thr[7], w1, __builtin_fma(thr[8], w2, g->dety0)));
  const double rz = __builtin_fma(thr[9], w0, __builtin_fma(thr[10], w1, __builtin_fma(thr[11], w2, g->detz0)));
  o->ry = ry;
  o->rz = rz;
  double ky, kz;
  if (!res_round(ry, &ky) || !res_round(rz, &kz)) return 2;
  /* ky, kz are the rounded pixel: floor(r + 0.5) equals round-half-away for
   * every r that is not a half-integer, and negative values are off anyway */
  if (!(ky >= 1.0 && ky <= g->detysize && kz >= 1.0 && kz <= g->detzsize)) return 0;
  if (ky >= g->bsy0 && ky <= g->bsy1 && kz >= g->bsz0 && kz <= g->bsz1) return 0;
  o->dy = (int32_t)ky;
  o->dz = (int32_t)kz;
  return 1;
}

/* The exact forward_comp.m:28-65 pipeline of one event (omega index om,
 * reflection h): returns 1 and the stack index when the event is selected;
 * yz receives the unrounded pixel coordinates of a selected event.  The
 * per (voxel, omega) terms are recomputed here with the function of the
 * exact pipeline (grs_vw_terms).  Kept out of line: it runs for about 1e-4 of the events, and inlining it would raise
 * the register pressure of the whole kernel. */
__device__ __noinline__ int exact_event(const grs_dgeo *g, const double *pos, const double *cs, const double *gwt,
                                        uint32_t ev, uint32_t *idx, double *yz) {
  const int32_t om = (int32_t)(ev >> 8), h = (int32_t)(ev & 255u);
  grs_vw v;
  grs_vw_terms(g, pos, cs[2 * om], cs[2 * om + 1], &v);
  const double *t = gwt + h * GRS_GWT;
  double Gt[3];
  grs_Gt(v.c, v.s, t[0], t[1], t[2], Gt);
  double Glen;
  double beta = grs_event_beta(&v, Gt, &Glen);
  int32_t dy = 0, dz = 0;
  double w[2] = {0.0, 0.0};
  if (!grs_event_finish_yz(g, &v, Gt, Glen, beta, &dy, &dz, w)) return 0;
  *idx = (uint32_t)((int64_t)(dz - 1) + (int64_t)(dy - 1) * g->nz + (int64_t)om * g->nz * g->ny);
  yz[0] = w[0];
  yz[1] = w[1];
  return 1;
}

/* Omega solve of one reflection.  With (g0, g1) the in-plane part of a
 * vector that rotates with omega, |G| = gn and r = |(g0, g1)|, the term
 * r cos(omega + phi) / gn, phi = atan2(g1, g0), must lie in [lo_cb, hi_cb]
 * (the caller folds every other part of cos(beta), forward_comp.m:31, and its
 * bounds into the window), so omega + phi must lie in the arcs where
 * cos(psi) is in [lo_cb gn / r, hi_cb gn / r] (gn / r from a refined
 * reciprocal square root).  The arc angles come from one call of
 * arc_ends (errors below 5e-12 rad); the arcs are widened by 1e-6 rad and
 * the index bounds by 1e-6 steps.  The ranges of rotation indices [i0, i1]
 * land in six fixed slots (two arcs x three periods) and are disjoint, in
 * no particular order.  Returns the bit mask of the slots in use (at most
 * four) or -1 when more would be needed (the caller then scans every
 * event). */
__device__ __forceinline__ int om_ranges(const grs_dgeo *g, double g0, double g1, double gn, double lo_cb,
                                         double hi_cb, int32_t nrot, double idom, int32_t *i0, int32_t *i1) {
  const double twopi = 2.0 * GRS_PI, eps = 1e-6;
  const double r2 = g0 * g0 + g1 * g1;
  const double ir = r2 > 0.0 ? fast_rsq(r2) : 0.0;
  const double r = r2 > 0.0 ? r2 * ir : r2 * 0.0; /* NaN stays NaN */
  if (lo_cb != lo_cb || hi_cb != hi_cb || r != r || !(gn > 0.0)) return -1;
  if (!(hi_cb > lo_cb)) return 0;
  double lo[2], hi[2];
  int niv = 0;
  bool whole = false;
  if (!(r > 1e-12 * gn)) {
    /* |cos(beta) x component| <= r / gn <= 1e-12 at every omega */
    whole = lo_cb <= 1e-12 && hi_cb >= -1e-12;
    if (!whole) return 0;
  } else {
    const double gr = gn * ir;
    const double a = lo_cb * gr, b = hi_cb * gr;
    if (!(a < 1.0) || !(b > -1.0)) return 0;
    const arc_angles an = arc_ends(a, b, g1, g0);
    const double pa = an.pa, pb = an.pb, phi = an.phi;
    /* the gaps between the two arcs, around psi = 0 and psi = pi: one of at
     * most 1e-5 rotation steps is closed (its events join the candidates),
     * so that the index ranges of different arcs and periods never share
     * an index */
    const bool gap0 = (2.0 * (pb - eps)) * idom <= 1e-5;
    const bool gap_pi = (twopi - 2.0 * (pa + eps)) * idom <= 1e-5;
    if (gap0 && gap_pi) {
      whole = true;
    } else if (gap0) {
      lo[0] = (-pa - eps) - phi;
      hi[0] = (pa + eps) - phi;
      niv = 1;
    } else if (gap_pi) {
      lo[0] = (pb - eps) - phi;
      hi[0] = (twopi - pb + eps) - phi;
      niv = 1;
    } else {
      lo[0] = (pb - eps) - phi;
      hi[0] = (pa + eps) - phi;
      lo[1] = (-pa - eps) - phi;
      hi[1] = (-pb + eps) - phi;
      niv = 2;
    }
  }
  if (whole) {
    i0[0] = 0;
    i1[0] = nrot - 1;
    return 1;
  }
  /* slot 3 k + j: arc k in period m + j.  The arcs and their periods are
   * more than 1e-5 steps apart, so the ranges are disjoint as they come;
   * their order does not matter to the caller */
  uint32_t mask = 0;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    if (k >= niv) break;
    /* the last period whose arc ends at or before om0.  The product rounds
     * differently from the quotient only next to an integer, where m may be
     * one less (that arc ends 2 pi below om0 and is skipped; the periods
     * the scan can need, m..m+2 then, are still within the four steps) or
     * one more (the period passed over ends about 2 pi below om0) */
    const double m = __builtin_floor((g->om0 - hi[k]) * (1.0 / twopi));
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double L = lo[k] + twopi * (m + (double)j), U = hi[k] + twopi * (m + (double)j);
      double f0 = __builtin_ceil((L - g->om0) * idom - 1e-6);
      double f1 = __builtin_floor((U - g->om0) * idom + 1e-6);
      f0 = f0 > 0.0 ? f0 : 0.0;
      f1 = f1 < (double)(nrot - 1) ? f1 : (double)(nrot - 1);
      const bool ok = L <= g->om_last + 1e-9 && U >= g->om0 - 1e-9 && f0 <= f1;
      if (j == 3) {
        if (ok) return -1; /* a scan longer than the periods kept here */
      } else {
        i0[3 * k + j] = (int32_t)f0;
        i1[3 * k + j] = (int32_t)f1;
        mask |= (ok ? 1u : 0u) << (3 * k + j);
      }
    }
  }
  if (__popc(mask) > 4) return -1;
  return (int)mask;
}

template <bool MEDIAN, int MAP, bool AUDIT>
__global__ __launch_bounds__(64 * GRS_WPB, GRS_FAST_MIN_WAVES) void k_fast(grs_eval_args a, int32_t H) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int64_t per_wave = fast_wave_doubles(H);
  double *gwt = smem + wib * per_wave;
  uint32_t *ring1 = (uint32_t *)(gwt + (int64_t)H * GRS_GWT); /* phase-A survivors (omega << 8 | hkl) */
  uint32_t *ring2 = ring1 + GRS_QCAP;                          /* events for the exact pipeline */
  uint32_t *ring3 = ring2 + GRS_QCAP;                          /* selected events: stack index */
  uint32_t *rstart = ring3 + GRS_QCAP;                         /* omega-solve ranges: (i0 << 8 | hkl) */
  uint32_t *rcum = rstart + 4 * H;                             /* exclusive prefix of range lengths */
  double *Pw = gwt + (int64_t)H * GRS_GWT + GRS_RING_DOUBLES + GRS_RANGE_DOUBLES(H); /* S*pos of the voxel */
  /* S*U*B of the evaluation: the omega solve reads it from LDS, which keeps
   * its 18 registers free for the candidate loop */
  double *sbuf = Pw + 4;
  /* the geometry constants live in LDS: the loop bodies read ~30 of them per
   * batch, and neither scalar registers (spilled) nor vector loads (the
   * compiler cannot prove them unchanged across the loop's stores) serve */
  __shared__ grs_dgeo sg;
  /* b1_classify's window thresholds (outer, then inner margins) and the
   * detector rows scaled to pixels */
  __shared__ double b1thr[12];
  if (threadIdx.x == 0) {
    sg = *a.g;
    const double lm = 1e-9 * sg.lambda_max;
    b1thr[0] = sg.cos_beta_max - 1e-12;
    b1thr[1] = sg.lambda_min - lm;
    b1thr[2] = sg.lambda_max + lm;
    b1thr[3] = sg.cos_beta_max + 1e-12;
    b1thr[4] = sg.lambda_min + lm;
    b1thr[5] = sg.lambda_max - lm;
    for (int k = 0; k < 3; k++) {
      b1thr[6 + k] = -sg.R[3 * k + 1] * sg.ipy;
      b1thr[9 + k] = -sg.R[3 * k + 2] * sg.ipz;
    }
  }
  __syncthreads();
  const grs_dgeo *g = &sg;
  const int wpb = (int)(blockDim.x >> 6);
  /* one voxel for every evaluation: its per-omega beam terms are computed
   * once per block into LDS behind the per-wave regions (the launch sizes
   * the allocation with the same rule) */
  const double *btab = nullptr;
  if (MAP == GRS_MAP_ONE_VOXEL && a.nrot == g->nrot && fast_tab_doubles(MAP, a.nrot, H, wpb) > 0 &&
      a.e0 < a.n_eval) {
    double *bt = smem + (int64_t)wpb * per_wave;
    double P[3];
    voxel_P(g, a.pos, P);
    for (int32_t om = (int32_t)threadIdx.x; om < g->nrot; om += (int32_t)blockDim.x) {
      om_terms w;
      omega_terms(g, P, a.cs[2 * om], a.cs[2 * om + 1], &w);
      double *r = bt + GRS_BT * om;
      r[0] = w.u0;
      r[1] = w.u1;
      r[2] = w.u2;
      r[3] = w.k0;
      r[4] = w.d1;
      r[5] = w.d2;
      r[6] = w.L2;
      r[7] = w.tnum;
      r[8] = w.ov1;
      r[9] = w.ov2;
      r[10] = a.cs[2 * om];
      r[11] = a.cs[2 * om + 1];
    }
    __syncthreads();
    btab = bt;
  }
  /* blocks are dealt to the 8 XCDs in turn (block b on XCD b mod 8): with
   * xcd_map, block b takes the place b / 8 + (b mod 8) * gridDim.x / 8 in
   * the evaluation order, so that each XCD's L2 sees a contiguous run of
   * evaluations (neighbouring samples of a CSR row share detector tiles) */
  int64_t vblock = blockIdx.x;
  if (a.xcd_map && (gridDim.x & 7u) == 0u) vblock = (int64_t)(blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  const int64_t gwave = vblock * wpb + wib;
  const int64_t nwaves = (int64_t)gridDim.x * wpb;
  uint32_t *wlist = MEDIAN ? a.scratch + gwave * a.scratch_stride : nullptr;
  const int32_t nrot = g->nrot;
  const int64_t nz = g->nz, nzny = g->nz * g->ny;
  const uint32_t nz32 = (uint32_t)nz, nzny32 = (uint32_t)nzny; /* the stack holds fewer than 2^32 entries */
  /* counts-only on the tiled mask: the selected events carry tiled bit numbers */
  const bool tiled = !MEDIAN && a.bits_tiled && !a.count_keys;
  const uint32_t ntz = (uint32_t)((nz + 31) >> 5), nty = (uint32_t)((g->ny + 31) >> 5);
  const uint32_t *hbits = tiled ? a.hit_tbits : a.hit_bits;
  const double m4pi = -4.0 * GRS_PI;
  /* phase A margins: see the header comment */
  const double dcb = 1e-12;
  const double cbm = g->cos_beta_max - dcb;
  const double lmin = g->lambda_min, lmax = g->lambda_max;
  const double lrel = 1e-9 * lmax;
  const double inv4pi = 1.0 / (4.0 * GRS_PI);
  const double idom = g->om_solve ? 1.0 / g->dom : 0.0;
  /* starting (omega, hkl) of each lane in the flattened event space */
  const int32_t om0 = lane / H, h0 = lane - (lane / H) * H;

  int64_t row_hint = -1;
  for (int64_t e = a.e0 + gwave; e < a.n_eval; e += nwaves) {
    int64_t vi, row, jloc;
    double sub[9];
    int ok = eval_setup<MAP>(a, e, &vi, sub, &row, &jloc, &row_hint);
    if (ok <= 0) {
      if (!MEDIAN) {
        /* a skipped orientation counts as C = NaN for the argmax; in a grid
         * it has C = 0 */
        if (MAP == GRS_MAP_GRID) {
          if (lane == 0) grid_counts_out(a, e, 0, 0, false);
        } else if (lane == 0 && ok == 0) {
          atomicMax(a.csr.row_best + row, (unsigned long long)(((1ull << GRS_KEY_JBITS) - 1) - (uint64_t)jloc));
        }
      } else if (lane == 0) {
        a.out_ns[e] = 0;
        a.out_ni[e] = 0;
        a.out_C[e] = 0.0;
        a.out_med[e] = 0.0f;
      }
      continue;
    }
    /* per-reflection terms; with the omega solve also the candidate ranges */
    const bool solve = g->om_solve != 0;
    if (lane == 0) {
      voxel_P(g, a.pos + 3 * vi, Pw);
#pragma unroll
      for (int k = 0; k < 9; k++) sbuf[k] = sub[k];
    }
    wave_sync();
    double vW = 0.0, vu2lo = 0.0, vu2hi = 0.0, sK = 1.0, val = 0.0, vbe = 0.0, vga = 0.0;
    double vzt = 1e300, vzb = -1e300, vsn = 1.0, vcn = 0.0, vte = 0.0, vya = 1e300, vdl = 0.0;
    if (solve) {
      const double *b = a.vb + GRS_VB * vi;
      vW = b[0];
      vu2lo = b[1];
      vu2hi = b[2];
      sK = __builtin_sqrt(b[3]);
      val = b[4];
      vbe = b[5];
      vga = b[6];
      vzt = b[7];
      vzb = b[8];
      vsn = b[9];
      vcn = b[10];
      vte = b[11];
      vya = b[12];
      vdl = b[13];
    }
    uint32_t R = 0;
    bool overflow = false;
    for (int hb = 0; hb < H; hb += 64) {
      const int h = hb + lane;
      int nr_h = 0; /* bit mask of the valid slots of ri0/ri1, or -1 */
      int32_t ri0[6], ri1[6];
      if (h < H) {
        double gw[3];
        grs_Gw(sbuf, a.hkl + 3 * h, gw);
        double n2 = (gw[0] * gw[0] + gw[1] * gw[1]) + gw[2] * gw[2];
        const double ign = fast_rsq(n2);
        const double gn = n2 * ign;
        double *t = gwt + h * GRS_GWT;
        t[0] = gw[0];
        t[1] = gw[1];
        t[2] = gw[2];
        t[3] = ign;
        t[4] = n2;
        double ml = lrel + (1e-12 * (4.0 * GRS_PI)) * ign; /* absolute lambda margin */
        if (solve) {
          /* cos(beta) window: beta and lambda bounds with the phase-A margins,
           * the detector cone and the detector rows and columns.  cos(beta) = u . G/gn lies within rg W
           * of r' cos(omega + phi') / gn + c0 + u2 g2 / gn, where
           * (g0', g1') = (g0 + ga g1, g1 - ga g0) has modulus r' and angle
           * phi' and u2 spans [vb[1], vb[2]] (the voxel bounds), so the
           * window moves onto the first term. */
          double lo_cb = cbm;
          double lw = -(lmax + ml) * gn * inv4pi;
          lo_cb = lw > lo_cb ? lw : lo_cb;
          lo_cb = -sK > lo_cb ? -sK : lo_cb;
          const double z = gw[2] * ign;
          const double st = det_sin_theta(z, -lo_cb, vzt, vzb, vya, vdl, vsn, vcn, vte);
          lo_cb = -st > lo_cb ? -st : lo_cb;
          double hi_cb = dcb;
          double hw = -(lmin - ml) * gn * inv4pi;
          hi_cb = hw < hi_cb ? hw : hi_cb;
          const double q2 = gw[0] * gw[0] + gw[1] * gw[1];
          const double rg = q2 > 0.0 ? q2 * fast_rsq(q2) * ign : 0.0;
          const double c0 = (val * gw[0] + vbe * gw[1]) * (0.5 * ign);
          const double sa = z * vu2lo, sb = z * vu2hi;
          const double slo = sa < sb ? sa : sb, shi = sa < sb ? sb : sa;
          const double w = rg * vW;
          nr_h = om_ranges(g, gw[0] + vga * gw[1], gw[1] - vga * gw[0], gn, ((lo_cb - c0) - shi) - w - 1e-12,
                           ((hi_cb - c0) - slo) + w + 1e-12, nrot, idom, ri0, ri1);
        }
      }
      if (solve) {
        overflow = overflow || (__ballot(nr_h < 0) != 0);
        const uint32_t msk = nr_h > 0 ? (uint32_t)nr_h : 0u;
        const int cnt = __popc(msk);
        int incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
          int v = __shfl_up(incl, off, 64);
          if (lane >= off) incl += v;
        }
        uint32_t pos = R + (uint32_t)(incl - cnt);
#pragma unroll
        for (int k = 0; k < 6; k++) {
          if ((msk >> k) & 1u) {
            rstart[pos] = ((uint32_t)ri0[k] << 8) | (uint32_t)h;
            rcum[pos] = (uint32_t)(ri1[k] - ri0[k] + 1);
            pos++;
          }
        }
        R += (uint32_t)__shfl(incl, 63, 64);
      }
    }
    wave_sync();
    uint32_t total = 0;
    if (solve && !overflow) {
      /* lengths -> exclusive prefix sums */
      for (uint32_t cb = 0; cb < R; cb += 64) {
        const uint32_t r = cb + lane;
        const uint32_t len = r < R ? rcum[r] : 0u;
        uint32_t incl = len;
        for (int off = 1; off < 64; off <<= 1) {
          uint32_t v = (uint32_t)__shfl_up((int)incl, off, 64);
          if (lane >= off) incl += v;
        }
        wave_sync();
        if (r < R) rcum[r] = total + incl - len;
        total += (uint32_t)__shfl((int)incl, 63, 64);
        wave_sync();
      }
      if (lane == 0) rcum[R] = total;
      wave_sync();
    }

    const double *vpos = a.pos + 3 * vi;
    uint32_t head1 = 0, tail1 = 0, head2 = 0, tail2 = 0, head3 = 0, tail3 = 0;
    int32_t ns = 0, ni = 0;
    /* the gather of the previous batch is consumed one batch later */
    bool p_sel = false, p_hit = false;
    uint32_t p_key = 0;

    auto consume = [&]() {
      uint64_t ms = __ballot(p_sel);
      if (MEDIAN && p_sel) wlist[ns + (int32_t)lane_prefix(ms)] = p_key;
      ns += (int32_t)__popcll(ms);
      bool hit;
      if (MEDIAN) hit = p_sel && g->has_hit && p_key <= g->hit_key;
      else hit = p_hit;
      ni += (int32_t)__popcll(__ballot(hit));
      p_sel = false;
      p_hit = false;
    };

    /* gather: distance key (or hit bit) of n selected events of ring3 */
    auto gather = [&](uint32_t n) {
      consume();
      if ((uint32_t)lane < n) {
        uint32_t idx = ring3[(head3 + lane) & (GRS_QCAP - 1)];
        p_sel = true;
        if (MEDIAN) p_key = a.keys[idx];
        else if (a.count_keys) p_hit = g->has_hit && a.keys[idx] <= g->hit_key;
        else p_hit = (hbits[idx >> 5] >> (idx & 31u)) & 1u;
      }
      head3 += n;
    };

    auto push3 = [&](bool sel, uint32_t idx) {
      uint64_t m = __ballot(sel);
      if (sel) ring3[(tail3 + lane_prefix(m)) & (GRS_QCAP - 1)] = idx;
      tail3 += (uint32_t)__popcll(m);
      if (tail3 - head3 >= 64) {
        wave_sync();
        gather(64);
        wave_sync();
      }
    };

    /* phase B2: exact pipeline on n events of ring2 */
    auto b2 = [&](uint32_t n) {
      bool sel = false;
      uint32_t idx = 0;
      if ((uint32_t)lane < n) {
        double yz[2];
        sel = exact_event(a.g, vpos, a.cs, gwt, ring2[(head2 + lane) & (GRS_QCAP - 1)], &idx, yz) != 0;
        if (sel && tiled) {
          const uint32_t omx = idx / nzny32, rem = idx - omx * nzny32, y0 = rem / nz32;
          idx = grs_tiled_bit(omx, y0 + 1u, rem - y0 * nz32 + 1u, ntz, nty);
        }
      }
      head2 += n;
      push3(sel, idx);
    };

    /* phase B1: resolve n events of ring1 or queue them for the exact pipeline */
    auto b1_core = [&](bool valid, uint32_t ev, uint32_t n) {
      int st = 0;
      uint32_t idx = 0;
      bool bad = false, inwin = false;
      double dev = 0.0;
      if (valid) {
        int32_t om = (int32_t)(ev >> 8), h = (int32_t)(ev & 255u);
        const double *t = gwt + h * GRS_GWT;
        b1_out o;
        o.ry = o.rz = 0.0;
        o.dy = o.dz = 0;
        o.win = 0;
        const double *brow = btab ? btab + GRS_BT * om : nullptr;
        const double cc = brow ? brow[10] : a.cs[2 * om], ss = brow ? brow[11] : a.cs[2 * om + 1];
        st = b1_classify(g, cc, ss, Pw, brow, t, b1thr, &o);
        if (st == 1)
          idx = tiled ? grs_tiled_bit((uint32_t)om, (uint32_t)o.dy, (uint32_t)o.dz, ntz, nty)
                      : (uint32_t)(o.dz - 1) + (uint32_t)(o.dy - 1) * nz32 + (uint32_t)om * nzny32;
        inwin = o.win != 0;
        if (AUDIT) {
          double yz[2] = {0.0, 0.0};
          uint32_t eidx = 0;
          const int es = exact_event(a.g, vpos, a.cs, gwt, ev, &eidx, yz);
          bad = (st == 0 && es) || (st == 1 && (!es || eidx != idx));
          if (st == 1 && es) dev = __builtin_fmax(__builtin_fabs(o.ry - yz[0]), __builtin_fabs(o.rz - yz[1]));
        }
      }
      if (AUDIT) {
        /* whole-wave reductions, outside the divergent block */
        uint64_t mb = __ballot(bad), m1 = __ballot(st == 1), m0 = __ballot(valid && st == 0);
        uint64_t m2 = __ballot(st == 2), mw = __ballot(inwin);
        for (int off = 32; off >= 1; off >>= 1) dev = __builtin_fmax(dev, __shfl_xor(dev, off, 64));
        if (lane == 0) {
          atomicAdd(a.audit + 0, (unsigned long long)n);
          atomicAdd(a.audit + 1, (unsigned long long)__popcll(m1));
          atomicAdd(a.audit + 2, (unsigned long long)__popcll(m0));
          atomicAdd(a.audit + 3, (unsigned long long)__popcll(m2));
          atomicAdd(a.audit + 4, (unsigned long long)__popcll(mb));
          atomicMax(a.audit + 5, (unsigned long long)__double_as_longlong(dev));
          atomicAdd(a.audit + 6, (unsigned long long)__popcll(mw));
        }
      }
      uint64_t m2 = __ballot(st == 2);
      if (st == 2) ring2[(tail2 + lane_prefix(m2)) & (GRS_QCAP - 1)] = ev;
      tail2 += (uint32_t)__popcll(m2);
      push3(st == 1, idx);
      if (tail2 - head2 >= 64) {
        wave_sync();
        b2(64);
        wave_sync();
      }
    };

    auto b1 = [&](uint32_t n) {
      const bool valid = (uint32_t)lane < n;
      const uint32_t ev = valid ? ring1[(head1 + lane) & (GRS_QCAP - 1)] : 0u;
      head1 += n;
      b1_core(valid, ev, n);
    };

    if (solve && !overflow) {
      /* candidates of the omega solve, 64 per step, straight to phase B1.
       * Ranges are non-empty and rcum is their exclusive prefix, so the range
       * of candidate base + l is rb plus the number of later ranges that
       * start at or before base + l, where rb is the last range starting
       * before base (0 at base 0).  Those starts lie in the next 64 ranges;
       * each marks its position in ring1 (unused on this path) and a ballot
       * turns the marks into a bit mask. */
      uint32_t rb = 0;
      for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t idx = base + (uint32_t)lane;
        const bool valid = idx < total;
        ring1[lane] = 0u;
        wave_sync();
        const uint32_t rn = rb + 1u + (uint32_t)lane;
        if (rn < R) {
          const uint32_t p = rcum[rn] - base;
          if (p < 64u) ring1[p] = 1u;
        }
        wave_sync();
        const uint64_t starts = __ballot(ring1[lane] != 0u);
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const uint32_t r = rb + (uint32_t)__popcll(starts & upto);
        rb += (uint32_t)__popcll(starts);
        uint32_t ev = 0;
        if (valid) {
          const uint32_t st0 = rstart[r];
          ev = (((st0 >> 8) + (idx - rcum[r])) << 8) | (st0 & 255u);
        }
        b1_core(valid, ev, total - base < 64u ? total - base : 64u);
      }
    }
    /* phase A (no omega solve): wave-uniform steps of 64 consecutive events;
     * lane l owns event l + 64*k, tracked as (omega, hkl) without divisions */
    int32_t om = om0, h = h0;
    const int32_t nsteps = (solve && !overflow) ? 0 : (nrot * H + 63) / 64;
    for (int32_t k = 0; k < nsteps; ++k) {
      bool pass = false;
      if (om < nrot) {
        const double c = a.cs[2 * om], s = a.cs[2 * om + 1];
        double swv[3], uv[3];
        beam_terms(g, Pw, c, s, swv, uv);
        const double u0 = uv[0], u1 = uv[1], u2 = uv[2];
        const double *t = gwt + h * GRS_GWT;
        double Gt0 = c * t[0] - s * t[1];
        double Gt1 = s * t[0] + c * t[1];
        double dot = __builtin_fma(u2, t[2], __builtin_fma(u1, Gt1, u0 * Gt0));
        double l4 = m4pi * dot;
        /* the phase-A thresholds of the reflection (margins: see the header comment) */
        const double gn = __builtin_sqrt(t[4]);
        const double ml = lrel + (1e-12 * (4.0 * GRS_PI)) * t[3];
        pass = (dot < dcb * gn) && (dot > cbm * gn) && (l4 > (lmin - ml) * t[4]) && (l4 < (lmax + ml) * t[4]);
      }
      uint64_t m = __ballot(pass);
      if (pass) ring1[(tail1 + lane_prefix(m)) & (GRS_QCAP - 1)] = ((uint32_t)om << 8) | (uint32_t)h;
      tail1 += (uint32_t)__popcll(m);
      if (tail1 - head1 >= 64) {
        wave_sync();
        b1(64);
        wave_sync();
      }
      h += 64;
      while (h >= H) {
        h -= H;
        om++;
      }
    }
    while (tail1 != head1) {
      uint32_t n = tail1 - head1;
      if (n > 64) n = 64;
      wave_sync();
      b1(n);
      wave_sync();
    }
    while (tail2 != head2) {
      uint32_t n = tail2 - head2;
      if (n > 64) n = 64;
      wave_sync();
      b2(n);
      wave_sync();
    }
    while (tail3 != head3) {
      uint32_t n = tail3 - head3;
      if (n > 64) n = 64;
      wave_sync();
      gather(n);
      wave_sync();
    }
    consume();
    if (MEDIAN) {
      __threadfence_block();
      float med = 2000.0f;
      if (ns > 0) med = wave_median(wlist, ns, lane);
      if (lane == 0) {
        a.out_ns[e] = ns;
        a.out_ni[e] = ni;
        a.out_C[e] = (double)ni / (double)ns;
        a.out_med[e] = med;
      }
    } else if (MAP == GRS_MAP_GRID) {
      if (lane == 0) grid_counts_out(a, e, ns, ni, true);
    } else if (lane == 0) {
      uint64_t jpart = ((1ull << GRS_KEY_JBITS) - 1) - (uint64_t)jloc;
      uint64_t cpart = 0;
      if (ns > 0) {
        double C = (double)ni / (double)ns;
        cpart = (uint64_t)__builtin_floor(C * 1073741824.0) + 1ull;
      }
      atomicMax(a.csr.row_best + row, (unsigned long long)((cpart << GRS_KEY_JBITS) | jpart));
    }
    wave_sync();
  }
}

/* ---------------------------------------------------------------------- */
/* batched compete: device-side steps between and after the two launches   */

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

/* winner of each CSR row from its packed argmax key: the index of the
 * first maximum in the expanded list (-1 for an empty row) */
__global__ void k_winner_list(const unsigned long long *row_best, const int64_t *row_eoff, int64_t m, int64_t *elist) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  if (row_eoff[r + 1] == row_eoff[r]) {
    elist[r] = -1;
    return;
  }
  const unsigned long long jmask = (1ull << GRS_KEY_JBITS) - 1;
  elist[r] = row_eoff[r] + (int64_t)(jmask - (row_best[r] & jmask));
}

/* best index in the row's expanded list and its quaternion (the local
 * sample of the grid, q0 from the vector part, 0 when |q_vec| > 1) */
__global__ void k_winner_out(const int64_t *elist, const int64_t *row_eoff, const int64_t *row_ptr,
                             const double *cand_q, int32_t nr, double gridsize, int64_t m, int64_t *best,
                             double *best_q) {
  GRS_NO_CONTRACT
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double *bq = best_q + 4 * r;
  if (elist[r] < 0) {
    best[r] = -1;
    bq[0] = bq[1] = bq[2] = bq[3] = 0.0;
    return;
  }
  const int64_t n1 = nr + 1, S = n1 * n1 * n1;
  const int64_t j = elist[r] - row_eoff[r];
  const int64_t cand = row_ptr[r] + j / S;
  const int64_t smp = j % S;
  const int64_t k[3] = {smp / (n1 * n1), (smp / n1) % n1, smp % n1};
  const double *q = cand_q + 4 * cand;
  const double half = (gridsize * (double)nr) / 2.0;
  double v[3];
  for (int t = 0; t < 3; t++) v[t] = (q[1 + t] - half) + (double)k[t] * gridsize;
  const double ss = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
  bq[0] = ss <= 1.0 ? __builtin_sqrt(1.0 - ss) : 0.0;
  bq[1] = v[0];
  bq[2] = v[1];
  bq[3] = v[2];
  best[r] = j;
}

/* bit i of the hit mask = (keys[i] <= hit_key) */
__global__ void k_pack_hits(const uint32_t *keys, int64_t n, uint32_t hit_key, int32_t has_hit, uint32_t *bits) {
  int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nw = (n + 31) >> 5;
  if (w >= nw) return;
  uint32_t b = 0;
  if (has_hit) {
    for (int k = 0; k < 32; k++) {
      int64_t i = (w << 5) + k;
      if (i < n && keys[i] <= hit_key) b |= 1u << k;
    }
  }
  bits[w] = b;
}

/* the same mask in 32 x 32 pixel tiles (grs_tiled_bit): thread per 32-bit
 * word, word = tile * 32 + (dy - 1) mod 32, bit = (dz - 1) mod 32; pixels
 * outside the stack are 0 */
__global__ void k_pack_hits_tiled(const uint32_t *keys, int64_t nz, int64_t ny, int64_t nw, uint32_t ntz,
                                  uint32_t nty, uint32_t hit_key, int32_t has_hit, uint32_t *tbits) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int64_t tile = w >> 5;
  const int64_t tz = tile % ntz, rest = tile / ntz;
  const int64_t ty = rest % nty, om = rest / nty;
  const int64_t y = ty * 32 + (w & 31);
  uint32_t b = 0;
  if (has_hit && y < ny) {
    const uint32_t *col = keys + (om * ny + y) * nz;
    for (int k = 0; k < 32; k++) {
      const int64_t z = tz * 32 + k;
      if (z < nz && col[z] <= hit_key) b |= 1u << k;
    }
  }
  tbits[w] = b;
}

/* ---------------------------------------------------------------------- */
/* k_plain                                                                 */

__device__ uint32_t plain_select(uint32_t *A, int64_t T, int32_t n, int32_t k) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t pivot = A[(int64_t)((lo + hi) >> 1) * T];
    int32_t i = lo, j = hi;
    while (i <= j) {
      while (A[(int64_t)i * T] < pivot) i++;
      while (A[(int64_t)j * T] > pivot) j--;
      if (i <= j) {
        uint32_t tmp = A[(int64_t)i * T];
        A[(int64_t)i * T] = A[(int64_t)j * T];
        A[(int64_t)j * T] = tmp;
        i++;
        j--;
      }
    }
    if (k <= j) hi = j;
    else if (k >= i) lo = i;
    else break;
  }
  return A[(int64_t)k * T];
}

__global__ __launch_bounds__(256) void k_plain(grs_eval_args a, int64_t T) {
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= T) return;
  const grs_dgeo *g = a.g;
  const int32_t nrot = g->nrot, H = g->H;
  uint32_t *buf = a.scratch + tid; /* element j at buf[j*T] */
  for (int64_t e = tid; e < a.n_eval; e += T) {
    int64_t vi, ui;
    eval_slots(a, e, &vi, &ui);
    if (!a.valid[ui]) {
      a.out_ns[e] = 0;
      a.out_ni[e] = 0;
      a.out_C[e] = 0.0;
      a.out_med[e] = 0.0f;
      continue;
    }
    const double *SUB = a.SUB + ui * 9;
    const double *pos = a.pos + vi * 3;
    int32_t ns = 0, ni = 0;
    for (int32_t om = 0; om < nrot; om++) {
      grs_vw v;
      grs_vw_terms(g, pos, a.cs[2 * om], a.cs[2 * om + 1], &v);
      for (int32_t h = 0; h < H; h++) {
        double gw[3], Gt[3], Glen;
        grs_Gw(SUB, a.hkl + 3 * h, gw);
        grs_Gt(v.c, v.s, gw[0], gw[1], gw[2], Gt);
        double beta = grs_event_beta(&v, Gt, &Glen);
        int32_t dy, dz;
        if (grs_event_finish(g, &v, Gt, Glen, beta, &dy, &dz)) {
          uint32_t key = a.keys[(int64_t)(dz - 1) + (int64_t)(dy - 1) * g->nz + (int64_t)om * g->nz * g->ny];
          buf[(int64_t)ns * T] = key;
          ns++;
          if (g->has_hit && key <= g->hit_key) ni++;
        }
      }
    }
    float med = 2000.0f;
    if (ns > 0) {
      int32_t k1 = (ns & 1) ? (ns >> 1) : (ns >> 1) - 1;
      uint32_t v1 = plain_select(buf, T, ns, k1);
      if (ns & 1) {
        med = key_to_float(v1);
      } else {
        uint32_t v2 = 0xFFFFFFFFu;
        for (int32_t j = k1 + 1; j < ns; j++) {
          uint32_t x = buf[(int64_t)j * T];
          v2 = x < v2 ? x : v2;
        }
        float fa = key_to_float(v1), fb = key_to_float(v2);
        med = fa + (fb - fa) / 2.0f;
      }
    }
    a.out_ns[e] = ns;
    a.out_ni[e] = ni;
    a.out_C[e] = (double)ni / (double)ns;
    a.out_med[e] = med;
  }
}

/* ---------------------------------------------------------------------- */
/* compete: first maximum of C per voxel row, NaN never wins              */

__global__ void k_compete_reduce(const int64_t *row_start, int64_t M, const int32_t *ns, const int32_t *ni,
                                 const double *C, const float *med, int64_t *best) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  int64_t a = row_start[r], b = row_start[r + 1];
  int64_t bi = -1;
  double bc = 0.0;
  for (int64_t j = a; j < b; j++) {
    double c = C[j];
    if (c == c && (bi < 0 || c > bc)) {
      bi = j;
      bc = c;
    }
  }
  if (bi < 0 && b > a) bi = a;
  best[r] = bi;
}

/* ---------------------------------------------------------------------- */
/* launchers                                                               */

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t grs_launch_omega_cs(const double *d_rot, int32_t nrot, double *d_cs, hipStream_t st) {
  if (nrot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_omega_cs, dim3(blocks_for(nrot, 256)), dim3(256), 0, st, d_rot, nrot, d_cs);
  return hipGetLastError();
}

hipError_t grs_launch_voxel_pos(const double *d_rows, int64_t M, int32_t stride, const double *d_p1, double *d_pos,
                                hipStream_t st) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_voxel_pos, dim3(blocks_for(M, 256)), dim3(256), 0, st, d_rows, M, stride, d_p1, d_pos);
  return hipGetLastError();
}

hipError_t grs_launch_voxel_bounds(const grs_dgeo *d_g, const double *d_pos, int64_t V, const double *d_cs,
                                   int32_t nrot, double *d_vb, hipStream_t st) {
  if (V <= 0) return hipSuccess;
  if (V > 0x7fffffff) return hipErrorInvalidValue;
  if (V >= 4096)
    hipLaunchKernelGGL(k_voxel_bounds_seq, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, st, d_g, d_pos, V, d_cs,
                       nrot, d_vb);
  else
    hipLaunchKernelGGL(k_voxel_bounds, dim3((unsigned)V), dim3(256), 0, st, d_g, d_pos, V, d_cs, nrot, d_vb);
  return hipGetLastError();
}

hipError_t grs_launch_sub_from_q(const grs_dgeo *d_g, const double *d_q, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sub_from_q, dim3(blocks_for(N, 256)), dim3(256), 0, st, d_g, d_q, N, d_SUB, d_valid);
  return hipGetLastError();
}

hipError_t grs_launch_sub_from_U(const grs_dgeo *d_g, const double *d_U, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sub_from_U, dim3(blocks_for(N, 256)), dim3(256), 0, st, d_g, d_U, N, d_SUB, d_valid);
  return hipGetLastError();
}

hipError_t grs_launch_sub_from_euler(const grs_dgeo *d_g, const double *d_eu, int64_t N, double *d_SUB,
                                     int32_t *d_valid, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sub_from_euler, dim3(blocks_for(N, 256)), dim3(256), 0, st, d_g, d_eu, N, d_SUB, d_valid);
  return hipGetLastError();
}

static size_t fast_lds(int32_t H, int wpb, const grs_eval_args *a) {
  return (size_t)((int64_t)wpb * fast_wave_doubles(H) + fast_tab_doubles(a->map, a->nrot, H, wpb)) * sizeof(double);
}

/* Resident blocks per CU of one k_fast instantiation at wpb waves per block
 * (cached per wpb). */
template <bool MEDIAN, int MAP, bool AUDIT>
static int fast_blocks_per_cu(size_t lds, int wpb) {
  static int cached[GRS_WPB + 1] = {0};
  static size_t cached_lds[GRS_WPB + 1] = {0};
  if (cached[wpb] == 0 || cached_lds[wpb] != lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fast<MEDIAN, MAP, AUDIT>, 64 * wpb, lds) != hipSuccess ||
        nb <= 0)
      nb = 1;
    cached[wpb] = nb;
    cached_lds[wpb] = lds;
  }
  return cached[wpb];
}

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return cus;
}

/* Persistent launch: at most as many waves as can be resident at once. */
template <bool MEDIAN, int MAP, bool AUDIT = false>
static hipError_t launch_fast_impl(const grs_eval_args *a, int32_t H, int64_t n_waves, int wpb, hipStream_t st) {
  if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) return hipErrorInvalidValue;
  /* every wave of the launch needs a scratch slot: fewer slots than a block
   * has waves take smaller blocks */
  while (wpb > 1 && n_waves < wpb) wpb >>= 1;
  size_t lds = fast_lds(H, wpb, a);
  if (lds > 65536) {
    /* above the default dynamic LDS limit of a launch */
    hipError_t e = hipFuncSetAttribute((const void *)k_fast<MEDIAN, MAP, AUDIT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  int64_t resident = (int64_t)device_cus() * fast_blocks_per_cu<MEDIAN, MAP, AUDIT>(lds, wpb);
  int64_t blocks = n_waves / wpb; /* n_waves is a multiple of 4: no wave beyond the scratch slots */
  if (blocks > resident) blocks = resident;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((k_fast<MEDIAN, MAP, AUDIT>), dim3((unsigned)blocks), dim3(64 * wpb), lds, st, *a, H);
  return hipGetLastError();
}

hipError_t grs_launch_fast(const grs_eval_args *a, int32_t H, int64_t n_waves, int32_t wpb, hipStream_t st) {
  if (a->n_eval <= a->e0) return hipSuccess;
  switch (a->map) {
    case GRS_MAP_ONE_VOXEL: return launch_fast_impl<true, GRS_MAP_ONE_VOXEL>(a, H, n_waves, wpb, st);
    case GRS_MAP_ONE_U: return launch_fast_impl<true, GRS_MAP_ONE_U>(a, H, n_waves, wpb, st);
    case GRS_MAP_PAIRS: return launch_fast_impl<true, GRS_MAP_PAIRS>(a, H, n_waves, wpb, st);
    case GRS_MAP_CSR_LIST: return launch_fast_impl<true, GRS_MAP_CSR_LIST>(a, H, n_waves, wpb, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t grs_launch_fast_counts(const grs_eval_args *a, int32_t H, int64_t n_waves, int32_t wpb, hipStream_t st) {
  if (a->n_eval <= a->e0) return hipSuccess;
  if (a->map == GRS_MAP_CSR) return launch_fast_impl<false, GRS_MAP_CSR>(a, H, n_waves, wpb, st);
  if (a->map == GRS_MAP_GRID && a->n_u > 0) return launch_fast_impl<false, GRS_MAP_GRID>(a, H, n_waves, wpb, st);
  return hipErrorInvalidValue;
}

/* One-voxel evaluation that also runs the exact pipeline on every event of
 * phase B1 and counts disagreements (a->audit). */
hipError_t grs_launch_fast_audit(const grs_eval_args *a, int32_t H, int64_t n_waves, hipStream_t st) {
  if (a->n_eval <= a->e0) return hipSuccess;
  if (a->map != GRS_MAP_ONE_VOXEL || !a->audit) return hipErrorInvalidValue;
  return launch_fast_impl<true, GRS_MAP_ONE_VOXEL, true>(a, H, n_waves, GRS_WPB, st);
}

hipError_t grs_launch_pack_hits(const uint32_t *d_keys, int64_t n, uint32_t hit_key, int32_t has_hit,
                                uint32_t *d_bits, hipStream_t st) {
  int64_t nw = (n + 31) >> 5;
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_hits, dim3(blocks_for(nw, 256)), dim3(256), 0, st, d_keys, n, hit_key, has_hit, d_bits);
  return hipGetLastError();
}

hipError_t grs_launch_pack_hits_tiled(const uint32_t *d_keys, int64_t nz, int64_t ny, int64_t nrot, uint32_t hit_key,
                                      int32_t has_hit, uint32_t *d_tbits, hipStream_t st) {
  const int64_t ntz = (nz + 31) >> 5, nty = (ny + 31) >> 5;
  const int64_t nw = ntz * nty * nrot * 32;
  if (nw <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_hits_tiled, dim3(blocks_for(nw, 256)), dim3(256), 0, st, d_keys, nz, ny, nw,
                     (uint32_t)ntz, (uint32_t)nty, hit_key, has_hit, d_tbits);
  return hipGetLastError();
}

hipError_t grs_launch_plain(const grs_eval_args *a, int64_t T, hipStream_t st) {
  if (a->n_eval <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_plain, dim3(blocks_for(T, 256)), dim3(256), 0, st, *a, T);
  return hipGetLastError();
}

hipError_t grs_launch_compete_reduce(const int64_t *d_row_start, int64_t M, const int32_t *d_ns, const int32_t *d_ni,
                                     const double *d_C, const float *d_med, int64_t *d_best, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compete_reduce, dim3(blocks_for(M, 256)), dim3(256), 0, st, d_row_start, M, d_ns, d_ni, d_C,
                     d_med, d_best);
  return hipGetLastError();
}

hipError_t grs_launch_math_eval(int fn, const double *d_x, double *d_y, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_math_eval, dim3(blocks_for(n, 256)), dim3(256), 0, st, fn, d_x, d_y, n);
  return hipGetLastError();
}

hipError_t grs_launch_copy_u32(const uint32_t *d_src, uint32_t *d_dst, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t b = blocks_for(n, 256);
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)b), dim3(256), 0, st, d_src, d_dst, n);
  return hipGetLastError();
}

hipError_t grs_launch_check_keys(const uint32_t *d_keys, int64_t n, uint32_t *d_bad, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t b = blocks_for(n, 256);
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(k_check_keys, dim3((unsigned)b), dim3(256), 0, st, d_keys, n, d_bad);
  return hipGetLastError();
}

int64_t grs_fast_wave_slots(int device) {
  (void)device;
  /* upper bound over the instantiations (at most 32 resident waves per
   * CU): the scratch is indexed by wave */
  return (int64_t)device_cus() * 64;
}

hipError_t grs_launch_fill_u64(unsigned long long *d_p, int64_t n, unsigned long long v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n, 256)), dim3(256), 0, st, d_p, n, v);
  return hipGetLastError();
}

hipError_t grs_launch_winner_list(const unsigned long long *d_row_best, const int64_t *d_row_eoff, int64_t m,
                                  int64_t *d_elist, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_winner_list, dim3(blocks_for(m, 256)), dim3(256), 0, st, d_row_best, d_row_eoff, m, d_elist);
  return hipGetLastError();
}

hipError_t grs_launch_winner_out(const int64_t *d_elist, const int64_t *d_row_eoff, const int64_t *d_row_ptr,
                                 const double *d_cand_q, int32_t nr, double gridsize, int64_t m, int64_t *d_best,
                                 double *d_best_q, hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_winner_out, dim3(blocks_for(m, 256)), dim3(256), 0, st, d_elist, d_row_eoff, d_row_ptr, d_cand_q,
                     nr, gridsize, m, d_best, d_best_q);
  return hipGetLastError();
}
