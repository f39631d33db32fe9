/*
 * grs_gt.hip -- the G-vector matching score ("Gt" indexing method).
 *
 * The reference's cuda_forward_Gt has no source in the repository and
 * ForwardCalc_OR_space_Gt.m:68 calls a calc_Gt_match that is absent; these
 * kernels follow calc_Gt_match_fun.m:1-152 (with calcExpectSpotNr.m:6-47):
 *
 * k_gt_proj   per projection: SamposW = Omega*S*pos (pos not shifted by
 *             RotAxisOffset, :24), the incoming-beam unit vector and the
 *             terms of calcExpectSpotNr with the centre at the effective
 *             detector distance Lsam2det_tr of the tilt (:12-16, :25-31).
 * k_gt_spots  per experimental spot (Spots{i}(:,9:10)): the diffracted-beam
 *             unit vector towards its pixel position in the untilted frame
 *             (:35-48), 2theta, lambda_j = 2 d_j sind(2theta/2) and the unit
 *             vector of Gt_spot_j = (2 pi/lambda_j)(Kout_tr - Kin) for each of
 *             the first four families j (:50-59, :88).
 * k_gt_score  one wavefront per orientation, lanes over reflections: Gt =
 *             Omega*S*U*B*hkl per projection; the expected-spot count
 *             (the forward_comp selection); for reflections of the first
 *             four families (:96-149) the smallest angle acos(n.g)*180/pi to
 *             the projection's spot unit vectors of the family: the largest
 *             dot product is found first and the angle taken as the minimum
 *             over every spot whose dot product is within 1e-12 of it (acos
 *             is monotone up to its last bit), a dot product above 1 giving
 *             0 (the real part MATLAB keeps); a match when the angle is
 *             below 1 degree and |Gt| is within 0.1 of 2 pi/d.  The matched
 *             angles are listed per wavefront in (projection, reflection)
 *             order; their median (rank counting across the wave, even count
 *             a + (b - a)/2) and mean (sequential sum) close the row
 *             [Nr_match, median, Nr_match/NrSpotExpectAll, mean].
 * k_gt_sort   per (projection, family), when every projection has at most
 *             GT_SORT_MAX spots: the keys (z components of the unit vectors,
 *             NaN last, not counted) sorted, the unit vectors packed in that
 *             order, and the start of every z bucket of width 2/GRS_GT_NB.
 *             k_gt_score then visits only the spots with |z - n_z| <= GT_WIN,
 *             found from the bucket below n_z - GT_WIN by a short forward
 *             scan (the same index a binary search gives): a spot within
 *             1.01 degrees of n has |w - n| = 2 sin(angle/2) <= GT_WIN, so
 *             the window holds every spot that can give an angle below 1
 *             degree and every spot within 1e-12 of the best dot product of
 *             such a spot; the row equals the full scan's (without such a
 *             spot neither finds a match).
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

#define GT_PROJ_DOUBLES 18
#define GT_SORT_MAX 4096 /* spots per projection up to which k_gt_sort keeps them in LDS */
#define GT_WIN 0.0177    /* > 2 sin(1.01 deg / 2) = 0.017627 */

/* sind(x): reduction to |r| <= 45 degrees by multiples of 90, then
 * sin or cos of (pi/180)*r with the quadrant's sign */
__device__ __forceinline__ double gt_sind(double x) {
  GRS_NO_CONTRACT
  const double n = grs_round(x / 90.0);
  const double r = x - n * 90.0;
  double m = n - 4.0 * __builtin_floor(n / 4.0);
  const double a = (GRS_PI / 180.0) * r;
  if (m == 0.0) return grs_sin(a);
  if (m == 1.0) return grs_cos(a);
  if (m == 2.0) return -grs_sin(a);
  return -grs_cos(a);
}

/* proj[i]: c s SW0 SW1 SW2 u0 u1 u2 ctr1 ctr2 cu1 cu2 L2 k0 tnum ov1 ov2 (17) */
__global__ void k_gt_proj(const grs_dgeo *g, const double *pos, const double *cs, int32_t nrot, double Ltr,
                          double *proj) {
  GRS_NO_CONTRACT
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrot) return;
  const double c = cs[2 * i], s = cs[2 * i + 1];
  const double *S = g->S;
  double OS[9];
  for (int j = 0; j < 3; j++) {
    OS[0 * 3 + j] = c * S[0 * 3 + j] - s * S[1 * 3 + j];
    OS[1 * 3 + j] = s * S[0 * 3 + j] + c * S[1 * 3 + j];
    OS[2 * 3 + j] = S[2 * 3 + j];
  }
  const double p0 = pos[0], p1 = pos[1], p2 = pos[2];
  const double SW0 = (OS[0] * p0 + OS[1] * p1) + OS[2] * p2;
  const double SW1 = (OS[3] * p0 + OS[4] * p1) + OS[5] * p2;
  const double SW2 = (OS[6] * p0 + OS[7] * p1) + OS[8] * p2;
  const double gp1 = g->Lss + SW0, gp2 = (SW1 - g->P0y) + g->RAO, gp3 = SW2 - g->P0z;
  const double ng = grs_sqrt((gp1 * gp1 + gp2 * gp2) + gp3 * gp3);
  const double ctr1 = (gp2 * Ltr) / gp1, ctr2 = (gp3 * Ltr) / gp1;
  const double alpha = grs_atan(grs_sqrt(gp2 * gp2 + gp3 * gp3) / gp1);
  const double nc = grs_sqrt((Ltr * Ltr + ctr1 * ctr1) + ctr2 * ctr2);
  const double *R = g->R;
  double *o = proj + GT_PROJ_DOUBLES * i;
  o[0] = c;
  o[1] = s;
  o[2] = SW0;
  o[3] = SW1;
  o[4] = SW2;
  o[5] = gp1 / ng;
  o[6] = gp2 / ng;
  o[7] = gp3 / ng;
  o[8] = ctr1;
  o[9] = ctr2;
  o[10] = ctr1 / nc;
  o[11] = ctr2 / nc;
  o[12] = (g->Lsd - SW0) / grs_cos(alpha);
  o[13] = g->Lsd - SW0;
  o[14] = (R[0] * (g->Lsd - SW0) + R[3] * ((g->dety00 - g->RAO) - SW1)) + R[6] * (g->detz00 - SW2);
  o[15] = (SW1 - g->dety00) + g->RAO;
  o[16] = SW2 - g->detz00;
}

/* units[12*s + 3*j + k]: family j's unit vector of spot s */
__global__ void k_gt_spots(const grs_dgeo *g, const double *proj, const int64_t *fptr, int32_t nrot,
                           const double *fsp, int64_t nsp, double Lsd_tr, double dety00_tr, double detz00_tr,
                           const double *dposs, int32_t nfam, double *units) {
  GRS_NO_CONTRACT
  const int64_t sidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nsp) return;
  int32_t lo = 0, hi = nrot;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) >> 1;
    if (fptr[mid] <= sidx) lo = mid;
    else hi = mid;
  }
  const double *P = proj + GT_PROJ_DOUBLES * lo;
  const double SW0 = P[2], SW1 = P[3], SW2 = P[4], k0 = P[5], k1 = P[6], k2 = P[7];
  const double y22 = (g->dety0 - fsp[3 * sidx]) * g->py;
  const double z22 = (g->detz0 - fsp[3 * sidx + 1]) * g->pz;
  const double a = Lsd_tr - SW0, b = (y22 - dety00_tr) - SW1, cc = (z22 - detz00_tr) - SW2;
  const double nk = grs_sqrt((a * a + b * b) + cc * cc);
  const double o0 = a / nk, o1 = b / nk, o2 = cc / nk;
  const double tt = (grs_acos((k0 * o0 + k1 * o1) + k2 * o2) * 180.0) / GRS_PI;
  const double half = gt_sind(tt / 2.0);
  const double f0 = o0 - k0, f1 = o1 - k1, f2 = o2 - k2;
  for (int j = 0; j < 4; j++) {
    double *u = units + 12 * sidx + 3 * j;
    if (j >= nfam) {
      u[0] = u[1] = u[2] = 0.0;
      continue;
    }
    const double lam = (2.0 * dposs[j]) * half;
    const double sc = (2.0 * GRS_PI) / lam;
    const double g0 = sc * f0, g1 = sc * f1, g2 = sc * f2;
    const double n = grs_sqrt((g0 * g0 + g1 * g1) + g2 * g2);
    u[0] = g0 / n;
    u[1] = g1 / n;
    u[2] = g2 / n;
  }
}

/* block (i, f): spots of projection i sorted by units[.., f, 2]; n2 = the
 * power of two >= the largest projection (dynamic LDS n2 * 12 bytes) */
__global__ __launch_bounds__(256) void k_gt_sort(const int64_t *fptr, const double *units, int64_t nsp, int32_t n2,
                                                 double *skey, double *su, int32_t *sorig, int32_t *bst,
                                                 int32_t *slen) {
  extern __shared__ double gsm[];
  double *key = gsm;
  int32_t *idx = (int32_t *)(gsm + n2);
  const int32_t i = blockIdx.x >> 2, f = blockIdx.x & 3;
  const int64_t s0 = fptr[i];
  const int32_t m = (int32_t)(fptr[i + 1] - s0);
  for (int32_t t = threadIdx.x; t < n2; t += blockDim.x) {
    double k = __builtin_inf();
    if (t < m) {
      const double z = units[12 * (s0 + t) + 3 * f + 2];
      k = z == z ? z : __builtin_inf();
    }
    key[t] = k;
    idx[t] = t;
  }
  __syncthreads();
  /* bitonic sort, ascending by (key, idx) */
  for (int32_t size = 2; size <= n2; size <<= 1) {
    for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (int32_t t = threadIdx.x; t < (n2 >> 1); t += blockDim.x) {
        const int32_t lo = 2 * t - (t & (stride - 1));
        const int32_t hi = lo + stride;
        const bool up = (lo & size) == 0;
        const double ka = key[lo], kb = key[hi];
        const int32_t ia = idx[lo], ib = idx[hi];
        const bool gt = ka > kb || (ka == kb && ia > ib);
        if (gt == up) {
          key[lo] = kb;
          key[hi] = ka;
          idx[lo] = ib;
          idx[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  int32_t valid = 0;
  for (int32_t t = threadIdx.x; t < m; t += blockDim.x) {
    const int64_t o = (int64_t)f * nsp + s0 + t;
    skey[o] = key[t];
    const double *w = units + 12 * (s0 + idx[t]) + 3 * f;
    su[3 * o] = w[0];
    su[3 * o + 1] = w[1];
    su[3 * o + 2] = w[2];
    sorig[o] = idx[t];
    valid += key[t] != __builtin_inf() ? 1 : 0;
  }
  /* bucket starts: entry b > 0 is the first key >= -1 + 2b/GRS_GT_NB (the
   * non-finite keys sort last, so it never passes the finite ones), entry 0
   * is 0 */
  for (int32_t b = threadIdx.x; b <= GRS_GT_NB; b += blockDim.x) {
    int32_t lo = 0;
    if (b > 0) {
      const double zb = -1.0 + (2.0 * (double)b) / (double)GRS_GT_NB;
      int32_t hi = n2;
      while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (key[mid] < zb) lo = mid + 1;
        else hi = mid;
      }
    }
    bst[(int64_t)(4 * i + f) * (GRS_GT_NB + 1) + b] = lo;
  }
  /* block sum of the finite keys */
  __shared__ int32_t red[256];
  red[threadIdx.x] = valid;
  __syncthreads();
  for (int32_t o = 128; o > 0; o >>= 1) {
    if ((int32_t)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) slen[4 * i + f] = red[0];
}

__device__ __forceinline__ uint64_t gt_lanes_below(uint64_t m, int lane) {
  return lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
}

/* 4 blocks per CU (128 VGPRs): the kernel waits on the window loads, and
 * 4 waves per SIMD hide more of that than 3 unspilled ones (33 vs 50 ms for
 * 8192 orientations x 124k spots) */
__global__ __launch_bounds__(256, 4) void k_gt_score(const grs_dgeo *g, const double *hkl, const int32_t *famidx,
                                                  const double *glen_poss, const double *SUB, const int32_t *valid,
                                                  int64_t N, const double *proj, const int64_t *fptr,
                                                  const double *units, const double *skey, const double *su,
                                                  const int32_t *sorig, const int32_t *bst, const int32_t *slen,
                                                  int64_t nsp, int32_t cost_mode, double *scratch, int64_t cap,
                                                  double *out) {
  GRS_NO_CONTRACT
  const int32_t H = g->H, nrot = g->nrot;
  const int lane = threadIdx.x & 63;
  const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  double *list = scratch + gwave * cap;
  for (int64_t u = gwave; u < N; u += nwaves) {
    double *o = out + 4 * u;
    if (!valid[u]) {
      if (lane == 0) o[0] = o[1] = o[2] = o[3] = 0.0;
      continue;
    }
    const double *sub = SUB + 9 * u;
    int32_t nm = 0, expect = 0;
    /* calc_Gt_match's cost sums: sum over the matches, in their order, of
     * (n_k - w_k)^2 per component k (ForwardCalc_OR_space_Gt.m:195-197) */
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int32_t i = 0; i < nrot; i++) {
      const double *P = proj + GT_PROJ_DOUBLES * i;
      grs_vw v;
      v.c = P[0];
      v.s = P[1];
      v.SW1 = P[3];
      v.SW2 = P[4];
      v.u0 = P[5];
      v.u1 = P[6];
      v.u2 = P[7];
      v.ctr1 = P[8];
      v.ctr2 = P[9];
      v.cu1 = P[10];
      v.cu2 = P[11];
      v.L2 = P[12];
      v.k0 = P[13];
      v.tnum = P[14];
      v.ov1 = P[15];
      v.ov2 = P[16];
      const int64_t s0 = fptr[i], s1 = fptr[i + 1];
      for (int32_t h0 = 0; h0 < H; h0 += 64) {
        const int32_t h = h0 + lane;
        bool sel = false, match = false;
        double amin = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
        if (h < H) {
          double gw[3], Gt[3], Glen;
          grs_Gw(sub, hkl + 3 * h, gw);
          grs_Gt(v.c, v.s, gw[0], gw[1], gw[2], Gt);
          const double beta = grs_event_beta(&v, Gt, &Glen);
          int32_t dy, dz;
          sel = grs_event_finish(g, &v, Gt, Glen, beta, &dy, &dz) != 0;
          const int32_t f = famidx[h];
          if (f >= 0 && s1 > s0) {
            const double n0 = Gt[0] / Glen, n1 = Gt[1] / Glen, n2 = Gt[2] / Glen;
            /* spots to visit: the z window of the family's sorted list
             * (packed unit vectors su, entries lo .. e - 1, found from the
             * bucket start below zlo), or every spot of the projection */
            const double *W;
            int64_t k0, k1;
            double best = 0.0;
            bool any = false;
            if (skey) {
              const int64_t base = (int64_t)f * nsp + s0;
              const int32_t mlen = slen[4 * i + f];
              const double zlo = n2 - GT_WIN, zhi = n2 + GT_WIN;
              int32_t b = 0;
              if (zlo == zlo) {
                const double q = __builtin_floor((zlo + 1.0) * (0.5 * (double)GRS_GT_NB)) - 1.0;
                b = q < 0.0 ? 0 : (q > (double)GRS_GT_NB ? GRS_GT_NB : (int32_t)q);
              }
              int32_t lo = bst[(int64_t)(4 * i + f) * (GRS_GT_NB + 1) + b];
              W = su + 3 * base;
              /* the key of a finite entry is its packed z (W[3k + 2]) */
              while (lo < mlen && W[3 * lo + 2] < zlo) lo++;
              int32_t k = lo;
              for (; k < mlen && W[3 * k + 2] <= zhi; k++) {
                const double *w = W + 3 * k;
                const double d = (n0 * w[0] + n1 * w[1]) + n2 * w[2];
                if (d == d && (!any || d > best)) {
                  best = d;
                  any = true;
                }
              }
              k0 = lo;
              k1 = k;
            } else {
              W = nullptr;
              k0 = s0;
              k1 = s1;
              for (int64_t k = k0; k < k1; k++) {
                const double *w = units + 12 * k + 3 * f;
                const double d = (n0 * w[0] + n1 * w[1]) + n2 * w[2];
                if (d == d && (!any || d > best)) {
                  best = d;
                  any = true;
                }
              }
            }
            if (any) {
              /* the smallest angle and, as MATLAB's min, the first spot (in
               * the projection's order) that has it: the full scan visits
               * spots in that order, the window carries their indices */
              amin = 1e300;
              int64_t aidx = INT64_MAX;
              const double *wbest = nullptr;
              const double lim = best - 1e-12;
              for (int64_t k = k0; k < k1; k++) {
                const double *w = W ? W + 3 * k : units + 12 * k + 3 * f;
                const double d = (n0 * w[0] + n1 * w[1]) + n2 * w[2];
                if (d >= lim) {
                  const double a = d > 1.0 ? 0.0 : (grs_acos(d) * 180.0) / GRS_PI;
                  const int64_t oi = W ? (int64_t)sorig[(int64_t)f * nsp + s0 + k] : k;
                  if (a < amin || (a == amin && oi < aidx)) {
                    amin = a;
                    aidx = oi;
                    wbest = w;
                  }
                }
              }
              match = amin < 1.0 && grs_fabs(Glen - glen_poss[f]) < 0.1;
              if (match && cost_mode) {
                const double x0 = n0 - wbest[0], x1 = n1 - wbest[1], x2 = n2 - wbest[2];
                e0 = x0 * x0;
                e1 = x1 * x1;
                e2 = x2 * x2;
              }
            }
          }
        }
        expect += __popcll(__ballot(sel));
        const uint64_t mm = __ballot(match);
        if (match) list[nm + __popcll(gt_lanes_below(mm, lane))] = amin;
        nm += __popcll(mm);
        if (cost_mode) {
          /* the matches of this batch in lane order: every lane adds them
           * in sequence (the wave-uniform loop keeps the shuffles whole) */
          for (uint64_t m = mm; m; m &= m - 1) {
            const int l = __ffsll((long long)m) - 1;
            c0 = c0 + __shfl(e0, l);
            c1 = c1 + __shfl(e1, l);
            c2 = c2 + __shfl(e2, l);
          }
        }
      }
    }
    __threadfence_block();
    double med = __builtin_nan(""), mean = __builtin_nan("");
    if (nm > 0) {
      /* ranks with ties broken by position: the elements of rank k1 and k2 */
      const int32_t k1 = (nm & 1) ? (nm >> 1) : (nm >> 1) - 1, k2 = nm >> 1;
      double v1 = 0.0, v2 = 0.0;
      int f1 = 0, f2 = 0;
      for (int32_t j = lane; j < nm; j += 64) {
        const double x = list[j];
        int32_t r = 0;
        for (int32_t t = 0; t < nm; t++) {
          const double y = list[t];
          r += (y < x || (y == x && t < j)) ? 1 : 0;
        }
        if (r == k1) {
          v1 = x;
          f1 = 1;
        }
        if (r == k2) {
          v2 = x;
          f2 = 1;
        }
      }
      const uint64_t b1 = __ballot(f1), b2 = __ballot(f2);
      const int l1 = __ffsll((long long)b1) - 1, l2 = __ffsll((long long)b2) - 1;
      v1 = __shfl(v1, l1);
      v2 = __shfl(v2, l2);
      med = (nm & 1) ? v1 : v1 + (v2 - v1) / 2.0;
      if (lane == 0) {
        double acc = 0.0;
        for (int32_t t = 0; t < nm; t++) acc = acc + list[t];
        mean = acc / (double)nm;
      }
    }
    if (lane == 0) {
      o[0] = (double)nm;
      o[2] = (double)nm / (double)expect;
      if (cost_mode) {
        /* calc_Gt_match: [Nr_match Gt_matched_cost Nr_match/NrSpotExpectAll
         * mean angle], cost 100 without a match (:198, :203) */
        o[1] = nm > 0 ? ((c0 + c1) + c2) / (double)nm : 100.0;
        o[3] = mean;
      } else {
        o[1] = med;
        o[3] = mean;
      }
    }
    __threadfence_block();
  }
}

hipError_t grs_launch_gt_proj(const grs_dgeo *d_g, const double *d_pos, const double *d_cs, int32_t nrot, double Ltr,
                              double *d_proj, hipStream_t st) {
  if (nrot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gt_proj, dim3((unsigned)((nrot + 63) / 64)), dim3(64), 0, st, d_g, d_pos, d_cs, nrot, Ltr,
                     d_proj);
  return hipGetLastError();
}

hipError_t grs_launch_gt_spots(const grs_dgeo *d_g, const double *d_proj, const int64_t *d_fptr, int32_t nrot,
                               const double *d_fsp, int64_t nsp, double Lsd_tr, double dety00_tr, double detz00_tr,
                               const double *d_dposs, int32_t nfam, double *d_units, hipStream_t st) {
  if (nsp <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gt_spots, dim3((unsigned)((nsp + 255) / 256)), dim3(256), 0, st, d_g, d_proj, d_fptr, nrot,
                     d_fsp, nsp, Lsd_tr, dety00_tr, detz00_tr, d_dposs, nfam, d_units);
  return hipGetLastError();
}

int64_t grs_gt_wave_slots(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
  return (int64_t)cus * 16;
}

hipError_t grs_launch_gt_sort(const int64_t *d_fptr, int32_t nrot, const double *d_units, int64_t nsp,
                              int64_t max_frame, double *d_skey, double *d_su, int32_t *d_sorig, int32_t *d_bst,
                              int32_t *d_slen, hipStream_t st) {
  if (nrot <= 0) return hipSuccess;
  if (max_frame > GT_SORT_MAX) return hipErrorInvalidValue;
  int32_t n2 = 64;
  while (n2 < max_frame) n2 <<= 1;
  hipLaunchKernelGGL(k_gt_sort, dim3((unsigned)(4 * nrot)), dim3(256), (size_t)n2 * 12, st, d_fptr, d_units, nsp, n2,
                     d_skey, d_su, d_sorig, d_bst, d_slen);
  return hipGetLastError();
}

int64_t grs_gt_sort_max() { return GT_SORT_MAX; }

hipError_t grs_launch_gt_score(const grs_dgeo *d_g, const double *d_hkl, const int32_t *d_famidx,
                               const double *d_glen_poss, const double *d_SUB, const int32_t *d_valid, int64_t N,
                               const double *d_proj, const int64_t *d_fptr, const double *d_units,
                               const double *d_skey, const double *d_su, const int32_t *d_sorig,
                               const int32_t *d_bst, const int32_t *d_slen, int64_t nsp, int32_t cost_mode,
                               double *d_scratch, int64_t cap, int64_t waves, double *d_out, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  int64_t w = waves < N ? waves : N;
  const int64_t blocks = (w + 3) / 4;
  hipLaunchKernelGGL(k_gt_score, dim3((unsigned)blocks), dim3(256), 0, st, d_g, d_hkl, d_famidx, d_glen_poss, d_SUB,
                     d_valid, N, d_proj, d_fptr, d_units, d_skey, d_su, d_sorig, d_bst, d_slen, nsp, cost_mode, d_scratch,
                     cap, d_out);
  return hipGetLastError();
}
