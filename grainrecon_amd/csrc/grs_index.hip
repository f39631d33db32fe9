/*
 * grs_index.hip -- device steps of seed indexing between the scoring launches.
 *
 * GrainIndexing_engine_gpu_comp.m:46-66 with
 * fit_OR_constrain_with_local_comp_cuda.m:19-116 for a batch of seed
 * positions, with no host round trip between the steps:
 *
 * k_topk           sortrows(Output,3,'descend') followed by rows 1:Nr_seed,
 *                  as a top-k of unique 64-bit sort keys per segment (one
 *                  block per seed): an 8-bit MSB radix select over the keys
 *                  (histograms in LDS) finds the k largest, which are then
 *                  compacted into LDS and sorted descending by a bitonic
 *                  network.
 * k_gather_centres index_seeds(j,5:8): the orientation-set rows of the seeds.
 * k_ls_grids       ori_local_sampling.m:22-24, the three colon grids of every
 *                  seed (k_local_sampling then builds the rows).
 * k_qall_owner,    fit_OR_constrain_with_local_comp_cuda.m:19-26: the rows
 * k_qall_build     of seed j occupy q_all rows (j-1)*len_j+1 .. j*len_j where
 *                  len_j is that seed's own row count; a later seed overwrites
 *                  an earlier one where the ranges overlap, rows no seed wrote
 *                  stay zero (group 0), and q_all has max(Nr_seed*len_1,
 *                  max_j j*len_j) rows.  The owner of a row is the largest j
 *                  whose range covers it.
 * k_group_max      :100-111: for every seed j present in column 9, the first
 *                  maximum of C (NaN ignored) over rows
 *                  1+nr_per_OR*(j-1) .. nr_per_OR*j, where nr_per_OR is the
 *                  row count of the LAST seed; its sort key for
 *                  sortrows(Output_max,3,'descend').
 * k_fit_starts     :116 x0 = Output_max(j,6:8) and :163 column 9 of the rows
 *                  chosen for fitting.
 * k_seed_out       the index_seeds rows as orientation indices and C.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

#define GRS_CLS_NAN (1ull << 31)

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

/* ---------------------------------------------------------------------- */
/* k_topk: the k largest of n unique keys per segment, sorted descending    */

__global__ __launch_bounds__(1024) void k_topk(const unsigned long long *keys, int64_t n, int64_t seg_stride,
                                               int32_t k, unsigned long long *out) {
  __shared__ uint32_t hist[256];
  __shared__ unsigned long long sel[GRS_TOPK_MAX];
  __shared__ unsigned long long s_prefix, s_mask;
  __shared__ int32_t s_need, s_done, s_cnt;
  const int tid = threadIdx.x, bd = blockDim.x;
  const unsigned long long *kp = keys + (int64_t)blockIdx.x * seg_stride;
  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_need = k;
    s_done = 0;
    s_cnt = 0;
  }
  __syncthreads();
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += bd) hist[i] = 0;
    __syncthreads();
    const unsigned long long pre = s_prefix, msk = s_mask;
    for (int64_t i = tid; i < n; i += bd) {
      const unsigned long long v = kp[i];
      if ((v & msk) == pre) atomicAdd(&hist[(v >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      const uint32_t need = (uint32_t)s_need;
      uint32_t cum = 0;
      int b = 255;
      for (; b > 0; b--) {
        if (cum + hist[b] >= need) break;
        cum += hist[b];
      }
      const uint32_t rest = need - cum;
      s_need = (int32_t)rest;
      s_prefix = pre | ((unsigned long long)b << shift);
      s_mask = msk | (0xFFull << shift);
      s_done = hist[b] == rest ? 1 : 0;
    }
    __syncthreads();
    if (s_done) break;
  }
  int32_t P2 = 1;
  while (P2 < k) P2 <<= 1;
  const unsigned long long pre = s_prefix, msk = s_mask;
  for (int64_t i = tid; i < n; i += bd) {
    const unsigned long long v = kp[i];
    if ((v & msk) >= pre) {
      const int32_t pos = atomicAdd(&s_cnt, 1);
      if (pos < P2) sel[pos] = v;
    }
  }
  __syncthreads();
  const int32_t cnt = s_cnt < P2 ? s_cnt : P2;
  for (int32_t i = cnt + tid; i < P2; i += bd) sel[i] = 0ull;
  __syncthreads();
  for (int32_t sz = 2; sz <= P2; sz <<= 1) {
    for (int32_t j = sz >> 1; j > 0; j >>= 1) {
      for (int32_t i = tid; i < P2; i += bd) {
        const int32_t l = i ^ j;
        if (l > i) {
          const unsigned long long x = sel[i], y = sel[l];
          const bool desc = (i & sz) == 0;
          if (desc ? (x < y) : (x > y)) {
            sel[i] = y;
            sel[l] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int32_t i = tid; i < k; i += bd) out[(int64_t)blockIdx.x * k + i] = sel[i];
}

/* ---------------------------------------------------------------------- */
/* seeds -> local sampling rows                                             */

__device__ __forceinline__ int64_t key_index(unsigned long long key) {
  return (int64_t)(0xFFFFFFFFull - (key & 0xFFFFFFFFull));
}

__global__ void k_gather_centres(const double *set_q, const unsigned long long *topkey, int64_t n, double *cen) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t u = key_index(topkey[i]);
#pragma unroll
  for (int c = 0; c < 4; c++) cen[4 * i + c] = set_q[4 * u + c];
}

__global__ void k_ls_grids(const double *cen, int64_t n, int32_t nr, double gridsize, int32_t cap_axis,
                           double *grids, int32_t *gn, uint32_t *bad) {
  GRS_NO_CONTRACT
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n) return;
  const int64_t c = t / 3;
  const int axis = (int)(t - 3 * c);
  const double cj = cen[4 * c + 1 + axis];
  const double half = (gridsize * (double)nr) / 2.0;
  int64_t m = grs_colon(cj - half, gridsize, cj + half, grids + (3 * c + axis) * GRS_LS_MAX_GRID, cap_axis);
  if (m < 0) {
    atomicOr(bad, 2u);
    m = 0;
  }
  gn[3 * c + axis] = (int32_t)m;
}

__global__ void k_qall_owner(const int64_t *ls_n, int64_t P, int64_t K, int64_t cap, int64_t Rcap, int32_t *owner,
                             unsigned long long *Rmax) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * K * cap) return;
  const int64_t p = t / (K * cap);
  const int64_t rem = t - p * K * cap;
  const int64_t j = rem / cap, r = rem - j * cap;
  const int64_t len = ls_n[p * K + j];
  if (r < len) atomicMax(owner + p * Rcap + j * len + r, (int32_t)(j + 1));
  if (r == 0) {
    atomicMax(Rmax + p, (unsigned long long)((j + 1) * len));
    if (j == 0) atomicMax(Rmax + p, (unsigned long long)(K * len));
  }
}

__global__ void k_qall_build(const double *ls_q, const int64_t *ls_n, const int32_t *owner,
                             const unsigned long long *Rmax, int64_t P, int64_t K, int64_t cap, int64_t Rcap,
                             double *q_rows, uint32_t *gcount) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t grp = -1; /* the group row t is counted in, or -1 */
  if (t < P * Rcap) {
    const int64_t p = t / Rcap, row = t - p * Rcap;
    double *q = q_rows + 4 * t;
    const int32_t j = owner[t];
    if (row >= (int64_t)Rmax[p]) {
      const double nan = __builtin_nan("");
      q[0] = q[1] = q[2] = q[3] = nan;
    } else if (j == 0) {
      q[0] = q[1] = q[2] = q[3] = 0.0;
    } else {
      const int64_t len = ls_n[p * K + j - 1];
      const double *src = ls_q + ((p * K + j - 1) * cap + (row - (int64_t)(j - 1) * len)) * 4;
      q[0] = src[0];
      q[1] = src[1];
      q[2] = src[2];
      q[3] = src[3];
      grp = p * K + j - 1;
    }
  }
  /* the row counts: one atomic per distinct group among the wave's rows
   * (a group's rows are contiguous, so a wave holds one or two) */
  const int lane = (int)(threadIdx.x & 63);
  uint64_t todo = __ballot(grp >= 0);
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int64_t g = __shfl(grp, lead, 64);
    const uint64_t same = __ballot(grp == g);
    if (lane == lead) atomicAdd(gcount + g, (uint32_t)__popcll(same));
    todo &= ~same;
  }
}

/* one wavefront per (seed position, group) */
__global__ __launch_bounds__(256) void k_group_max(const unsigned long long *key2, const int64_t *ls_n,
                                                   const uint32_t *gcount, int64_t P, int64_t K, int64_t Rcap,
                                                   unsigned long long *gkey, int64_t *grow) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= P * K) return;
  const int64_t p = w / K, j = w - p * K;
  const int64_t nper = ls_n[p * K + K - 1];
  const int64_t base = j * nper;
  unsigned long long cls = 1ull;
  int64_t row = -1;
  if (gcount[w] > 0) {
    /* (class << 32) | (2^32 - 1 - r): the maximum is the largest C, first row */
    unsigned long long best = 0ull;
    for (int64_t r = lane; r < nper; r += 64) {
      const unsigned long long c = key2[p * Rcap + base + r] >> 32;
      if (c == GRS_CLS_NAN) continue;
      const unsigned long long v = (c << 32) | (0xFFFFFFFFull - (unsigned long long)r);
      if (v > best) best = v;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)(best & 0xFFFFFFFFull), off);
      const uint32_t hi = __shfl_xor((uint32_t)(best >> 32), off);
      const unsigned long long o = ((unsigned long long)hi << 32) | lo;
      if (o > best) best = o;
    }
    if (best == 0ull) {
      /* every C of the block is NaN: max returns NaN at the first row */
      cls = GRS_CLS_NAN;
      row = base;
    } else {
      cls = best >> 32;
      row = base + (int64_t)(0xFFFFFFFFull - (best & 0xFFFFFFFFull));
    }
  }
  if (lane == 0) {
    gkey[w] = (cls << 32) | (0xFFFFFFFFull - (unsigned long long)j);
    grow[w] = row;
  }
}

__global__ void k_fit_starts(const unsigned long long *fitkey, const int64_t *grow, const double *q_rows,
                             const int32_t *owner, const int32_t *ns2, const int32_t *ni2, int64_t P, int64_t K,
                             int32_t n_fit, int64_t Rcap, double *x0, int32_t *fgroup, int32_t *fns, int32_t *fni) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * n_fit) return;
  const int64_t p = t / n_fit;
  const int64_t j = key_index(fitkey[t]);
  const int64_t row = grow[p * K + j];
  if (row < 0) {
    x0[3 * t] = x0[3 * t + 1] = x0[3 * t + 2] = 0.0;
    fgroup[t] = 0;
    fns[t] = 0;
    fni[t] = 0;
    return;
  }
  const int64_t r = p * Rcap + row;
  x0[3 * t] = q_rows[4 * r + 1];
  x0[3 * t + 1] = q_rows[4 * r + 2];
  x0[3 * t + 2] = q_rows[4 * r + 3];
  fgroup[t] = owner[r];
  fns[t] = ns2[r];
  fni[t] = ni2[r];
}

__global__ void k_seed_out(const unsigned long long *topkey, const int32_t *ns1, const int32_t *ni1,
                           const int32_t *set_valid, int64_t P, int64_t K, int64_t N, int32_t *seed_idx,
                           double *seed_C) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * K) return;
  const int64_t p = t / K;
  const int64_t u = key_index(topkey[t]);
  seed_idx[t] = (int32_t)u;
  const int32_t ns = ns1[p * N + u], ni = ni1[p * N + u];
  seed_C[t] = set_valid[u] ? (double)ni / (double)ns : 0.0;
}

/* ---------------------------------------------------------------------- */
/* launchers                                                               */

static unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t grs_launch_fill_u32(uint32_t *d_p, int64_t n, uint32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, d_p, n, v);
  return hipGetLastError();
}

hipError_t grs_launch_topk(const unsigned long long *d_keys, int64_t segs, int64_t n, int64_t seg_stride, int32_t k,
                           unsigned long long *d_out, hipStream_t st) {
  if (segs <= 0 || k <= 0) return hipSuccess;
  if (k > GRS_TOPK_MAX || k > n || n > 0xFFFFFFFFll) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_topk, dim3((unsigned)segs), dim3(1024), 0, st, d_keys, n, seg_stride, k, d_out);
  return hipGetLastError();
}

hipError_t grs_launch_gather_centres(const double *d_set_q, const unsigned long long *d_topkey, int64_t n,
                                     double *d_cen, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_centres, dim3(grid_for(n, 256)), dim3(256), 0, st, d_set_q, d_topkey, n, d_cen);
  return hipGetLastError();
}

hipError_t grs_launch_ls_grids(const double *d_cen, int64_t n, int32_t nr, double gridsize, int32_t cap_axis,
                               double *d_grids, int32_t *d_gn, uint32_t *d_bad, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (cap_axis < 1 || cap_axis > GRS_LS_MAX_GRID) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ls_grids, dim3(grid_for(3 * n, 256)), dim3(256), 0, st, d_cen, n, nr, gridsize, cap_axis,
                     d_grids, d_gn, d_bad);
  return hipGetLastError();
}

hipError_t grs_launch_qall(const double *d_ls_q, const int64_t *d_ls_n, int64_t P, int64_t K, int64_t cap,
                           int32_t *d_owner, unsigned long long *d_Rmax, double *d_q_rows, uint32_t *d_gcount,
                           hipStream_t st) {
  if (P <= 0 || K <= 0) return hipSuccess;
  const int64_t Rcap = K * cap;
  hipLaunchKernelGGL(k_qall_owner, dim3(grid_for(P * K * cap, 256)), dim3(256), 0, st, d_ls_n, P, K, cap, Rcap,
                     d_owner, d_Rmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_qall_build, dim3(grid_for(P * Rcap, 256)), dim3(256), 0, st, d_ls_q, d_ls_n, d_owner, d_Rmax,
                     P, K, cap, Rcap, d_q_rows, d_gcount);
  return hipGetLastError();
}

hipError_t grs_launch_group_max(const unsigned long long *d_key2, const int64_t *d_ls_n, const uint32_t *d_gcount,
                                int64_t P, int64_t K, int64_t Rcap, unsigned long long *d_gkey, int64_t *d_grow,
                                hipStream_t st) {
  if (P <= 0 || K <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_max, dim3(grid_for(P * K, 4)), dim3(256), 0, st, d_key2, d_ls_n, d_gcount, P, K, Rcap,
                     d_gkey, d_grow);
  return hipGetLastError();
}

hipError_t grs_launch_fit_starts(const unsigned long long *d_fitkey, const int64_t *d_grow, const double *d_q_rows,
                                 const int32_t *d_owner, const int32_t *d_ns2, const int32_t *d_ni2, int64_t P,
                                 int64_t K, int32_t n_fit, int64_t Rcap, double *d_x0, int32_t *d_fgroup,
                                 int32_t *d_fns, int32_t *d_fni, hipStream_t st) {
  if (P <= 0 || n_fit <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fit_starts, dim3(grid_for(P * n_fit, 256)), dim3(256), 0, st, d_fitkey, d_grow, d_q_rows,
                     d_owner, d_ns2, d_ni2, P, K, n_fit, Rcap, d_x0, d_fgroup, d_fns, d_fni);
  return hipGetLastError();
}

hipError_t grs_launch_seed_out(const unsigned long long *d_topkey, const int32_t *d_ns1, const int32_t *d_ni1,
                               const int32_t *d_set_valid, int64_t P, int64_t K, int64_t N, int32_t *d_seed_idx,
                               double *d_seed_C, hipStream_t st) {
  if (P <= 0 || K <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_seed_out, dim3(grid_for(P * K, 256)), dim3(256), 0, st, d_topkey, d_ns1, d_ni1, d_set_valid,
                     P, K, N, d_seed_idx, d_seed_C);
  return hipGetLastError();
}
