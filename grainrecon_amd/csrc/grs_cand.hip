/*
 * grs_cand.hip -- candidate generation on the GPU.
 *
 * k_local_sampling   ori_local_sampling.m:12-34: one block per centre builds
 *                    the centre row and the grid rows [sqrt(1-(b^2+c^2+d^2))
 *                    b c d] in LDS, sorts their indices with a bitonic network
 *                    (rows compared column by column, unique(q,'rows')
 *                    order), drops exact duplicates and writes the rows
 *                    compacted.  The (q1, q2, q3) grids come from the host
 *                    (MATLAB's colon operator, a few values per centre).
 * k_find_neighbors   find_neighbor_ORs.m:37-96: one wavefront per voxel scans
 *                    its search box and sets one bit per positive grain id in
 *                    an LDS bitset; an empty box is grown by 2 voxels and
 *                    scanned again; the first nr_max set bits (ascending ids,
 *                    the order of unique) are written out.
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

/* ---------------------------------------------------------------------- */
/* k_local_sampling                                                        */

struct ls_shared {
  double q0[GRS_LS_MAX_ROWS];
  int32_t idx[GRS_LS_MAX_ROWS];
  double grid[3][GRS_LS_MAX_GRID];
  int32_t part[256];
};

__device__ __forceinline__ void ls_row(const ls_shared &sh, const double *cen, int32_t n2, int32_t n3, int32_t r,
                                       double *q) {
  if (r == 0) {
    q[0] = cen[0];
    q[1] = cen[1];
    q[2] = cen[2];
    q[3] = cen[3];
    return;
  }
  const int32_t t = r - 1;
  q[0] = sh.q0[r];
  q[1] = sh.grid[0][t / (n2 * n3)];
  q[2] = sh.grid[1][(t / n3) % n2];
  q[3] = sh.grid[2][t % n3];
}

/* row a before row b: columns compared in order, then the original index */
__device__ __forceinline__ bool ls_less(const ls_shared &sh, const double *cen, int32_t n2, int32_t n3, int32_t R,
                                        int32_t a, int32_t b) {
  if (a >= R) return false;
  if (b >= R) return true;
  double qa[4], qb[4];
  ls_row(sh, cen, n2, n3, a, qa);
  ls_row(sh, cen, n2, n3, b, qb);
#pragma unroll
  for (int c = 0; c < 4; c++) {
    if (qa[c] < qb[c]) return true;
    if (qa[c] > qb[c]) return false;
  }
  return a < b;
}

__global__ __launch_bounds__(256) void k_local_sampling(const double *cen_all, const double *grids, const int32_t *gn,
                                                        int64_t K, double *out, int64_t cap, int64_t *out_n,
                                                        uint32_t *bad) {
  __shared__ ls_shared sh;
  const int tid = threadIdx.x;
  for (int64_t c = blockIdx.x; c < K; c += gridDim.x) {
    const double *cen = cen_all + 4 * c;
    const int32_t n1 = gn[3 * c], n2 = gn[3 * c + 1], n3 = gn[3 * c + 2];
    const int32_t R = 1 + n1 * n2 * n3;
    int32_t P = 1;
    while (P < R) P <<= 1;
    for (int i = tid; i < 3 * GRS_LS_MAX_GRID; i += blockDim.x)
      sh.grid[i / GRS_LS_MAX_GRID][i % GRS_LS_MAX_GRID] = grids[(3 * c) * GRS_LS_MAX_GRID + i];
    __syncthreads();
    for (int32_t r = tid; r < P; r += blockDim.x) {
      sh.idx[r] = r;
      if (r >= 1 && r < R) {
        const int32_t t = r - 1;
        const double b = sh.grid[0][t / (n2 * n3)], cc = sh.grid[1][(t / n3) % n2], d = sh.grid[2][t % n3];
        const double ss = (b * b + cc * cc) + d * d;
        if (ss > 1.0) atomicOr(bad, 1u);
        sh.q0[r] = __builtin_sqrt(1.0 - ss);
      }
    }
    __syncthreads();
    for (int32_t k = 2; k <= P; k <<= 1) {
      for (int32_t j = k >> 1; j > 0; j >>= 1) {
        for (int32_t i = tid; i < P; i += blockDim.x) {
          const int32_t l = i ^ j;
          if (l > i) {
            const int32_t a = sh.idx[i], b = sh.idx[l];
            const bool asc = (i & k) == 0;
            const bool sw = asc ? ls_less(sh, cen, n2, n3, R, b, a) : ls_less(sh, cen, n2, n3, R, a, b);
            if (sw) {
              sh.idx[i] = b;
              sh.idx[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    /* unique: keep the first row of every run of equal rows */
    const int32_t per = (R + blockDim.x - 1) / blockDim.x;
    const int32_t p0 = tid * per;
    int32_t cnt = 0;
    for (int32_t p = p0; p < p0 + per && p < R; p++) {
      bool keep = p == 0;
      if (!keep) {
        double qa[4], qb[4];
        ls_row(sh, cen, n2, n3, sh.idx[p], qa);
        ls_row(sh, cen, n2, n3, sh.idx[p - 1], qb);
        keep = !(qa[0] == qb[0] && qa[1] == qb[1] && qa[2] == qb[2] && qa[3] == qb[3]);
      }
      cnt += keep ? 1 : 0;
    }
    sh.part[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int32_t v = tid >= off ? sh.part[tid - off] : 0;
      __syncthreads();
      sh.part[tid] += v;
      __syncthreads();
    }
    int32_t pos = sh.part[tid] - cnt;
    for (int32_t p = p0; p < p0 + per && p < R; p++) {
      double qa[4];
      ls_row(sh, cen, n2, n3, sh.idx[p], qa);
      bool keep = p == 0;
      if (!keep) {
        double qb[4];
        ls_row(sh, cen, n2, n3, sh.idx[p - 1], qb);
        keep = !(qa[0] == qb[0] && qa[1] == qb[1] && qa[2] == qb[2] && qa[3] == qb[3]);
      }
      if (keep) {
        if (pos < cap) {
          double *o = out + (c * cap + pos) * 4;
          o[0] = qa[0];
          o[1] = qa[1];
          o[2] = qa[2];
          o[3] = qa[3];
        }
        pos++;
      }
    }
    if (tid == (int)blockDim.x - 1) out_n[c] = sh.part[tid];
    __syncthreads();
  }
}

/* ---------------------------------------------------------------------- */
/* k_find_neighbors                                                        */

__global__ __launch_bounds__(64) void k_find_neighbors(const int32_t *gid, int64_t nx, int64_t ny, int64_t nz,
                                                       const int32_t *vox, int64_t M, int32_t radius, int32_t nr_max,
                                                       int32_t W, int32_t *ids, int32_t *nout, uint32_t *bad) {
  extern __shared__ uint32_t bits[];
  const int lane = threadIdx.x;
  const int64_t dims[3] = {nx, ny, nz};
  for (int64_t m = blockIdx.x; m < M; m += gridDim.x) {
    int64_t lo[3], hi[3];
    bool found = false;
    for (int64_t tries = 0;; tries++) {
      const int64_t r = (int64_t)radius + 2 * tries;
      bool whole = true;
      for (int j = 0; j < 3; j++) {
        const int64_t p = vox[3 * m + j];
        lo[j] = p - r > 0 ? p - r : 1;
        hi[j] = p + r < dims[j] ? p + r : dims[j];
        whole = whole && lo[j] == 1 && hi[j] == dims[j];
      }
      for (int32_t w = lane; w < W; w += 64) bits[w] = 0u;
      __syncthreads();
      const int64_t bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1;
      const int64_t plane = bx * by;
      bool any = false;
      for (int64_t z = lo[2]; z <= hi[2]; z++) {
        const int32_t *gz = gid + (z - 1) * nx * ny;
        for (int64_t t = lane; t < plane; t += 64) {
          const int64_t iy = t / bx, ix = t - iy * bx;
          const int32_t id = gz[(lo[0] - 1 + ix) + (lo[1] - 1 + iy) * nx];
          if (id > 0) {
            atomicOr(&bits[id >> 5], 1u << (id & 31));
            any = true;
          }
        }
      }
      __syncthreads();
      found = __ballot(any) != 0;
      if (found) break;
      if (whole) break; /* no positive id anywhere: reported below */
    }
    if (!found) {
      if (lane == 0) {
        atomicOr(bad, 1u);
        nout[m] = 0;
      }
      __syncthreads();
      continue;
    }
    /* the first nr_max set bits in ascending order */
    int32_t base = 0;
    for (int32_t w0 = 0; w0 < W && base < nr_max; w0 += 64) {
      const int32_t w = w0 + lane;
      const uint32_t word = w < W ? bits[w] : 0u;
      const int32_t pc = __popc(word);
      int32_t incl = pc;
      for (int off = 1; off < 64; off <<= 1) {
        int32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
      }
      int32_t pos = base + incl - pc;
      uint32_t rest = word;
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        if (pos < nr_max) ids[m * nr_max + pos] = w * 32 + b;
        pos++;
      }
      base += __shfl(incl, 63, 64);
    }
    if (lane == 0) nout[m] = base < nr_max ? base : nr_max;
    __syncthreads();
  }
}

/* ---------------------------------------------------------------------- */
/* launchers                                                               */

hipError_t grs_launch_local_sampling(const double *d_cen, const double *d_grids, const int32_t *d_gn, int64_t K,
                                     double *d_out, int64_t cap, int64_t *d_n, uint32_t *d_bad, hipStream_t st) {
  if (K <= 0) return hipSuccess;
  int64_t blocks = K < 65536 ? K : 65536;
  hipLaunchKernelGGL(k_local_sampling, dim3((unsigned)blocks), dim3(256), 0, st, d_cen, d_grids, d_gn, K, d_out, cap,
                     d_n, d_bad);
  return hipGetLastError();
}

hipError_t grs_launch_find_neighbors(const int32_t *d_gid, int64_t nx, int64_t ny, int64_t nz, const int32_t *d_vox,
                                     int64_t M, int32_t radius, int32_t nr_max, int32_t W, int32_t *d_ids,
                                     int32_t *d_nout, uint32_t *d_bad, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  int64_t blocks = M < 65536 ? M : 65536;
  hipLaunchKernelGGL(k_find_neighbors, dim3((unsigned)blocks), dim3(64), sizeof(uint32_t) * (size_t)W, st, d_gid, nx,
                     ny, nz, d_vox, M, radius, nr_max, W, d_ids, d_nout, d_bad);
  return hipGetLastError();
}
