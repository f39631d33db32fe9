/*
 * grs_region.hip -- region growth around an indexed seed with the volumes
 * resident on the device (grow_indexed_region_cuda.m:29-151).
 *
 * k_region_list    :44-77: the voxels of the bounding box within Box_L of
 *                  its centre point, with their completeness (the seed's
 *                  replaced by the indexed completeness), mask and median
 *                  distance gathered from the volumes; only voxels inside the
 *                  mask are listed (the operator scores no others,
 *                  grow_indexed_region_parallel.m:61-90).
 * k_region_accept  the growth rule of cuda_grow_indexed_region
 *                  (grow_indexed_region_parallel.m:78-81) on the scores of
 *                  the listed voxels, and the write-back of :117-128
 *                  (completeness = Nr_intersect/Nr_simu, median distance,
 *                  Nr_intersect) into the volumes; the accepted voxels and
 *                  their completeness are appended to a list for the
 *                  weighted centre, which the host sums in MATLAB's find()
 *                  order.
 * Volumes are nx x ny x nz doubles with x fastest (MATLAB's I(x,y,z)).
 */
#include <hip/hip_runtime.h>

#include "grs_internal.h"

__global__ void k_region_list(const double *I, const double *mask, const double *dm, int64_t nx, int64_t ny,
                              int64_t lo0, int64_t lo1, int64_t lo2, int64_t bx, int64_t by, int64_t bz, double cp0,
                              double cp1, double cp2, double boxL, int64_t seed_lin, double seed_comp, double *rows,
                              uint32_t *counts) {
  GRS_NO_CONTRACT
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bx * by * bz) return;
  const int64_t ix = t % bx, iy = (t / bx) % by, iz = t / (bx * by);
  const double x = (double)(lo0 + ix), y = (double)(lo1 + iy), z = (double)(lo2 + iz);
  const double dx = x - cp0, dy = y - cp1, dz = z - cp2;
  const double d = grs_sqrt((dx * dx + dy * dy) + dz * dz);
  if (!(d <= boxL)) return;
  atomicAdd(counts + 1, 1u);
  const int64_t lin = (lo0 + ix - 1) + nx * ((lo1 + iy - 1) + ny * (lo2 + iz - 1));
  if (!(mask[lin] > 0.0)) return;
  const uint32_t k = atomicAdd(counts, 1u);
  double *r = rows + 6 * (int64_t)k;
  r[0] = x;
  r[1] = y;
  r[2] = z;
  r[3] = lin == seed_lin ? seed_comp : I[lin];
  r[4] = mask[lin];
  r[5] = dm[lin];
}

__global__ void k_region_accept(const double *rows, int64_t M, const int32_t *ns, const int32_t *ni, const double *C,
                                const float *med, const double *p1, int64_t nx, int64_t ny, double *vC, double *vDM,
                                double *vNI, int64_t *acc_lin, double *acc_C, uint32_t *count) {
  GRS_NO_CONTRACT
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const double *v = rows + 6 * i;
  const double Cn = C[i], dmv = (double)med[i];
  /* min([DisMedian maxDmedian]) with MATLAB's NaN rule */
  const double a = v[5], b = p1[13];
  const double lim = a != a ? b : (b != b ? a : (a < b ? a : b));
  if (!((Cn > v[3] && dmv <= lim) && (Cn >= p1[11] * (1.0 - p1[12])))) return;
  const int64_t lin = ((int64_t)v[0] - 1) + nx * (((int64_t)v[1] - 1) + ny * ((int64_t)v[2] - 1));
  const double Cw = (double)ni[i] / (double)ns[i];
  vC[lin] = Cw;
  vDM[lin] = dmv;
  vNI[lin] = (double)ni[i];
  const uint32_t k = atomicAdd(count, 1u);
  acc_lin[k] = lin;
  acc_C[k] = Cw;
}

hipError_t grs_launch_region_list(const double *d_I, const double *d_mask, const double *d_dm, int64_t nx, int64_t ny,
                                  const int64_t *lo, const int64_t *bdim, const double *cp, double boxL,
                                  int64_t seed_lin, double seed_comp, double *d_rows, uint32_t *d_counts,
                                  hipStream_t st) {
  const int64_t n = bdim[0] * bdim[1] * bdim[2];
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_region_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_I, d_mask, d_dm, nx, ny,
                     lo[0], lo[1], lo[2], bdim[0], bdim[1], bdim[2], cp[0], cp[1], cp[2], boxL, seed_lin, seed_comp,
                     d_rows, d_counts);
  return hipGetLastError();
}

hipError_t grs_launch_region_accept(const double *d_rows, int64_t M, const int32_t *d_ns, const int32_t *d_ni,
                                    const double *d_C, const float *d_med, const double *d_p1, int64_t nx, int64_t ny,
                                    double *d_vC, double *d_vDM, double *d_vNI, int64_t *d_acc_lin, double *d_acc_C,
                                    uint32_t *d_count, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_region_accept, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, d_rows, M, d_ns, d_ni,
                     d_C, d_med, d_p1, nx, ny, d_vC, d_vDM, d_vNI, d_acc_lin, d_acc_C, d_count);
  return hipGetLastError();
}

/* out[i] = v[lin[i]] (lin checked by the caller) */
__global__ void k_gather_f64(const double *v, const int64_t *lin, int64_t n, double *out) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) out[i] = v[lin[i]];
}

hipError_t grs_launch_gather_f64(const double *d_v, const int64_t *d_lin, int64_t n, double *d_out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  hipLaunchKernelGGL(k_gather_f64, dim3((unsigned)blocks), dim3(256), 0, st, d_v, d_lin, n, d_out);
  return hipGetLastError();
}

/* v[lin[i]] = val[i] (lin checked by the caller, no index repeated) */
__global__ void k_scatter_f64(double *v, const int64_t *lin, const double *val, int64_t n) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) v[lin[i]] = val[i];
}

hipError_t grs_launch_scatter_f64(double *d_v, const int64_t *d_lin, const double *d_val, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = (n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536;
  hipLaunchKernelGGL(k_scatter_f64, dim3((unsigned)blocks), dim3(256), 0, st, d_v, d_lin, d_val, n);
  return hipGetLastError();
}
