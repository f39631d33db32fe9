/* Internal declarations shared by the kernels and the C-ABI layer. */
#ifndef GRS_INTERNAL_H
#define GRS_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grs_geom.h"

/* How an evaluation index maps to (voxel slot, orientation slot). */
enum grs_eval_map {
  GRS_MAP_ONE_VOXEL = 0, /* voxel 0, orientation e      (cuda_forward_comp) */
  GRS_MAP_ONE_U = 1,     /* voxel e, orientation 0      (cuda_grow_indexed_region) */
  GRS_MAP_PAIRS = 2,     /* voxel eval_vox[e], orientation eval_u[e] (compete) */
};

typedef struct grs_eval_args {
  const grs_dgeo *g;       /* device copy of the geometry */
  const double *hkl;       /* H x 3 */
  const double *cs;        /* nrot x 2: cos, sin of omega */
  const double *vw;        /* V x nrot x GRS_VW_DOUBLES (fast kernel) */
  const double *pos;       /* V x 3 voxel positions in mm (plain kernel) */
  const double *SUB;       /* NU x 9 */
  const int32_t *valid;    /* NU: 0 -> orientation row skipped */
  const uint32_t *keys;    /* distance stack as ordered keys, nz*ny*nrot */
  int64_t n_eval;
  int32_t map;
  const int32_t *eval_vox, *eval_u;
  uint32_t *scratch;       /* per-wave (fast) or per-thread (plain) key lists */
  int64_t scratch_stride;  /* keys per list = nrot*H */
  int32_t *out_ns, *out_ni;
  double *out_C;
  float *out_med;
} grs_eval_args;

/* Launchers (return hipError_t). */
hipError_t grs_launch_vw_table(const grs_dgeo *d_g, const double *d_pos, int64_t V, const double *d_cs, int32_t nrot,
                               double *d_vw, hipStream_t st);
hipError_t grs_launch_voxel_pos(const double *d_rows, int64_t M, int32_t row_stride, const double *d_p1,
                                double *d_pos, hipStream_t st);
hipError_t grs_launch_sub_from_q(const grs_dgeo *d_g, const double *d_q, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st);
hipError_t grs_launch_sub_from_U(const grs_dgeo *d_g, const double *d_U, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st);
hipError_t grs_launch_sub_from_euler(const grs_dgeo *d_g, const double *d_euler, int64_t N, double *d_SUB,
                                     int32_t *d_valid, hipStream_t st);
hipError_t grs_launch_omega_cs(const double *d_rot, int32_t nrot, double *d_cs, hipStream_t st);
hipError_t grs_launch_fast(const grs_eval_args *a, int32_t H, int64_t n_waves, hipStream_t st);
hipError_t grs_launch_plain(const grs_eval_args *a, int64_t n_threads, hipStream_t st);
hipError_t grs_launch_compete_reduce(const int64_t *d_row_start, int64_t M, const int32_t *d_ns, const int32_t *d_ni,
                                     const double *d_C, const float *d_med, int64_t *d_best, hipStream_t st);
hipError_t grs_launch_math_eval(int fn, const double *d_x, double *d_y, int64_t n, hipStream_t st);
hipError_t grs_launch_check_keys(const uint32_t *d_keys, int64_t n, uint32_t *d_bad, hipStream_t st);

hipError_t grs_launch_render(const grs_dgeo *d_g, const double *d_pos, const int32_t *d_grain, int64_t M,
                             int32_t nrot, const double *d_SUB, const double *d_hkl, const double *d_cs,
                             uint8_t *d_mask, hipStream_t st);
hipError_t grs_launch_edt_cols(const uint8_t *d_mask, int64_t nz, int64_t ncols, int32_t *d_g1, hipStream_t st);
hipError_t grs_launch_edt_rows(const int32_t *d_g1, int64_t nz, int64_t ny, int64_t slice0, int64_t nslices,
                               int32_t *d_v, double *d_z, float *d_out, hipStream_t st);

/* Wave slots the fast kernel keeps resident (its scratch is sized from this). */
int64_t grs_fast_wave_slots(int device);

#endif
