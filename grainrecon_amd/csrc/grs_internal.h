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
  GRS_MAP_CSR = 3,       /* CSR rows of candidate orientations expanded by local
                            sampling; e is the global expanded index */
  GRS_MAP_CSR_LIST = 4,  /* as GRS_MAP_CSR for the expanded indices elist[e]
                            (elist[e] < 0: empty row, zero output) */
  GRS_MAP_GRID = 5,      /* voxel v = e / n_u, orientation (e % n_u) + v*u_stride:
                            every voxel against one shared list (u_stride 0) or
                            against its own block of n_u rows (u_stride n_u) */
};

/* CSR candidate rows with local sampling (ori_local_sampling.m:22-34): row r
 * (a voxel) owns candidates row_ptr[r] .. row_ptr[r+1]-1 (centre quaternions
 * cand_q[4c..4c+3]); each candidate expands to S = (nr+1)^3 samples
 * q_i = (c_i - (gridsize*nr)/2) + k_i*gridsize, k in [0, nr]^3 (q1 slowest),
 * q0 recomputed.  row_eoff[r] = S * row_ptr[r] is the first expanded index. */
typedef struct grs_csr {
  const int64_t *row_ptr;
  const int64_t *row_eoff;
  const double *cand_q;
  int64_t nrows;
  int32_t nr;
  int32_t S;
  double gridsize;
  const int64_t *elist;          /* GRS_MAP_CSR_LIST */
  unsigned long long *row_best;  /* counts-only: per-row packed argmax key */
} grs_csr;

typedef struct grs_eval_args {
  const grs_dgeo *g;       /* device copy of the geometry */
  const double *hkl;       /* H x 3 */
  const double *cs;        /* nrot x 2: cos, sin of omega */
  const double *vb;        /* V x GRS_VB omega-solve bounds (fast kernel) */
  const double *pos;       /* V x 3 voxel positions in mm */
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
  const uint32_t *hit_bits; /* counts-only kernels: bit i = (keys[i] <= hit_key) */
  grs_csr csr;
  unsigned long long *audit; /* GRS_AUDIT_WORDS counters, audit launches only */
  int64_t e0;              /* first evaluation of this launch (evaluations e0 .. n_eval-1) */
  int32_t count_keys;      /* counts-only kernels: 1 = hit test on the keys, 0 = on hit_bits */
  int64_t n_u, u_stride;   /* GRS_MAP_GRID */
  unsigned long long *out_key; /* counts-only GRS_MAP_GRID: per-evaluation sort key */
  int32_t nrot;            /* rotation angles (must equal the geometry's); 0 turns the one-voxel beam table off */
  const uint32_t *hit_tbits; /* counts-only kernels: the hit mask in 32 x 32 pixel tiles (grs_tiled_bit) */
  int32_t bits_tiled;      /* counts-only kernels with count_keys = 0: 1 = test hit_tbits, 0 = hit_bits */
  int32_t xcd_map;         /* 1 = consecutive waves of the evaluation order on one XCD (grid a multiple of 8 blocks) */
} grs_eval_args;

/* Bit of pixel (dz, dy) (1-based) of projection om in the tiled hit mask:
 * 32 x 32 pixel tiles of 128 bytes, z fastest inside a 32-bit word (one tile
 * column), the tile's 32 words for 32 consecutive dy, tiles z-fastest, then
 * dy, then omega; ntz = ceil(nz / 32), nty = ceil(ny / 32). */
__host__ __device__ inline uint32_t grs_tiled_bit(uint32_t om, uint32_t dy, uint32_t dz, uint32_t ntz, uint32_t nty) {
  const uint32_t y0 = dy - 1u, z0 = dz - 1u;
  return ((((om * nty + (y0 >> 5)) * ntz + (z0 >> 5)) << 10) | ((y0 & 31u) << 5)) | (z0 & 31u);
}

/* Audit counters of the fast kernel's resolution stage: events seen by phase
 * B1, resolved as selected, resolved as not selected, sent to the exact
 * pipeline, disagreements with the exact pipeline, the largest
 * |coordinate difference| (pixels, as the bits of a non-negative double), and
 * events that passed the angle and wavelength windows into the spot
 * construction. */
#define GRS_AUDIT_WORDS 7

/* Packed per-row argmax key of the counts-only kernel: larger is better.
 * ((floor(C*2^30) + 1) << 33) | (2^33-1-j) for a scored candidate j with
 * Nr_simu > 0, (2^33-1-j) when Nr_simu = 0 (C = NaN never beats a number).
 * Distinct ratios with denominators <= 23170 differ by more than 2^-29, so the
 * fixed-point part orders C exactly; equal C values tie and the smaller j
 * (MATLAB max: first maximum) wins. */
#define GRS_KEY_JBITS 33
#define GRS_KEY_MAX_EVENTS 23170
/* largest k of the device top-k (sortrows + first rows) */
#define GRS_TOPK_MAX 4096
/* doubles per voxel of the omega-solve bounds (k_voxel_bounds) */
#define GRS_VB 14

/* Launchers (return hipError_t). */
/* d_vb: GRS_VB doubles per voxel, the omega-solve bounds */
hipError_t grs_launch_voxel_bounds(const grs_dgeo *d_g, const double *d_pos, int64_t V, const double *d_cs,
                                   int32_t nrot, double *d_vb, hipStream_t st);
hipError_t grs_launch_voxel_pos(const double *d_rows, int64_t M, int32_t row_stride, const double *d_p1,
                                double *d_pos, hipStream_t st);
hipError_t grs_launch_sub_from_q(const grs_dgeo *d_g, const double *d_q, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st);
hipError_t grs_launch_sub_from_U(const grs_dgeo *d_g, const double *d_U, int64_t N, double *d_SUB, int32_t *d_valid,
                                 hipStream_t st);
hipError_t grs_launch_sub_from_euler(const grs_dgeo *d_g, const double *d_euler, int64_t N, double *d_SUB,
                                     int32_t *d_valid, hipStream_t st);
hipError_t grs_launch_omega_cs(const double *d_rot, int32_t nrot, double *d_cs, hipStream_t st);
/* wpb: waves per block, 1, 2 or 4; n_waves: a multiple of 4 (wave slots of the scratch) */
hipError_t grs_launch_fast(const grs_eval_args *a, int32_t H, int64_t n_waves, int32_t wpb, hipStream_t st);
hipError_t grs_launch_fast_counts(const grs_eval_args *a, int32_t H, int64_t n_waves, int32_t wpb, hipStream_t st);
hipError_t grs_launch_fast_audit(const grs_eval_args *a, int32_t H, int64_t n_waves, hipStream_t st);
hipError_t grs_launch_fill_u64(unsigned long long *d_p, int64_t n, unsigned long long v, hipStream_t st);
hipError_t grs_launch_winner_list(const unsigned long long *d_row_best, const int64_t *d_row_eoff, int64_t m,
                                  int64_t *d_elist, hipStream_t st);
hipError_t grs_launch_winner_out(const int64_t *d_elist, const int64_t *d_row_eoff, const int64_t *d_row_ptr,
                                 const double *d_cand_q, int32_t nr, double gridsize, int64_t m, int64_t *d_best,
                                 double *d_best_q, hipStream_t st);
hipError_t grs_launch_pack_hits(const uint32_t *d_keys, int64_t n, uint32_t hit_key, int32_t has_hit,
                                uint32_t *d_bits, hipStream_t st);
hipError_t grs_launch_pack_hits_tiled(const uint32_t *d_keys, int64_t nz, int64_t ny, int64_t nrot, uint32_t hit_key,
                                      int32_t has_hit, uint32_t *d_tbits, hipStream_t st);
hipError_t grs_launch_plain(const grs_eval_args *a, int64_t n_threads, hipStream_t st);
hipError_t grs_launch_compete_reduce(const int64_t *d_row_start, int64_t M, const int32_t *d_ns, const int32_t *d_ni,
                                     const double *d_C, const float *d_med, int64_t *d_best, hipStream_t st);
hipError_t grs_launch_math_eval(int fn, const double *d_x, double *d_y, int64_t n, hipStream_t st);
hipError_t grs_launch_copy_u32(const uint32_t *d_src, uint32_t *d_dst, int64_t n, hipStream_t st);
hipError_t grs_launch_check_keys(const uint32_t *d_keys, int64_t n, uint32_t *d_bad, hipStream_t st);

/* candidate generation */
#define GRS_LS_MAX_ROWS 2048 /* centre + grid rows of one local sampling */
#define GRS_LS_MAX_GRID 64   /* values per (q1, q2, q3) grid */
hipError_t grs_launch_local_sampling(const double *d_cen, const double *d_grids, const int32_t *d_gn, int64_t K,
                                     double *d_out, int64_t cap, int64_t *d_n, uint32_t *d_bad, hipStream_t st);
hipError_t grs_launch_find_neighbors(const int32_t *d_gid, int64_t nx, int64_t ny, int64_t nz, const int32_t *d_vox,
                                     int64_t M, int32_t radius, int32_t nr_max, int32_t W, int32_t *d_ids,
                                     int32_t *d_nout, uint32_t *d_bad, hipStream_t st);

hipError_t grs_launch_render(const grs_dgeo *d_g, const double *d_pos, const int32_t *d_grain, int64_t M,
                             int32_t nrot, const double *d_SUB, const double *d_hkl, const double *d_cs,
                             uint8_t *d_mask, hipStream_t st);
hipError_t grs_launch_sim_events(const grs_dgeo *d_g, const double *d_pos, const int32_t *d_grain, int64_t M,
                                 int32_t nrot, const double *d_SUB, const double *d_hkl, const double *d_cs, int64_t cap,
                                 unsigned long long *d_counter, int32_t *d_vox, int32_t *d_om, int32_t *d_h,
                                 double *d_yz, hipStream_t st);
hipError_t grs_launch_edt_cols(const uint8_t *d_mask, int64_t nz, int64_t ncols, int32_t *d_g1, hipStream_t st);
hipError_t grs_launch_edt_rows(const int32_t *d_g1, int64_t nz, int64_t ny, int64_t slice0, int64_t nslices,
                               int32_t *d_v, double *d_z, float *d_out, hipStream_t st);

hipError_t grs_launch_spot_residual(const void *fitc, int64_t n, const double *hkl, const double *omega,
                                    const double *meas, const int32_t *grain, const double *gr6, const double *up,
                                    double *res, double *contrib, double *jac, hipStream_t st);
hipError_t grs_launch_grad_reduce(const double *contrib, int64_t n, const int64_t *goff, int64_t G, double *out7,
                                  double *out6, hipStream_t st);
hipError_t grs_launch_spot_normal(const void *fitc, int64_t G, const double *hkl, const double *omega,
                                  const double *meas, const int64_t *goff, const double *gr6, double delta,
                                  double *pergrain, double *partial, double *out37, hipStream_t st);
size_t grs_fitc_size(void);

/* Wave slots the fast kernel keeps resident (its scratch is sized from this). */
int64_t grs_fast_wave_slots(int device);

/* seed indexing (grs_index.hip) */
hipError_t grs_launch_fill_u32(uint32_t *d_p, int64_t n, uint32_t v, hipStream_t st);
hipError_t grs_launch_topk(const unsigned long long *d_keys, int64_t segs, int64_t n, int64_t seg_stride, int32_t k,
                           unsigned long long *d_out, hipStream_t st);
hipError_t grs_launch_gather_centres(const double *d_set_q, const unsigned long long *d_topkey, int64_t n,
                                     double *d_cen, hipStream_t st);
hipError_t grs_launch_ls_grids(const double *d_cen, int64_t n, int32_t nr, double gridsize, int32_t cap_axis,
                               double *d_grids, int32_t *d_gn, uint32_t *d_bad, hipStream_t st);
hipError_t grs_launch_qall(const double *d_ls_q, const int64_t *d_ls_n, int64_t P, int64_t K, int64_t cap,
                           int32_t *d_owner, unsigned long long *d_Rmax, double *d_q_rows, uint32_t *d_gcount,
                           hipStream_t st);
hipError_t grs_launch_group_max(const unsigned long long *d_key2, const int64_t *d_ls_n, const uint32_t *d_gcount,
                                int64_t P, int64_t K, int64_t Rcap, unsigned long long *d_gkey, int64_t *d_grow,
                                hipStream_t st);
hipError_t grs_launch_fit_starts(const unsigned long long *d_fitkey, const int64_t *d_grow, const double *d_q_rows,
                                 const int32_t *d_owner, const int32_t *d_ns2, const int32_t *d_ni2, int64_t P,
                                 int64_t K, int32_t n_fit, int64_t Rcap, double *d_x0, int32_t *d_fgroup,
                                 int32_t *d_fns, int32_t *d_fni, hipStream_t st);
hipError_t grs_launch_seed_out(const unsigned long long *d_topkey, const int32_t *d_ns1, const int32_t *d_ni1,
                               const int32_t *d_set_valid, int64_t P, int64_t K, int64_t N, int32_t *d_seed_idx,
                               double *d_seed_C, hipStream_t st);

/* spot pairing (grs_pair.hip) */
hipError_t grs_launch_pair_events(const grs_dgeo *d_g, const double *d_hkl, const double *d_cs,
                                  const uint32_t *d_keys, const double *d_pos, const double *d_SUB, int64_t M,
                                  int64_t E, const int64_t *d_fptr, const double *d_fsp, int32_t *d_flag,
                                  int64_t *d_sid, double *d_lam, int32_t *d_dyz, float *d_dis, hipStream_t st);
hipError_t grs_launch_pair_scan(const int32_t *d_flag, int64_t M, int64_t E, int32_t *d_off, int64_t *d_count,
                                hipStream_t st);
hipError_t grs_launch_pair_rows(const grs_dgeo *d_g, const double *d_hkl, const double *d_rot, const double *d_pos,
                                const double *d_U, int64_t M, int64_t E, const double *d_fsp, const int32_t *d_flag,
                                const int64_t *d_sid, const double *d_lam, const int32_t *d_dyz, const float *d_dis,
                                const int32_t *d_off, const int64_t *d_row_ptr, double *d_rows, hipStream_t st);

/* region growth (grs_region.hip) */
hipError_t grs_launch_region_list(const double *d_I, const double *d_mask, const double *d_dm, int64_t nx, int64_t ny,
                                  const int64_t *lo, const int64_t *bdim, const double *cp, double boxL,
                                  int64_t seed_lin, double seed_comp, double *d_rows, uint32_t *d_counts,
                                  hipStream_t st);
hipError_t grs_launch_region_accept(const double *d_rows, int64_t M, const int32_t *d_ns, const int32_t *d_ni,
                                    const double *d_C, const float *d_med, const double *d_p1, int64_t nx, int64_t ny,
                                    double *d_vC, double *d_vDM, double *d_vNI, int64_t *d_acc_lin, double *d_acc_C,
                                    uint32_t *d_count, hipStream_t st);

/* out[i] = v[lin[i]] */
hipError_t grs_launch_gather_f64(const double *d_v, const int64_t *d_lin, int64_t n, double *d_out, hipStream_t st);
/* v[lin[i]] = val[i], no index repeated */
hipError_t grs_launch_scatter_f64(double *d_v, const int64_t *d_lin, const double *d_val, int64_t n, hipStream_t st);

/* G-vector matching score (grs_gt.hip) */
hipError_t grs_launch_gt_proj(const grs_dgeo *d_g, const double *d_pos, const double *d_cs, int32_t nrot, double Ltr,
                              double *d_proj, hipStream_t st);
hipError_t grs_launch_gt_spots(const grs_dgeo *d_g, const double *d_proj, const int64_t *d_fptr, int32_t nrot,
                               const double *d_fsp, int64_t nsp, double Lsd_tr, double dety00_tr, double detz00_tr,
                               const double *d_dposs, int32_t nfam, double *d_units, hipStream_t st);
int64_t grs_gt_wave_slots(int device);
/* spot order by the z of each family's unit vectors, per projection (at most
 * grs_gt_sort_max() spots per projection); d_skey/d_sidx 4*nsp, d_slen 4*nrot */
/* per (projection, family): keys (unit-vector z) sorted, the unit vectors
 * packed in that order (3 doubles), GRS_GT_NB + 1 bucket starts, the count
 * of finite keys */
#define GRS_GT_NB 256
/* d_sorig: per sorted entry, the spot's index within its projection */
hipError_t grs_launch_gt_sort(const int64_t *d_fptr, int32_t nrot, const double *d_units, int64_t nsp,
                              int64_t max_frame, double *d_skey, double *d_su, int32_t *d_sorig, int32_t *d_bst,
                              int32_t *d_slen, hipStream_t st);
int64_t grs_gt_sort_max();
/* d_skey == nullptr: every spot of a projection is visited.  cost_mode != 0:
 * rows [Nr_match, matched cost, ratio, mean] (calc_Gt_match), else
 * [Nr_match, median angle, ratio, mean] (calc_Gt_match_fun) */
hipError_t grs_launch_gt_score(const grs_dgeo *d_g, const double *d_hkl, const int32_t *d_famidx,
                               const double *d_glen_poss, const double *d_SUB, const int32_t *d_valid, int64_t N,
                               const double *d_proj, const int64_t *d_fptr, const double *d_units,
                               const double *d_skey, const double *d_su, const int32_t *d_sorig,
                               const int32_t *d_bst, const int32_t *d_slen, int64_t nsp, int32_t cost_mode,
                               double *d_scratch, int64_t cap, int64_t waves, double *d_out, hipStream_t st);

/* fit_grain_COM: the M started Nelder-Mead fits run to completion on the
 * device, one block each (grs_com.hip); d_cs holds 2 doubles and d_dis
 * GRS_NM_NMAX + 1 doubles per pair row */
struct grs_nm;
hipError_t grs_launch_com_nm(const grs_dgeo *d_g, const double *d_rows, const int64_t *d_rptr, int64_t M,
                             int64_t total, double *d_cs, double *d_dis, grs_nm *d_nm, hipStream_t st);

#endif
