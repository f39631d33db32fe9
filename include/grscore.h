/*
 * grscore.h -- C ABI of the MI355X (gfx950) forward-completeness scoring
 * engine for LabDCT grain mapping.
 *
 * Plain C types only (pointers, sizes, doubles); no HIP or torch types cross
 * this boundary.  Every function returns a grs_status (0 on success, a
 * negative code otherwise) and never aborts the calling process; the text of
 * the last error is available from grs_last_error().
 *
 * Two families of entry points:
 *
 *  1. Legacy MEX operators, one per native operator of GrainRecon.  They take
 *     exactly the single-precision row vectors that the MATLAB wrappers build
 *     and return the same single-precision output layout, so a MEX gateway
 *     can forward its mxArrays unchanged.  The inputs are
 *     promoted to double and scored in FP64.
 *       grs_cuda_forward_comp          replaces cuda_forward_comp
 *           (call sites ForwardCalc_OR_space_comp_cuda.m:32,
 *            fit_OR_constrain_with_local_comp_cuda.m:47,
 *            fit_OR_constrain_with_local_comp_cuda_2.m:62,
 *            forward_calc_gridding_OR.m:17)
 *       grs_cuda_grow_indexed_region   replaces cuda_grow_indexed_region
 *           (grow_indexed_region_cuda.m:105-107,
 *            Forward_calc_comp_allvoxel_cuda.m:57-59)
 *       grs_cuda_forward_comp_compete  replaces cuda_forward_comp_compete
 *           (grow_unindexed_compete_comp_cuda.m:39-41)
 *     The MEX operators receive no beam-centre pixel (dety0, detz0) and no hit
 *     threshold; these entries use dety0 = detysize/2 + 0.5,
 *     detz0 = detzsize/2 + 0.5 (the convention of
 *     revise_all_unindexed_voxel_cuda.m:40-42) and a threshold of 0.5 pixel
 *     (minEucDis = 0.5*mean([pixelysize pixelzsize]), setup_para.m:34,91,133).
 *
 *  2. Native FP64 entries on a context that keeps the geometry and the
 *     distance stack resident in HBM between calls (the MEX path re-uploads
 *     the stack on every call and resets the device afterwards).
 *
 * Layout conventions shared by both families:
 *  - 3x3 matrices are row-major 9-vectors (MATLAB reshape(X',1,9));
 *  - hkl is nrhkl triples (h,k,l);
 *  - the distance stack is MATLAB column-major proj_bin_bw(detz, dety, i):
 *    element (iz, iy, i) (1-based) at (iz-1) + (iy-1)*nz + (i-1)*nz*ny,
 *    with nz >= detzsize rows and ny >= detysize columns; values must be
 *    non-negative and not NaN (bwdist distances in pixels);
 *  - a distance d is a hit when (double)d <= minEucDis/mean([py pz]).
 */
#ifndef GRSCORE_H
#define GRSCORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t grs_status;
#define GRS_OK 0
#define GRS_ERR_ARG (-1)      /* invalid argument, shape or value */
#define GRS_ERR_DEVICE (-2)   /* a HIP runtime call failed */
#define GRS_ERR_STATE (-3)    /* geometry or distance stack not set */
#define GRS_ERR_NOMEM (-4)    /* device allocation failed */

#define GRS_METHOD_FAST 0  /* wave-per-evaluation filtered kernel (default) */
#define GRS_METHOD_PLAIN 1 /* thread-per-evaluation port (baseline) */

typedef struct grs_ctx grs_ctx;

/* Geometry and selection parameters: the argument list of forward_comp.m:4-6
 * (forward_comp.m is the CPU scorer every native operator reproduces). */
typedef struct grs_geometry {
  double S[9];        /* sample-frame matrix S */
  double B[9];        /* reciprocal-lattice matrix B = FormB(cell) */
  double RotDet[9];   /* detector rotation get_det_R(tilt_x, tilt_y, tilt_z) */
  double Lsam2sou, Lsam2det;      /* mm */
  double P0y, P0z, RotAxisOffset; /* mm */
  double dety00, detz00;          /* detector offset, mm */
  double pixelysize, pixelzsize;  /* mm */
  double dety0, detz0;            /* beam centre, pixels */
  double detysize, detzsize;      /* pixels */
  double BeamStopY[2], BeamStopZ[2]; /* inclusive pixel ranges; [0 -1] = none */
  double thetamax;                /* degrees */
  double lambda_min, lambda_max;  /* Angstrom */
  double minEucDis;               /* mm */
} grs_geometry;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on HIP device `device`. */
grs_status grs_create(int32_t device, grs_ctx **out);
void grs_destroy(grs_ctx *ctx);
/* Text of the last error on ctx (ctx may be NULL for the legacy entries). */
const char *grs_last_error(const grs_ctx *ctx);
const char *grs_version(void);

/* Geometry, reflections and rotation angles (degrees) used by later calls. */
grs_status grs_set_geometry(grs_ctx *ctx, const grs_geometry *geo, const double *hkl, int64_t nrhkl,
                            const double *rot_angles, int64_t nrot);

/* Distance stack (see layout above).  src_on_device != 0 means `proj` is a
 * device pointer; the data is copied either way. */
grs_status grs_set_distance_stack(grs_ctx *ctx, const float *proj, int64_t nz, int64_t ny, int64_t nrot,
                                  int32_t src_on_device);

/* ---- native FP64 operators ---------------------------------------------- */

/* One voxel at pos (mm) against N orientations given as rows [q0 q1 q2 q3];
 * q0 is recomputed from the vector part and rows with |q_vec| > 1 give zeros
 * (ForwardCalc_OR_space_comp.m:22-53).  Outputs per row: Nr_simu,
 * Nr_intersect, C = Nr_intersect/Nr_simu (NaN when Nr_simu = 0) and the
 * median distance (2000 when no spot lands on the detector).  Host pointers.
 * Replaces cuda_forward_comp (ForwardCalc_OR_space_comp_cuda.m:32). */
grs_status grs_forward_comp(grs_ctx *ctx, const double *pos, const double *q, int64_t N, int32_t *nr_simu,
                            int32_t *nr_intersect, double *completeness, float *dis_median, int32_t method);

/* Same on device-resident orientations and outputs (pos stays on the host). */
grs_status grs_forward_comp_device(grs_ctx *ctx, const double *pos, const double *d_q, int64_t N,
                                   int32_t *d_nr_simu, int32_t *d_nr_intersect, double *d_completeness,
                                   float *d_dis_median, int32_t method);

/* Allocation-free form of grs_forward_comp_device (fast method).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The call makes no hipMalloc/hipFree and no allocating
 * library call; its working memory (voxel position and bounds, S*U*B of a
 * chunk of GRS_NOALLOC_CHUNK orientations, per-wave median key lists) lives in
 * one static device array of GRS_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends.  The context must already hold the geometry and the
 * distance stack (grs_set_geometry / grs_set_distance_stack allocate).  Any N
 * is accepted (chunks of GRS_NOALLOC_CHUNK); geometries whose nrot*nrhkl
 * events per evaluation leave room for fewer than 4 waves' key lists
 * (nrot*nrhkl > ~3.9e6) are rejected with GRS_ERR_ARG.  The key lists bound
 * the number of resident waves (about 1000 at 181 omega x 86 reflections), so
 * it runs slower than grs_forward_comp_device on large N; results are
 * identical.  All calls on one GPU share the array: calls that overlap on one
 * GPU (two contexts, two host threads) race. */
#define GRS_NOALLOC_SCRATCH_BYTES ((int64_t)64 << 20)
#define GRS_NOALLOC_CHUNK ((int64_t)16384)
grs_status grs_forward_comp_device_noalloc(grs_ctx *ctx, const double *pos, const double *d_q, int64_t N,
                                           int32_t *d_nr_simu, int32_t *d_nr_intersect, double *d_completeness,
                                           float *d_dis_median);

/* One orientation U (row-major) against M voxels; vox rows are
 * [x y z C mask DisMedian] with 1-based voxel indices, param1 the 14 values
 * [RecVolumePixel(:,1)' simap_data_flag VoxSize Center' Dimension seed_comp
 * drop_off maxDmedian] (grow_indexed_region_cuda.m:71-95).  out rows are
 * [x y z Nr_simu Nr_intersect dis_median growth_flag]; voxels with mask <= 0
 * are not scored (grow_indexed_region_parallel.m:69) and give zeros.
 * growth_flag = C > C_old && dis_median <= min(DisMedian, maxDmedian) &&
 * C >= seed_comp*(1-drop_off) (grow_indexed_region_parallel.m:78-81; the
 * operator receives no Ninter column, so the equal-C clause is not applied).
 * Replaces cuda_grow_indexed_region (grow_indexed_region_cuda.m:105). */
grs_status grs_grow_indexed_region(grs_ctx *ctx, const double *vox, int64_t M, const double *U,
                                   const double *param1, double *out, int32_t method);

/* M voxels, rows [x y z start end] (1-based inclusive indices into id_all,
 * find_neighbor_ORs.m:26-29) against the candidate grains id_all (1-based);
 * euler holds 3 angles (degrees) per grain id.  param1 = the first 11 values
 * of the grow param1.  out rows [i phi1 PHI phi2 GrainId C Nr_intersect
 * dis_median] for the first maximum of C (revise_all_unindexed_voxel.m:152).
 * Replaces cuda_forward_comp_compete (grow_unindexed_compete_comp_cuda.m:39). */
grs_status grs_forward_comp_compete(grs_ctx *ctx, const double *vox, int64_t M, const double *id_all, int64_t n_id,
                                    const double *euler, int64_t n_grains, const double *param1, double *out,
                                    int32_t method);

/* Batched compete over CSR rows with local orientation sampling (one launch
 * over every candidate of every row).  Row r is a voxel at pos[3r..3r+2] (mm)
 * whose candidate orientations are cand_q[4c..4c+3] for c in
 * row_ptr[r] .. row_ptr[r+1]-1 (M+1 offsets, row_ptr[0] = 0, row_ptr[M] = K).
 * Each candidate expands to (nr+1)^3 samples on the grid of
 * ori_local_sampling.m:22-30: q_i = (c_i - (gridsize*nr)/2) + k_i*gridsize,
 * k in [0, nr]^3 with q1 slowest, q0 recomputed.  Per row: the first maximum
 * of C over the expanded list (revise_all_unindexed_voxel.m:152), as its index
 * j in that list (best, -1 for an empty row), its quaternion (best_q), counts,
 * C and median.  The scoring launch counts hits on a bit-packed mask and
 * reduces a packed 64-bit argmax key per row; only the winners are re-scored
 * for the median.  Each chunk of rows runs on the device without host
 * round trips, as a replayed HIP graph (grs_set_graphs).  Requires nrot*nrhkl <= 23170.  No reference counterpart:
 * the reference scores these lists in serial 30000-voxel MEX calls
 * (GrainGrow_engine_compete_comp.m:36-45) without local sampling. */
grs_status grs_compete_csr(grs_ctx *ctx, const double *pos, int64_t M, const int64_t *row_ptr, const double *cand_q,
                           int64_t K, int32_t nr, double gridsize, int64_t *best, double *best_q, int32_t *nr_simu,
                           int32_t *nr_intersect, double *completeness, float *dis_median);

/* ---- spot-centroid residual and its adjoint (geometry / orientation fits) - */

/* Matched spots kept resident for the fits: spot i has reflection hkl[3i..],
 * rotation angle omega[i] (degrees), measured centroid meas[2i..2i+1]
 * (dety, detz pixels) and belongs to grain grain[i] (0-based, sorted
 * non-decreasing, < G).  The SpotsPair columns of
 * Forward_simu_spots_exp.m:323-325 (hkl 14-16, omega 18, centroid 19-20). */
grs_status grs_set_spots(grs_ctx *ctx, int64_t n, const double *hkl, const double *omega, const double *meas,
                         const int32_t *grain, int64_t G);

/* Residual r (2n: unrounded simulated minus measured centroid, pixels) of the
 * resident spots for geometry geo7 = [Lsam2sou Lsam2det dety00 detz00 tilt_x
 * tilt_y tilt_z] (mm, degrees; geo_fitting.m:8-17) and per-grain parameters
 * grains6 = [r1 r2 r3 x y z] (Rodrigues vector, centre of mass in mm;
 * fit_grain_COM.m:116-118), the other constants from the context geometry.
 * With grad7/grad6 the hand-written adjoint adds the vector-Jacobian product
 * J^T u (u = upstream, or r itself when upstream is NULL: the gradient of
 * 0.5*|r|^2); with jac the 2n x 13 Jacobian rows [d/d geo7, d/d grains6 of the
 * spot's grain].  Replaces the fminsearchbnd objectives geo_fitting.m:5-79 and
 * fit_grain_COM.m:113-170 with a differentiable, unrounded residual. */
grs_status grs_spot_residual(grs_ctx *ctx, const double *geo7, const double *grains6, double *resid,
                             const double *upstream, double *grad7, double *grad6, double *jac);

/* Gauss-Newton normal equations of the same residual for a Levenberg-Marquardt
 * step over geo7 and grains6, reduced on the device in a fixed order: Agg
 * (7 x 7) = sum w Jg^T Jg and bg (7) = sum w Jg^T r over all spots; per grain
 * g, Akk (G x 6 x 6) = sum w Jk^T Jk, Agk (G x 7 x 6) = sum w Jg^T Jk and bk
 * (G x 6) = sum w Jk^T r over its spots (row-major blocks); sums = [sum rho,
 * sum |r|^2].  delta > 0 weighs a spot by w = 1/sqrt(|r|^2 + delta^2) and
 * rho = sqrt(|r|^2 + delta^2) - delta (a smooth stand-in for the Euclidean
 * distance that geo_fitting.m:70-78 and fit_grain_COM.m:84-86 average);
 * delta = 0 gives least squares (w = 1, rho = |r|^2 / 2).  Replaces, with
 * the adjoint, the objective evaluations of fminsearchbnd in
 * L_shift_tilt_fitting.m:23-50 and fit_grain_COM.m:5-29. */
grs_status grs_spot_normal(grs_ctx *ctx, const double *geo7, const double *grains6, double delta, double *Agg,
                           double *bg, double *Akk, double *Agk, double *bk, double *sums);

/* ---- legacy single-precision MEX operators (stateless) ------------------ */

/* cuda_forward_comp(pos, q, S, B, hkl, proj, rot, RotDet, param):
 * param = [nr_hkl nr_rot Lsam2sou Lsam2det P0y P0z dety00 detz00 pixelysize
 * pixelzsize detysize detzsize BeamStopY(2) BeamStopZ(2) thetamax lambda_min
 * lambda_max RotAxisOffset] (ForwardCalc_OR_space_comp_cuda.m:17-27).
 * out: 4 floats per orientation [Nr_simu Nr_intersect C dis_median]. */
grs_status grs_cuda_forward_comp(const float *pos, const float *q, int64_t n_q, const float *S, const float *B,
                                 const float *hkl, int64_t n_hkl, const float *proj, int64_t n_proj,
                                 const float *rot, int64_t n_rot, const float *RotDet, const float *param,
                                 int64_t n_param, float *out);

/* cuda_grow_indexed_region(vox6, U, S, B, hkl, proj, rot, RotDet, param, param1):
 * out: 7 floats per voxel (grow_indexed_region_cuda.m:101-115). */
grs_status grs_cuda_grow_indexed_region(const float *vox, int64_t n_vox, const float *U, const float *S,
                                        const float *B, const float *hkl, int64_t n_hkl, const float *proj,
                                        int64_t n_proj, const float *rot, int64_t n_rot, const float *RotDet,
                                        const float *param, int64_t n_param, const float *param1,
                                        int64_t n_param1, float *out);

/* cuda_forward_comp_compete(vox5, id_all, Euler, S, B, hkl, proj, rot,
 * RotDet, param, param1): out 8 floats per voxel
 * (grow_unindexed_compete_comp_cuda.m:17-48). */
grs_status grs_cuda_forward_comp_compete(const float *vox, int64_t n_vox, const float *id_all, int64_t n_id,
                                         const float *euler, int64_t n_euler, const float *S, const float *B,
                                         const float *hkl, int64_t n_hkl, const float *proj, int64_t n_proj,
                                         const float *rot, int64_t n_rot, const float *RotDet,
                                         const float *param, int64_t n_param, const float *param1,
                                         int64_t n_param1, float *out);

/* ---- input preparation (no counterpart among the MEX operators) --------- */

/* Forward-simulates binary spot masks with the same geometry: every selected
 * (omega, hkl) event of voxel i (position pos[3i..3i+2] mm, orientation
 * U[9*grain[i]..], row-major) sets its pixel in the context's detzsize x
 * detysize x nrot mask (MATLAB order).  clear != 0 zeroes the mask first. */
grs_status grs_render_spots(grs_ctx *ctx, const double *pos, const int32_t *grain, int64_t M, const double *U,
                            int64_t G, int32_t clear);
/* Lists every selected event of each voxel (position pos[3i..] mm,
 * orientation U[9*grain[i]..]) with its omega index, hkl index and unrounded
 * pixel coordinates (-dety22/py + dety0, -detz22/pz + detz0), sorted by
 * (voxel, omega, hkl): the simulated spot list of Forward_simu_spots_exp.m
 * with the forward_comp geometry.  Writes min(total, cap) events and the
 * total count to n_out. */
grs_status grs_simulate_spots(grs_ctx *ctx, const double *pos, const int32_t *grain, int64_t M, const double *U,
                              int64_t G, int64_t cap, int32_t *vox_out, int32_t *omega_out, int32_t *hkl_out,
                              double *yz_out, int64_t *n_out);
/* Copies a binary mask (nonzero = spot pixel, MATLAB order, nz x ny x nrot)
 * into the context's mask.  src_on_device as for grs_set_distance_stack. */
grs_status grs_set_binary_stack(grs_ctx *ctx, const uint8_t *mask, int64_t nz, int64_t ny, int64_t nrot,
                                int32_t src_on_device);
/* Exact Euclidean distance transform of the context's mask (bwdist of
 * GrainMapping_fun.m:73-75, as single) installed as the distance stack. */
grs_status grs_distance_from_mask(grs_ctx *ctx);
/* Copies the current mask / distance stack to host memory (n elements). */
grs_status grs_get_mask(grs_ctx *ctx, uint8_t *out, int64_t n);
grs_status grs_get_distance_stack(grs_ctx *ctx, float *out, int64_t n);

/* Device memory helpers on ctx's device (for callers without a GPU runtime
 * of their own, e.g. the benchmark keeping orientations resident in HBM). */
grs_status grs_device_alloc(grs_ctx *ctx, int64_t bytes, void **d_out);
grs_status grs_device_free(grs_ctx *ctx, void *d);
grs_status grs_copy_to_device(grs_ctx *ctx, void *d_dst, const void *src, int64_t bytes);
grs_status grs_copy_to_host(grs_ctx *ctx, void *dst, const void *d_src, int64_t bytes);
/* Device-to-device copy by a plain copy kernel (bytes % 4 == 0); calibrates
 * memory-traffic counters against a transfer of known size. */
grs_status grs_device_copy(grs_ctx *ctx, void *d_dst, const void *d_src, int64_t bytes);

/* ---- orientation fit ------------------------------------------------------- */

/* M independent orientation fits (fit_OR_constrain_with_local_comp_cuda.m:
 * 116-126 with fit_OR_fun_local.m:2-99): fit m minimises
 * 1 - Nr_intersect/Nr_simu of the voxel at pos[3m..] over x = (q1, q2, q3)
 * (q0 = sqrt(1 - (q1^2+q2^2+q3^2))), starting at x0[3m..] inside the bounds
 * max(x0 - half_width, -1) .. min(x0 + half_width, 1) (the reference uses
 * 0.02), by fminsearchbnd with MaxFunEvals, MaxIter, TolFun, TolX (the
 * reference passes 2000, 2000, 1e-10 and the default 1e-4).  Each fit follows
 * the sequential method exactly; the evaluations of all active fits are
 * scored together in one device launch per step (while the batch is small,
 * with each requested reflection's three possible follow-ups, so that an
 * expansion or contraction needs no launch of its own; the counts returned
 * are the method's own).  A point with no selected
 * event or |q_vec| > 1 has the value NaN.  Outputs: x_out (3 per fit), fval,
 * the number of evaluations and of iterations. */
grs_status grs_fit_or_local(grs_ctx *ctx, const double *pos, const double *x0, int64_t M, double half_width,
                            int32_t max_fun_evals, int32_t max_iter, double tol_fun, double tol_x, double *x_out,
                            double *fval, int32_t *n_fun_evals, int32_t *n_iter);

/* ---- seed indexing ---------------------------------------------------------- */

/* Makes the orientation set of the seed indexing (OR.q of
 * GrainIndexing_engine_gpu_comp.m:46, N rows [q0 q1 q2 q3]) resident on the
 * device; its S*U*B matrices (q0 recomputed from the vector part) follow the
 * current geometry. */
grs_status grs_set_orientation_set(grs_ctx *ctx, const double *q, int64_t N);

/* GrainIndexing_engine_gpu_comp.m:46-66 for n_pos seed positions (pos,
 * n_pos x 3 mm),
 * each with fit_OR_constrain_with_local_comp_cuda.m:12-165 (Nr_fit = n_fit):
 *   scores of every orientation of the set at the seed; the first
 *   min(nr_seed, N) rows of sortrows(Output,3,'descend') (NaN first, ties in
 *   set order); ori_local_sampling(seed q, nr, gridsize) around each, laid out
 *   as q_all (seed j at rows (j-1)*len_j+1 .. j*len_j, later seeds
 *   overwriting, unwritten rows zero); their scores; per seed the first
 *   maximum of C over rows 1+nr_per_OR*(j-1) .. nr_per_OR*j (nr_per_OR the
 *   last seed's row count); sortrows(Output_max,3,'descend'); fminsearchbnd
 *   of 1 - C from the first n_fit rows (x0 = their q1..q3, bounds
 *   x0 -/+ half_width clipped to [-1, 1], MaxFunEvals, MaxIter, TolFun, TolX);
 *   Euler angles / Rodrigues vector / quaternion of each fit (:133-147), the
 *   verification counts with U = quaternion2U(quat) (the counts and median of
 *   forward_comp_pair_spots.m:3-93, 2000 when nothing is selected) and the fit
 *   of highest C (:161, first maximum, NaN ignored).
 * The steps up to the fit starts run on the device without host round trips
 * (captured in a HIP graph and replayed for further batches of the same
 * shape); the fits of all seeds are batched.
 * Outputs: indexing (n_pos x 18) = [seed number (1-based), pos, Nr_simu,
 * Nr_intersect, C, dis_median, Euler ZXZ (degrees), Rodrigues (3),
 * quaternion (4)]; candidate (n_pos) = column 9 of the fitted Output_max row
 * (the 1-based rank of the seed whose local sampling gave it, 0 for an
 * unwritten row); optional fit_info (n_pos x 9) = [x0 (3), x (3), fval,
 * n_fun_evals, n_iter]; optional seed_idx / seed_C (n_pos x nr_seed): 0-based
 * orientation-set indices of the index_seeds rows and their C (-1 / NaN
 * beyond N).  Requires nrot*nrhkl <= 23170, nr <= 11, min(nr_seed, N) <= 4096. */
grs_status grs_index_seeds(grs_ctx *ctx, const double *pos, int64_t n_pos, int32_t nr_seed, int32_t nr, double gridsize,
                           int32_t n_fit, double half_width, int32_t max_fun_evals, int32_t max_iter, double tol_fun,
                           double tol_x, double *indexing, int32_t *candidate, double *fit_info, int32_t *seed_idx,
                           double *seed_C);

/* The orientation bookkeeping of fit_OR_constrain_with_local_comp_cuda.m:
 * 133-148 on the host for M fitted vector parts x (M x 3): out10 (M x 10) =
 * [Euler ZXZ (degrees, in [0, 360)), Rodrigues (3), quaternion (4)], U (M x 9,
 * quaternion2U of that quaternion), ok (M) = 0 where rod2angle's gimbal-lock
 * branch or a complex intermediate would be needed (outputs NaN). */
grs_status grs_fit_orientations(const double *x, int64_t M, double *out10, double *U, int32_t *ok);

/* ---- spot pairing ------------------------------------------------------------ */

/* Experimental spots per projection (Spots{i}(:,9:11) of
 * forward_comp_pair_spots.m:73-79): nframes + 1 offsets (frame i owns rows
 * frame_ptr[i] .. frame_ptr[i+1]-1 of spots) and n rows [centroid y,
 * centroid z, area] in pixels.  nframes must equal the number of rotation
 * angles when pairing. */
grs_status grs_set_frame_spots(grs_ctx *ctx, const int64_t *frame_ptr, int64_t nframes, const double *spots,
                               int64_t n);

/* forward_comp_pair_spots.m:3-93 for M (pos, U) pairs (pos M x 3 mm, U M x 9
 * row-major): Nr_simu, Nr_intersect and dis_median (2000 when nothing is
 * selected) in ns, ni, med; row_ptr (M + 1) the offsets of each pair's rows;
 * rows (cap x 23) the HittedSpots rows [reshape(U,1,9) pos hkl Energy
 * rot_angle spot_y spot_z spot_area dety detz dis] in (projection,
 * reflection) order, with pos(2) - RotAxisOffset (:16) and every value
 * rounded to single precision as MATLAB's concatenation with the single dis
 * does.  A hit is paired with the nearest spot of its projection (first
 * minimum) when that distance is below min(sqrt(area), 15).  rows = NULL
 * returns row_ptr only; GRS_ERR_ARG when the rows exceed cap.  Replaces the
 * host loop of forward_comp_pair_spots.m (called from
 * fit_OR_constrain_with_local_comp_cuda.m:89,151). */
grs_status grs_pair_spots(grs_ctx *ctx, const double *pos, const double *U, int64_t M, int32_t *ns, int32_t *ni,
                          float *med, int64_t *row_ptr, double *rows, int64_t cap);

/* ---- region growth with resident volumes ---------------------------------------- */

/* Reconstruction volumes kept on the device, nx x ny x nz doubles with x
 * fastest (MATLAB's V(x,y,z) memory order): which = 0 Completeness, 1 Mask,
 * 2 Dismedian, 3 Ninter (the DS fields GrainMapping_engine_gpu_comp.m:
 * 22-32 passes to the growth).  Setting a volume of another size drops the
 * others.  grs_get_volume copies one back (n = nx*ny*nz). */
grs_status grs_set_volume(grs_ctx *ctx, int32_t which, const double *v, int64_t nx, int64_t ny, int64_t nz);
grs_status grs_get_volume(grs_ctx *ctx, int32_t which, double *v, int64_t n);
/* out[i] = volume `which` at linear index lin[i] (x fastest, 0-based; every
 * index checked against nx*ny*nz): a caller mirroring the volumes on the
 * host fetches only the voxels a growth changed (no reference counterpart;
 * the reference keeps its volumes on the host). */
grs_status grs_get_volume_at(grs_ctx *ctx, int32_t which, const int64_t *lin, int64_t n, double *out);
/* volume `which` at linear index lin[i] = values[i] (indices checked and
 * distinct): the write-back of a growth done on the host, e.g. the
 * connected-voxel growth (no reference counterpart; the reference keeps its
 * volumes on the host). */
grs_status grs_set_volume_at(grs_ctx *ctx, int32_t which, const int64_t *lin, int64_t n, const double *values);

/* grow_indexed_region_cuda.m:29-151 around the 1-based seed voxel with the
 * orientation U (row-major) on the resident volumes: Completeness(seed) =
 * indexed_comp; the voxels of the box of half width
 * Box_L = ceil(spot_L*mean([pixelysize pixelzsize]) / ((Lsam2det/Lsam2sou)*VoxSize))
 * within Box_L of its centre point are scored (inside the mask) and grown by
 * the operator's rule (seed_comp = indexed_comp, drop_off, maxDmedian); the
 * accepted voxels' Completeness, Dismedian and Ninter are written into the
 * volumes (the caller's Completeness(ind) = Completeness_out(ind), :232-234).
 * p1 = [RecVolumePixel(:,1)' simap_data_flag VoxSize tomo_scale.Center'
 * tomo_scale.Dimension] (11).  Outputs: the completeness-weighted centre
 * test of :134-151 (center: round(weighted centre) with replace_center = 1,
 * or the initial centre with 0), n_indexed = the size of the grown region J
 * (seed included), its voxels (1-based [x y z], find order) in indexed
 * (cap x 3, the first cap of them), and n_checked = the voxels of the
 * sphere.  Replaces the wrapper around cuda_grow_indexed_region
 * (GrainMapping_engine_gpu_comp.m:222-227). */
grs_status grs_grow_region(grs_ctx *ctx, const int32_t *seed, double indexed_comp, const double *U, double spot_L,
                           double drop_off, double maxD, double maxDmedian, const double *p1, double *center,
                           int32_t *replace_center, int64_t *n_indexed, int32_t *indexed, int64_t cap,
                           int64_t *n_checked);

/* ---- grain centre-of-mass fit ------------------------------------------------------ */

/* fit_grain_COM.m:5-29 (fit_grain_COM_only = 0) for M grains: the pair rows
 * of grain m (rows[23*row_ptr[m] ..], the grs_pair_spots layout without the
 * leading seed column) give x0 = [position (columns 11:13 of hittedSpots_pair),
 * u2euler_corr(reshape(columns 2:10, 3, 3))] and the bounds x0 -/+ 0.1 mm and
 * max(x0-3, 0) .. min(x0+3, 360) degrees; fminsearchbnd (MaxFunEvals,
 * MaxIter, TolFun, TolX; the reference passes 2000, 500, 1e-10 and the
 * default 1e-4) minimises fit_grain_COM_fun: the mean over the pairs of the
 * distance between the rounded simulated pixel of the pair's (hkl, omega)
 * and its measured centroid.  Each fit runs its whole Nelder-Mead on the
 * device, one wave per grain, all grains in one launch.  x_out (M x 6) =
 * FitOutput, fval, the evaluation and iteration counts.  max_iter is at most
 * 100000 (it bounds a wave's run time). */
grs_status grs_fit_grain_com(grs_ctx *ctx, const double *rows, const int64_t *row_ptr, int64_t M,
                             int32_t max_fun_evals, int32_t max_iter, double tol_fun, double tol_x, double *x_out,
                             double *fval, int32_t *n_fun_evals, int32_t *n_iter);

/* rod = angle2rod(euler*pi/180, 'ZXZ'), quat = rod2quat(rod) for M ZXZ angle
 * triples in degrees (GrainMapping_engine_gpu_comp.m:148-149; host only):
 * rod_quat (M x 7) = [rod(3) quat(4)], ok = 0 where acos((tr-1)/2) would be
 * complex (rod zero). */
grs_status grs_euler_to_rod_quat(const double *euler_deg, int64_t M, double *rod_quat, int32_t *ok);

/* euler2u.m for M ZXZ angle triples in degrees converted as (deg*pi)/180
 * (host only; the engine's FP64 functions): U (M x 9) row-major. */
grs_status grs_euler2u_deg(const double *euler_deg, int64_t M, double *U);

/* u2euler_corr.m for M matrices given column-major (host only). */
grs_status grs_u2euler(const double *U_colmajor, int64_t M, double *euler_deg);

/* ---- G-vector matching score ("Gt" indexing method) ------------------------------ */

/* The score behind cuda_forward_Gt (ForwardCalc_OR_space_Gt_cuda.m:7-47),
 * whose CUDA source is not in the reference: for the seed position pos and N
 * orientation rows q ([q0 q1 q2 q3], q0 recomputed; |q_vec| > 1 gives a zero
 * row), the reflections of the first min(nfam, 4) families (fam_sq =
 * hkl_family_square, d_possible in Angstrom) matched against the experimental
 * spots of each projection (grs_set_frame_spots, centroids in pixels;
 * projection i is rotation angle i): a match when the angle between Gt and
 * the nearest spot's diffraction vector (the first such spot of the
 * projection on a tie) is below 1 degree and |Gt| is within 0.1 of
 * 2 pi/d_possible.  Nr_match/NrSpotExpectAll uses calcExpectSpotNr.m with
 * the tilt-corrected distance.  semantics selects the reference function:
 *   GRS_GT_MATCH      calc_Gt_match (ForwardCalc_OR_space_Gt.m:117-205, the
 *                     CPU path whose Output(:,2:4) the MEX returns, :72):
 *                     pos(:,2) shifted by -RotAxisOffset (:135); out (N x 4) =
 *                     [Nr_match, Gt_matched_cost, Nr_match/NrSpotExpectAll,
 *                     mean matched angle (degrees)], Gt_matched_cost = the
 *                     mean over the matches of |n - w|^2 between the unit Gt
 *                     and the matched spot's unit vector (:195-198), 100 with
 *                     no match (:203);
 *   GRS_GT_MATCH_FUN  calc_Gt_match_fun.m:1-152 (used by
 *                     fit_OR_constrain_with_local_Gt.m:48,105): pos not
 *                     shifted; out = [Nr_match, median matched angle,
 *                     Nr_match/NrSpotExpectAll, mean matched angle].
 * Angle statistics are NaN when nothing matches. */
#define GRS_GT_MATCH 0
#define GRS_GT_MATCH_FUN 1
grs_status grs_forward_gt(grs_ctx *ctx, const double *pos, const double *q, int64_t N, const double *fam_sq,
                          const double *d_possible, int32_t nfam, int32_t semantics, double *out);

/* ---- candidate generation ------------------------------------------------ */

/* ori_local_sampling.m:12-34 for K centres [q0 q1 q2 q3] (used as given,
 * :16): the grids c - gridsize*nr/2 : gridsize : c + gridsize*nr/2 of q1, q2,
 * q3 (MATLAB colon operator), the centre row followed by the rows
 * [sqrt(1-(b^2+c^2+d^2)) b c d], sorted and de-duplicated like
 * unique(q,'rows').  Centre k's rows go to out_q[4*(k*cap + r) ..], their
 * number to out_n[k].  At most 2048 rows per centre before de-duplication;
 * cap must hold them.  Replaces the host loops of ori_local_sampling.m
 * (called from fit_OR_constrain_with_local_comp_cuda.m:20 and
 * GrainMapping_engine_gpu_comp.m:87). */
grs_status grs_local_sampling(grs_ctx *ctx, const double *centres, int64_t K, int32_t nr, double gridsize,
                              double *out_q, int64_t cap, int64_t *out_n);

/* find_neighbor_ORs.m:3-96: grain_id is the nx x ny x nz GrainId volume
 * (x fastest, DS_in.GrainId(x,y,z)); vox holds M rows of 1-based [x y z].
 * For each voxel: the positive ids in the box of half-width search_radius
 * clipped to the volume (grown by 2 voxels per try while empty), unique in
 * ascending order, at most nr_max (the reference's 30).  id_all (capacity
 * M*nr_max) receives them voxel after voxel, ind M rows of 1-based
 * [first last] positions, n_total their count.  Grain ids must be below
 * 524288.  Replaces the host search of GrainGrow_engine_compete_comp.m:47 and
 * revise_all_unindexed_voxel_cuda.m:88. */
grs_status grs_find_neighbor_ors(grs_ctx *ctx, const int32_t *grain_id, int64_t nx, int64_t ny, int64_t nz,
                                 const int32_t *vox, int64_t M, int32_t search_radius, int32_t nr_max,
                                 int32_t *id_all, int64_t *ind, int64_t *n_total);

/* ---- measurement and diagnostics ---------------------------------------- */

/* Enables HIP-event timing of the scoring kernel launches on ctx's stream.
 * The events are read only by grs_kernel_time (or when 64 timed launches are
 * pending), so a timed call does not wait for its kernels. */
grs_status grs_set_timing(grs_ctx *ctx, int32_t enable);
/* Accumulated scoring-kernel time (ms) and launch count since the last reset;
 * reset != 0 clears them after reading. */
grs_status grs_kernel_time(grs_ctx *ctx, double *ms_total, int64_t *launches, int32_t reset);
/* Launch autotuning (default on).  The first call of a workload class --
 * kind 0, the kernel with the median (grs_forward_comp*, grow, compete,
 * winners), kind 1, the counts-only CSR kernel of grs_compete_csr, or kind 2,
 * the counts-only grid kernel of grs_index_seeds -- for the current geometry
 * (reflection and rotation counts) that is large enough splits its
 * evaluations among the configurations, times them and keeps the fastest:
 * waves per block (8, 4, 2, 1) x the block-to-evaluation mapping (blocks in
 * launch order, or each XCD given a contiguous run of evaluations), and for
 * kinds 1 and 2 x the hit-test source (GRS_MASK_BITS: the bit mask, 1 bit per
 * pixel in stack order; GRS_MASK_KEYS: the distance keys; GRS_MASK_TILED:
 * the bit mask in 32 x 32 pixel tiles, one 128-byte line per tile, so that
 * neighbouring spots of nearby orientations share lines).  Results never
 * depend on it.  No reference counterpart (the reference has one fixed
 * launch shape). */
#define GRS_MASK_BITS 0
#define GRS_MASK_KEYS 1
#define GRS_MASK_TILED 2
#define GRS_TUNE_MAX 24
grs_status grs_set_autotune(grs_ctx *ctx, int32_t enable);
/* The chosen configuration of a kind (GRS_ERR_STATE before tuning or
 * pinning): waves per block, hit-test source (GRS_MASK_*), XCD mapping (0/1),
 * and the measured milliseconds per evaluation of each of its n_configs
 * candidates (ms_per_eval holds GRS_TUNE_MAX values; order: for mapping 0
 * then 1, for each source in GRS_MASK_* order (kind 0: the bit mask only;
 * the tiled mask only when it exists), waves per block 8, 4, 2, 1). */
grs_status grs_get_tuning(grs_ctx *ctx, int32_t kind, int32_t *waves_per_block, int32_t *mask_layout, int32_t *xcd_map,
                          int32_t *n_configs, double *ms_per_eval);
/* Pins the launch configuration of a kind (no tuning for it; n_configs 0):
 * waves per block 1, 2, 4 or 8, a GRS_MASK_* source (GRS_MASK_BITS for kind 0),
 * XCD mapping 0 or 1.  waves_per_block = 0 forgets the kind's configuration,
 * so that the next large call tunes again.  Whatever the configuration, a
 * launch of fewer waves than the device has CUs (the orientation fits'
 * steps) runs one wave per block.  No reference counterpart. */
grs_status grs_set_launch_config(grs_ctx *ctx, int32_t kind, int32_t waves_per_block, int32_t mask_layout,
                                 int32_t xcd_map);

/* HIP graphs for grs_compete_csr and grs_index_seeds (default on): each
 * chunk's device pipeline (for the compete: key clearing, bounds, counts-only
 * scoring, winners, median re-scoring, winner outputs; for the indexing: the
 * steps up to the fit starts) is captured once and replayed while its shape
 * and device buffers are unchanged; enable = 0 launches the same kernels
 * directly. */
grs_status grs_set_graphs(grs_ctx *ctx, int32_t enable);
/* Graph captures and replays since the context was created. */
grs_status grs_graph_stats(grs_ctx *ctx, int64_t *captures, int64_t *replays);

/* Audit of the fast kernel's resolution stage (one-voxel calls only, i.e.
 * grs_forward_comp / grs_forward_comp_device with the fast method): every
 * event that the kernel decides without the exact forward_comp.m pipeline is
 * also run through that pipeline and compared.  enable != 0 clears and arms
 * the counters.  No reference counterpart (a diagnostic of this engine). */
grs_status grs_set_audit(grs_ctx *ctx, int32_t enable);
/* counts[6]: events examined, resolved as selected, resolved as not selected,
 * sent to the exact pipeline, disagreements with it, events that passed the
 * angle and wavelength windows into the spot construction; max_dev: the largest
 * difference (pixels) between the resolved and the exact unrounded spot
 * coordinates of the selected events. */
grs_status grs_get_audit(grs_ctx *ctx, int64_t *counts, double *max_dev);
/* Evaluates the engine's FP64 math on the device: fn 0 sin, 1 cos, 2 acos,
 * 3 atan, 4 round (half away from zero), 5 sqrt, 6 reciprocal, 7 asin,
 * 8 filter-grade acos, 9 filter-grade atan2(sin x, cos x); the omega solve's
 * own forms: 10 acos, 11 atan2(sin x, cos x), 12 refined reciprocal,
 * 13 refined reciprocal square root.  Host arrays. */
grs_status grs_math_device(grs_ctx *ctx, int32_t fn, const double *x, double *y, int64_t n);
/* Blocks until all work on ctx's stream is done. */
grs_status grs_synchronize(grs_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* GRSCORE_H */
