/*
 * grs_capi.hip -- C ABI of the scoring engine (declared in include/grscore.h).
 *
 * The context keeps the geometry, the reflection list, cos/sin of the
 * rotation angles and the distance stack resident in HBM; per-call buffers
 * grow on demand and are reused.  All work runs on the context's own stream.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/grscore.h"
#include "grs_internal.h"
#include "grs_nm.h"
#include "grs_orient.h"

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
};

}  // namespace

/* a launch configuration of k_fast: waves per block, hit-test source of the
 * counts-only kinds (GRS_MASK_*) and the block-to-evaluation mapping */
struct tune_cfg {
  int32_t wpb, mask, xcd;
};

struct grs_tune_state {
  bool done = false;
  tune_cfg cfg = {8, GRS_MASK_BITS, 0};
  int32_t ncfg = 0;
  double ms_per_eval[GRS_TUNE_MAX] = {0};
};

struct grs_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  char err[512] = {0};
  bool geo_set = false, stack_set = false;
  grs_dgeo hgeo;
  std::vector<double> hkl;
  int32_t H = 0, nrot = 0;
  int64_t snz = 0, sny = 0, snrot = 0;
  DevBuf d_geo, d_hkl, d_rot, d_cs, d_keys;
  DevBuf d_q, d_sub, d_valid, d_pos, d_vb, d_scratch, d_ns, d_ni, d_C, d_med;
  DevBuf d_rows, d_p1, d_U, d_eu, d_rowstart, d_best, d_evox, d_eu_idx, d_bad, d_x, d_y;
  DevBuf d_mask, d_g1, d_edt_v, d_edt_z, d_grain;
  DevBuf d_bits, d_tbits, d_rowptr, d_roweoff, d_candq, d_rowbest, d_elist;
  DevBuf c_cen, c_grid, c_gn, c_out, c_n, c_gid, c_vox, c_ids, c_nout;
  bool bits_valid = false; /* hit mask matches the current stack and threshold */
  bool tbits_ok = false;    /* the tiled hit mask exists (its bit numbers fit 32 bits) */
  /* resident matched-spot set of the fits */
  DevBuf f_hkl, f_om, f_meas, f_grain, f_goff, f_gr6, f_up, f_res, f_contrib, f_jac, f_out, f_ne;
  int64_t f_n = 0, f_G = 0;
  int64_t mnz = 0, mny = 0, mnrot = 0; /* shape of d_mask */
  bool timing = false;
  bool audit = false;
  DevBuf d_audit;
  /* HIP graph of the batched-compete pipeline of one chunk, replayed while
   * its shape and buffers are unchanged */
  bool graphs = true;
  /* launch autotuning of k_fast per (kind, H, nrot): waves per block and,
   * for the counts-only kind, the hit-test source (bit mask or keys) */
  bool autotune = true;
  std::map<uint64_t, struct grs_tune_state> tunes;
  std::vector<hipEvent_t> tev;
  hipGraphExec_t g_exec = nullptr;
  std::vector<uint64_t> g_key;
  int64_t g_captures = 0, g_replays = 0;
  DevBuf d_bestidx, d_bestq;
  /* resident orientation set of the seed indexing (N x 4 as given) with its
   * S*U*B matrices for geometry generation o_gen */
  DevBuf o_q, o_sub, o_valid;
  int64_t o_n = 0;
  uint64_t geo_gen = 0, o_gen = 0;
  /* seed-indexing pipeline buffers and its graphs (per shape and buffers) */
  DevBuf x_pos, x_vb, x_key1, x_ns1, x_ni1, x_top, x_cen, x_grid, x_gn, x_ls, x_lsn, x_owner, x_Rmax, x_q2, x_sub2,
      x_valid2, x_key2, x_ns2, x_ni2, x_gcount, x_gkey, x_grow, x_fitkey, x_x0, x_fgroup, x_fns, x_fni, x_sidx, x_sC,
      x_bad;
  std::map<std::vector<uint64_t>, hipGraphExec_t> gi_cache;
  /* experimental spots per projection (CSR) for the spot pairing */
  DevBuf p_fptr, p_fsp, p_flag, p_sid, p_lam, p_dyz, p_dis, p_off, p_cnt, p_rowptr, p_rows;
  int64_t p_nframes = -1, p_n = 0, p_maxframe = 0;
  /* reconstruction volumes (completeness, mask, median distance, Nr_intersect)
   * and the region-growth buffers */
  DevBuf v_vol[4], r_rows, r_counts, r_acc_lin, r_acc_C;
  int64_t v_dim[3] = {0, 0, 0};
  /* G-vector matching score buffers */
  DevBuf t_pos, t_proj, t_units, t_fam, t_dposs, t_glen, t_scratch, t_out, t_skey, t_su, t_bst, t_slen, t_sorig;
  DevBuf v_lin, v_out; /* grs_get_volume_at */
  /* grain centre-of-mass fit buffers */
  DevBuf m_rows, m_rptr, m_nm, m_cs, m_dis;
  DevBuf f_in, f_cnt; /* fit_batch: packed per-step inputs and counts */
  void *h_pin = nullptr; /* pinned staging of fit_batch (hipHostMalloc) */
  size_t h_pin_cap = 0;
  /* device address of this GPU's allocation-free scratch (nullptr if the
   * lookup failed; noalloc_err then holds the reason) */
  unsigned char *noalloc = nullptr;
  char noalloc_err[160] = {0};
  bool v_set[4] = {false, false, false, false};
  /* HIP-event timing of the scoring launches (grs_set_timing): one event pair
   * per timed region, read back only when the totals are asked for or the
   * pool is full, so that timing never makes a call wait for its kernels */
  std::vector<hipEvent_t> tpairs; /* begin, end of pair k at 2k, 2k+1 */
  int64_t t_pending = 0;          /* pairs recorded and not yet read */
  double ms_total = 0.0;
  int64_t launches = 0;
};

/* event pairs created with the context (no call creates events later, the
 * allocation-free entry included) */
static const int64_t kTimingPairs = 64;

/* adds the elapsed times of the recorded pairs to the totals */
static hipError_t timing_harvest(grs_ctx *c) {
  for (int64_t k = 0; k < c->t_pending; k++) {
    hipError_t e = hipEventSynchronize(c->tpairs[2 * k + 1]);
    if (e != hipSuccess) return e;
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, c->tpairs[2 * k], c->tpairs[2 * k + 1]);
    if (e != hipSuccess) return e;
    c->ms_total += ms;
    c->launches += 1;
  }
  c->t_pending = 0;
  return hipSuccess;
}

static hipError_t timing_begin(grs_ctx *c) {
  if (c->t_pending == kTimingPairs) {
    hipError_t e = timing_harvest(c);
    if (e != hipSuccess) return e;
  }
  return hipEventRecord(c->tpairs[2 * c->t_pending], c->st);
}

static hipError_t timing_end(grs_ctx *c) {
  hipError_t e = hipEventRecord(c->tpairs[2 * c->t_pending + 1], c->st);
  if (e == hipSuccess) c->t_pending += 1;
  return e;
}

static grs_status fail(grs_ctx *c, grs_status code, const char *fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(c, call)                                                                                        \
  do {                                                                                                         \
    hipError_t e_ = (call);                                                                                    \
    if (e_ != hipSuccess) return fail((c), GRS_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));    \
  } while (0)

static grs_status ensure(grs_ctx *c, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.cap >= bytes) return GRS_OK;
  if (b.p) {
    hipStreamSynchronize(c->st);
    hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    (void)hipGetLastError();
    return fail(c, GRS_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  b.cap = bytes;
  return GRS_OK;
}

#define ENSURE(c, buf, bytes)                      \
  do {                                             \
    grs_status s_ = ensure((c), (buf), (bytes));   \
    if (s_ != GRS_OK) return s_;                   \
  } while (0)

template <typename T>
static T *P(DevBuf &b) {
  return static_cast<T *>(b.p);
}

static void free_buf(DevBuf &b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

/* largest non-negative float key k with (double)as_float(k) <= thr */
static void hit_key_for(double thr, uint32_t *key, int32_t *has) {
  if (!(thr >= 0.0)) {
    *has = 0;
    *key = 0;
    return;
  }
  float f = (float)thr;
  if ((double)f > thr) f = std::nextafter(f, -INFINITY);
  if (f < 0.0f || std::signbit(f)) f = 0.0f;
  uint32_t k;
  std::memcpy(&k, &f, 4);
  *has = 1;
  *key = k;
}

/* Working memory of grs_forward_comp_device_noalloc: one copy per GPU (the
 * code object's data), held until the process ends.  In host code this name
 * is not a device address; the address of the current device's copy comes
 * from hipGetSymbolAddress (noalloc_lookup) and only that is passed on. */
static __device__ __align__(256) unsigned char scratch[GRS_NOALLOC_SCRATCH_BYTES];

#define GRS_NOALLOC_MAX_DEVICES 64
static std::mutex g_noalloc_mu;
static unsigned char *g_noalloc_ptr[GRS_NOALLOC_MAX_DEVICES];

/* Device address of `scratch` on `device`, looked up once per device with
 * that device current (the caller has made it current). */
static unsigned char *noalloc_lookup(int device, char *err, size_t errlen) {
  if (device < 0 || device >= GRS_NOALLOC_MAX_DEVICES) {
    snprintf(err, errlen, "device %d is beyond the %d devices the allocation-free entry supports", device,
             GRS_NOALLOC_MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      snprintf(err, errlen, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
      return nullptr;
    }
    g_noalloc_ptr[device] = static_cast<unsigned char *>(p);
  }
  return g_noalloc_ptr[device];
}

static double nanmin2(double a, double b) {
  if (a != a) return b;
  if (b != b) return a;
  return b < a ? b : a;
}

extern "C" {

const char *grs_version(void) { return "grscore 0.1 (gfx950)"; }

static char g_legacy_err[512];

const char *grs_last_error(const grs_ctx *ctx) { return ctx ? ctx->err : g_legacy_err; }

grs_status grs_create(int32_t device, grs_ctx **out) {
  if (!out) return GRS_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GRS_ERR_DEVICE;
  if (device < 0 || device >= n) return GRS_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return GRS_ERR_DEVICE;
  grs_ctx *c = new grs_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GRS_ERR_DEVICE;
  }
  for (int64_t k = 0; k < 2 * kTimingPairs; k++) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) {
      for (hipEvent_t e : c->tpairs) hipEventDestroy(e);
      hipStreamDestroy(c->st);
      delete c;
      return GRS_ERR_DEVICE;
    }
    c->tpairs.push_back(ev);
  }
  /* the allocation-free entry's scratch on this device (a failed lookup
   * only disables that entry, which then reports the reason) */
  c->noalloc = noalloc_lookup(device, c->noalloc_err, sizeof(c->noalloc_err));
  *out = c;
  return GRS_OK;
}

void grs_destroy(grs_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->st);
  DevBuf *bufs[] = {&c->d_geo, &c->d_hkl, &c->d_rot, &c->d_cs, &c->d_keys, &c->d_q, &c->d_sub, &c->d_valid,
                    &c->d_pos, &c->d_vb, &c->d_scratch, &c->d_ns, &c->d_ni, &c->d_C, &c->d_med, &c->d_rows,
                    &c->d_p1, &c->d_U, &c->d_eu, &c->d_rowstart, &c->d_best, &c->d_evox, &c->d_eu_idx,
                    &c->d_bad, &c->d_x, &c->d_y, &c->d_mask, &c->d_g1, &c->d_edt_v, &c->d_edt_z,
                    &c->d_grain, &c->d_bits, &c->d_tbits, &c->d_rowptr, &c->d_roweoff, &c->d_candq, &c->d_rowbest,
                    &c->d_elist, &c->f_hkl, &c->f_om, &c->f_meas, &c->f_grain, &c->f_goff, &c->f_gr6,
                    &c->f_up, &c->f_res, &c->f_contrib, &c->f_jac, &c->f_out, &c->f_ne, &c->c_cen, &c->c_grid,
                    &c->c_gn, &c->c_out, &c->c_n, &c->c_gid, &c->c_vox, &c->c_ids, &c->c_nout, &c->d_audit,
                    &c->d_bestidx, &c->d_bestq, &c->o_q, &c->o_sub, &c->o_valid, &c->x_pos, &c->x_vb,
                    &c->x_key1, &c->x_ns1, &c->x_ni1, &c->x_top, &c->x_cen, &c->x_grid, &c->x_gn, &c->x_ls,
                    &c->x_lsn, &c->x_owner, &c->x_Rmax, &c->x_q2, &c->x_sub2, &c->x_valid2, &c->x_key2,
                    &c->x_ns2, &c->x_ni2, &c->x_gcount, &c->x_gkey, &c->x_grow, &c->x_fitkey, &c->x_x0,
                    &c->x_fgroup, &c->x_fns, &c->x_fni, &c->x_sidx, &c->x_sC, &c->x_bad, &c->p_fptr,
                    &c->p_fsp, &c->p_flag, &c->p_sid, &c->p_lam, &c->p_dyz, &c->p_dis, &c->p_off, &c->p_cnt,
                    &c->p_rowptr, &c->p_rows, &c->v_vol[0], &c->v_vol[1], &c->v_vol[2], &c->v_vol[3],
                    &c->r_rows, &c->r_counts, &c->r_acc_lin, &c->r_acc_C, &c->t_pos, &c->t_proj, &c->t_units,
                    &c->t_fam, &c->t_dposs, &c->t_glen, &c->t_scratch, &c->t_out, &c->m_rows, &c->m_rptr,
                    &c->m_nm, &c->t_skey, &c->t_su, &c->t_bst, &c->t_slen, &c->t_sorig, &c->v_lin,
                    &c->v_out, &c->f_in, &c->f_cnt, &c->m_cs, &c->m_dis};
  if (c->g_exec) hipGraphExecDestroy(c->g_exec);
  for (auto &kv : c->gi_cache) hipGraphExecDestroy(kv.second);
  for (hipEvent_t e : c->tev) hipEventDestroy(e);
  for (DevBuf *b : bufs) free_buf(*b);
  if (c->h_pin) hipHostFree(c->h_pin);
  for (hipEvent_t e : c->tpairs) hipEventDestroy(e);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
}

static grs_status upload_geo(grs_ctx *c) {
  c->hgeo.H = c->H;
  c->hgeo.nrot = c->nrot;
  c->hgeo.nz = c->snz;
  c->hgeo.ny = c->sny;
  ENSURE(c, c->d_geo, sizeof(grs_dgeo));
  HIPCHK(c, hipMemcpyAsync(c->d_geo.p, &c->hgeo, sizeof(grs_dgeo), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_set_geometry(grs_ctx *c, const grs_geometry *geo, const double *hkl, int64_t nrhkl,
                            const double *rot, int64_t nrot) {
  if (!c) return GRS_ERR_ARG;
  if (!geo || !hkl || !rot) return fail(c, GRS_ERR_ARG, "null argument");
  if (nrhkl < 1 || nrot < 1) return fail(c, GRS_ERR_ARG, "need at least one reflection and one rotation angle");
  if (nrhkl * nrot >= (int64_t)1 << 30) return fail(c, GRS_ERR_ARG, "nrhkl*nrot = %lld events is too large",
                                                   (long long)(nrhkl * nrot));
  if (!(geo->pixelysize > 0.0 && geo->pixelzsize > 0.0)) return fail(c, GRS_ERR_ARG, "pixel sizes must be > 0");
  if (!(geo->detysize >= 1.0 && geo->detzsize >= 1.0 && geo->detysize < 2147483647.0 &&
        geo->detzsize < 2147483647.0))
    return fail(c, GRS_ERR_ARG, "detector size out of range");
  HIPCHK(c, hipSetDevice(c->device));
  grs_dgeo &g = c->hgeo;
  std::memset(&g, 0, sizeof(g));
  std::memcpy(g.S, geo->S, sizeof(g.S));
  std::memcpy(g.B, geo->B, sizeof(g.B));
  std::memcpy(g.R, geo->RotDet, sizeof(g.R));
  g.Lss = geo->Lsam2sou;
  g.Lsd = geo->Lsam2det;
  g.L = geo->Lsam2sou + geo->Lsam2det;
  g.P0y = geo->P0y;
  g.P0z = geo->P0z;
  g.RAO = geo->RotAxisOffset;
  g.dety00 = geo->dety00;
  g.detz00 = geo->detz00;
  g.py = geo->pixelysize;
  g.pz = geo->pixelzsize;
  g.dety0 = geo->dety0;
  g.detz0 = geo->detz0;
  g.detysize = geo->detysize;
  g.detzsize = geo->detzsize;
  g.bsy0 = geo->BeamStopY[0];
  g.bsy1 = geo->BeamStopY[1];
  g.bsz0 = geo->BeamStopZ[0];
  g.bsz1 = geo->BeamStopZ[1];
  g.beta_max = ((90.0 + geo->thetamax * 4.0) / 180.0) * GRS_PI;
  g.lambda_min = geo->lambda_min;
  g.lambda_max = geo->lambda_max;
  g.pio2 = GRS_PI / 2.0;
  g.thr = geo->minEucDis / ((geo->pixelysize + geo->pixelzsize) / 2.0);
  g.cos_beta_max = std::cos(g.beta_max);
  g.ipy = 1.0 / g.py;
  g.ipz = 1.0 / g.pz;
  /* omega solve: uniform ascending rotation angles spanning at most one turn
   * (plus two steps); corners of the detector in the frame of SamposW:
   * D0 + Y r2 + Z r3 with D0 = (Lsd, dety00 - RAO, detz00) and (Y, Z) the
   * detector coordinates of pixel positions 2 pixels outside the area */
  g.theta_win = g.beta_max - g.pio2;
  g.om_solve = 0;
  if (nrot >= 2) {
    const double step = (rot[nrot - 1] - rot[0]) / (double)(nrot - 1);
    bool uni = step > 0.0 && (rot[nrot - 1] - rot[0]) <= 360.0 + 2.0 * step;
    for (int64_t i = 0; uni && i < nrot; i++)
      uni = std::fabs(rot[i] - (rot[0] + (double)i * step)) <= 1e-9 * (1.0 + std::fabs(rot[i]));
    if (uni) {
      g.om_solve = 1;
      g.om0 = rot[0] * (GRS_PI / 180.0);
      g.dom = step * (GRS_PI / 180.0);
      g.om_last = rot[nrot - 1] * (GRS_PI / 180.0);
    }
  }
  {
    const double r2[3] = {g.R[1], g.R[4], g.R[7]}, r3[3] = {g.R[2], g.R[5], g.R[8]};
    const double D0[3] = {g.Lsd, g.dety00 - g.RAO, g.detz00};
    const double ys[2] = {0.5 - 2.0, g.detysize + 0.5 + 2.0}, zs[2] = {0.5 - 2.0, g.detzsize + 0.5 + 2.0};
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const double Y = (g.dety0 - ys[a]) * g.py, Z = (g.detz0 - zs[b]) * g.pz;
        for (int k = 0; k < 3; k++) g.corner[3 * (2 * a + b) + k] = D0[k] + Y * r2[k] + Z * r3[k];
      }
  }
  hit_key_for(g.thr, &g.hit_key, &g.has_hit);
  c->H = (int32_t)nrhkl;
  c->nrot = (int32_t)nrot;
  c->bits_valid = false;
  c->hkl.assign(hkl, hkl + 3 * nrhkl);
  ENSURE(c, c->d_hkl, sizeof(double) * 3 * nrhkl);
  ENSURE(c, c->d_rot, sizeof(double) * nrot);
  ENSURE(c, c->d_cs, sizeof(double) * 2 * nrot);
  HIPCHK(c, hipMemcpyAsync(c->d_hkl.p, hkl, sizeof(double) * 3 * nrhkl, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_rot.p, rot, sizeof(double) * nrot, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_omega_cs(P<double>(c->d_rot), (int32_t)nrot, P<double>(c->d_cs), c->st));
  grs_status s = upload_geo(c);
  if (s != GRS_OK) return s;
  c->geo_set = true;
  c->geo_gen += 1;
  return GRS_OK;
}

grs_status grs_set_distance_stack(grs_ctx *c, const float *proj, int64_t nz, int64_t ny, int64_t nrot,
                                  int32_t on_device) {
  if (!c) return GRS_ERR_ARG;
  if (!proj || nz < 1 || ny < 1 || nrot < 1) return fail(c, GRS_ERR_ARG, "invalid distance stack shape");
  if (nz > ((int64_t)1 << 32) / ny / nrot)
    return fail(c, GRS_ERR_ARG, "distance stack of %lld x %lld x %lld exceeds 2^32 entries", (long long)nz,
                (long long)ny, (long long)nrot);
  HIPCHK(c, hipSetDevice(c->device));
  int64_t n = nz * ny * nrot;
  ENSURE(c, c->d_keys, sizeof(uint32_t) * n);
  HIPCHK(c, hipMemcpyAsync(c->d_keys.p, proj, sizeof(float) * n,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
  ENSURE(c, c->d_bad, sizeof(uint32_t));
  HIPCHK(c, hipMemsetAsync(c->d_bad.p, 0, sizeof(uint32_t), c->st));
  HIPCHK(c, grs_launch_check_keys(P<uint32_t>(c->d_keys), n, P<uint32_t>(c->d_bad), c->st));
  uint32_t bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (bad) {
    c->stack_set = false;
    return fail(c, GRS_ERR_ARG, "distance stack holds negative or NaN values");
  }
  c->snz = nz;
  c->sny = ny;
  c->snrot = nrot;
  c->stack_set = true;
  c->bits_valid = false;
  if (c->geo_set) return upload_geo(c);
  return GRS_OK;
}

static grs_status check_ready(grs_ctx *c) {
  if (!c->geo_set) return fail(c, GRS_ERR_STATE, "geometry not set");
  if (!c->stack_set) return fail(c, GRS_ERR_STATE, "distance stack not set");
  if ((double)c->snz < c->hgeo.detzsize || (double)c->sny < c->hgeo.detysize || c->snrot < c->nrot)
    return fail(c, GRS_ERR_ARG,
                "distance stack %lld x %lld x %lld is smaller than detzsize x detysize x nrot = %g x %g x %d",
                (long long)c->snz, (long long)c->sny, (long long)c->snrot, c->hgeo.detzsize, c->hgeo.detysize,
                c->nrot);
  if (c->hgeo.nz != c->snz || c->hgeo.ny != c->sny) return upload_geo(c);
  return GRS_OK;
}

/* Runs the scoring kernel on evaluations already described on the device.
 * d_pos holds V voxel positions; d_sub/d_valid the orientations. */
/* ---- launch autotuning ------------------------------------------------------ */

/* configurations: waves per block x block-to-evaluation mapping, and for
 * the counts-only kinds x hit-test source (bit mask, keys, tiled bit mask) */
static int tune_configs(const grs_ctx *c, int kind, tune_cfg *out) {
  static const int32_t wpbs[4] = {8, 4, 2, 1};
  int n = 0;
  for (int32_t x = 0; x < 2; x++)
    for (int32_t m = 0; m < (kind >= 1 ? 3 : 1); m++) {
      if (m == GRS_MASK_TILED && !c->tbits_ok) continue;
      for (int w = 0; w < 4; w++) out[n++] = tune_cfg{wpbs[w], m, x};
    }
  return n;
}

static void apply_cfg(grs_eval_args *a, const tune_cfg &k) {
  a->count_keys = k.mask == GRS_MASK_KEYS;
  a->bits_tiled = k.mask == GRS_MASK_TILED;
  a->xcd_map = k.xcd;
}

/* evaluations per tuning part: enough to keep 12 resident waves per CU
 * busy for four evaluations (the 17 parts of the median kernel then fit a
 * 262,144-orientation seeding call) */
static int64_t tune_part_min(const grs_ctx *c) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus <= 0) cus = 256;
  return (int64_t)cus * 12 * 4;
}

static uint64_t tune_key(const grs_ctx *c, int kind) {
  return (uint64_t)kind | ((uint64_t)c->H << 4) | ((uint64_t)c->nrot << 16);
}

/* the configuration a (kind, H, nrot) class runs with, encoded for graph keys */
static uint64_t tune_code(const grs_ctx *c, int kind) {
  auto it = c->tunes.find(tune_key(c, kind));
  if (it == c->tunes.end() || !it->second.done) return 0;
  const tune_cfg &k = it->second.cfg;
  return 1 + (uint64_t)k.wpb * 16 + (uint64_t)k.mask * 2 + (uint64_t)k.xcd;
}

/* Launches k_fast over evaluations [0, n_eval) of a (kind 0: with median,
 * kind 1: counts-only CSR, kind 2: counts-only grid).  The first call of a (kind, H, nrot) class with
 * enough evaluations splits the range into 2*ncfg+1 equal parts: the first
 * warms the caches with the default configuration, the others run the
 * configurations in the order 0..ncfg-1, ncfg-1..0 (so that each sees an
 * early and a late part); every part is timed with HIP events and the
 * configuration with the least time per evaluation is kept.  Every part computes
 * its own evaluations, so the results do not depend on the configuration. */
static grs_status launch_tuned(grs_ctx *c, grs_eval_args a, int kind, int64_t waves) {
  grs_tune_state &t = c->tunes[tune_key(c, kind)];
  tune_cfg cfgs[GRS_TUNE_MAX];
  const int ncfg = tune_configs(c, kind, cfgs);
  const int64_t E = a.n_eval;
  auto launch = [&](const grs_eval_args &x, int32_t wpb) -> hipError_t {
    return kind >= 1 ? grs_launch_fast_counts(&x, c->H, waves, wpb, c->st) : grs_launch_fast(&x, c->H, waves, wpb, c->st);
  };
  if (!c->autotune || t.done || E < (int64_t)(2 * ncfg + 1) * tune_part_min(c)) {
    tune_cfg k = t.done ? t.cfg : tune_cfg{8, GRS_MASK_BITS, 0};
    if (k.mask == GRS_MASK_TILED && !c->tbits_ok) k.mask = GRS_MASK_BITS;
    apply_cfg(&a, k);
    /* fewer waves than CUs (the steps of the orientation fits): one wave
     * per block, so that every evaluation has a CU to itself */
    const int32_t wpb = waves < grs_fast_wave_slots(c->device) / 64 ? 1 : k.wpb;
    HIPCHK(c, launch(a, wpb));
    return GRS_OK;
  }
  /* part 0 warms caches with the default configuration and is not timed;
   * parts 1 .. 2*ncfg run the configurations in mirrored order */
  const int nparts = 2 * ncfg + 1;
  while ((int)c->tev.size() < nparts + 1) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->tev.push_back(e);
  }
  const int64_t part = E / nparts;
  double ms[GRS_TUNE_MAX] = {0};
  int64_t cnt[GRS_TUNE_MAX] = {0};
  int cfg_of[2 * GRS_TUNE_MAX + 1];
  for (int p = 0; p < nparts; p++) {
    const int q = p - 1;
    const int k = p == 0 ? 0 : (q < ncfg ? q : 2 * ncfg - 1 - q);
    cfg_of[p] = k;
    grs_eval_args x = a;
    x.e0 = (int64_t)p * part;
    x.n_eval = p == nparts - 1 ? E : (int64_t)(p + 1) * part;
    apply_cfg(&x, cfgs[k]);
    HIPCHK(c, hipEventRecord(c->tev[p], c->st));
    HIPCHK(c, launch(x, cfgs[k].wpb));
  }
  HIPCHK(c, hipEventRecord(c->tev[nparts], c->st));
  HIPCHK(c, hipEventSynchronize(c->tev[nparts]));
  for (int p = 1; p < nparts; p++) {
    float f = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&f, c->tev[p], c->tev[p + 1]));
    const int64_t n = (p == nparts - 1 ? E : (int64_t)(p + 1) * part) - (int64_t)p * part;
    ms[cfg_of[p]] += f;
    cnt[cfg_of[p]] += n;
  }
  int best = 0;
  for (int k = 0; k < GRS_TUNE_MAX; k++) t.ms_per_eval[k] = 0.0;
  for (int k = 0; k < ncfg; k++) {
    t.ms_per_eval[k] = ms[k] / (double)(cnt[k] > 0 ? cnt[k] : 1);
    if (t.ms_per_eval[k] < t.ms_per_eval[best]) best = k;
  }
  t.cfg = cfgs[best];
  t.ncfg = ncfg;
  t.done = true;
  return GRS_OK;
}

static bool tune_pending(const grs_ctx *c, int kind, int64_t E, int64_t waves) {
  if (!c->autotune) return false;
  auto it = c->tunes.find(tune_key(c, kind));
  if (it != c->tunes.end() && it->second.done) return false;
  tune_cfg cfgs[GRS_TUNE_MAX];
  const int ncfg = tune_configs(c, kind, cfgs);
  (void)waves;
  return E >= (int64_t)(2 * ncfg + 1) * tune_part_min(c);
}

static grs_status run_eval(grs_ctx *c, int32_t map, int64_t n_eval, int64_t V, const int32_t *d_evox,
                           const int32_t *d_eu, int32_t method, int32_t *d_ns, int32_t *d_ni, double *d_C,
                           float *d_med, bool bounds_ready = false) {
  if (n_eval <= 0) return GRS_OK;
  const int64_t stride = (int64_t)c->nrot * c->H;
  grs_eval_args a;
  std::memset(&a, 0, sizeof(a));
  a.g = P<grs_dgeo>(c->d_geo);
  a.hkl = P<double>(c->d_hkl);
  a.cs = P<double>(c->d_cs);
  a.pos = P<double>(c->d_pos);
  a.SUB = P<double>(c->d_sub);
  a.valid = P<int32_t>(c->d_valid);
  a.keys = P<uint32_t>(c->d_keys);
  a.n_eval = n_eval;
  a.map = map;
  a.eval_vox = d_evox;
  a.eval_u = d_eu;
  a.scratch_stride = stride;
  a.out_ns = d_ns;
  a.out_ni = d_ni;
  a.out_C = d_C;
  a.out_med = d_med;
  if (method == GRS_METHOD_FAST) {
    if (c->H > 240) return fail(c, GRS_ERR_ARG, "the fast method supports at most 240 reflections (got %d)", c->H);
    if (!bounds_ready) { /* else d_vb holds the bounds of these V positions */
      ENSURE(c, c->d_vb, sizeof(double) * GRS_VB * (size_t)V);
      HIPCHK(c, grs_launch_voxel_bounds(P<grs_dgeo>(c->d_geo), P<double>(c->d_pos), V, P<double>(c->d_cs), c->nrot,
                                        P<double>(c->d_vb), c->st));
    }
    int64_t waves = grs_fast_wave_slots(c->device);
    if (waves > n_eval) waves = ((n_eval + 3) / 4) * 4;
    ENSURE(c, c->d_scratch, sizeof(uint32_t) * (size_t)waves * stride);
    a.vb = P<double>(c->d_vb);
    a.scratch = P<uint32_t>(c->d_scratch);
    a.nrot = c->nrot;
    const bool audit = c->audit && map == GRS_MAP_ONE_VOXEL;
    if (audit) a.audit = P<unsigned long long>(c->d_audit);
    if (c->timing) HIPCHK(c, timing_begin(c));
    if (audit) {
      HIPCHK(c, grs_launch_fast_audit(&a, c->H, waves, c->st));
    } else {
      grs_status ts = launch_tuned(c, a, 0, waves);
      if (ts != GRS_OK) return ts;
    }
    if (c->timing) HIPCHK(c, timing_end(c));
  } else if (method == GRS_METHOD_PLAIN) {
    int64_t T = n_eval < 32768 ? n_eval : 32768;
    ENSURE(c, c->d_scratch, sizeof(uint32_t) * (size_t)T * stride);
    a.scratch = P<uint32_t>(c->d_scratch);
    if (c->timing) HIPCHK(c, timing_begin(c));
    HIPCHK(c, grs_launch_plain(&a, T, c->st));
    if (c->timing) HIPCHK(c, timing_end(c));
  } else {
    return fail(c, GRS_ERR_ARG, "unknown method %d", method);
  }
  return GRS_OK;
}

static grs_status forward_comp_impl(grs_ctx *c, const double *pos, const double *q, int64_t N, bool q_dev,
                                    int32_t *ns, int32_t *ni, double *C, float *med, bool out_dev,
                                    int32_t method) {
  if (!c) return GRS_ERR_ARG;
  if (N < 0 || !pos || (N > 0 && (!q || !ns || !ni || !C || !med))) return fail(c, GRS_ERR_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (N == 0) return GRS_OK;
  const double *d_q = q;
  if (!q_dev) {
    ENSURE(c, c->d_q, sizeof(double) * 4 * N);
    HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, sizeof(double) * 4 * N, hipMemcpyHostToDevice, c->st));
    d_q = P<double>(c->d_q);
  }
  ENSURE(c, c->d_sub, sizeof(double) * 9 * N);
  ENSURE(c, c->d_valid, sizeof(int32_t) * N);
  ENSURE(c, c->d_pos, sizeof(double) * 3);
  HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), d_q, N, P<double>(c->d_sub), P<int32_t>(c->d_valid),
                                  c->st));
  int32_t *d_ns = ns, *d_ni = ni;
  double *d_C = C;
  float *d_med = med;
  if (!out_dev) {
    ENSURE(c, c->d_ns, sizeof(int32_t) * N);
    ENSURE(c, c->d_ni, sizeof(int32_t) * N);
    ENSURE(c, c->d_C, sizeof(double) * N);
    ENSURE(c, c->d_med, sizeof(float) * N);
    d_ns = P<int32_t>(c->d_ns);
    d_ni = P<int32_t>(c->d_ni);
    d_C = P<double>(c->d_C);
    d_med = P<float>(c->d_med);
  }
  s = run_eval(c, GRS_MAP_ONE_VOXEL, N, 1, nullptr, nullptr, method, d_ns, d_ni, d_C, d_med);
  if (s != GRS_OK) return s;
  if (!out_dev) {
    HIPCHK(c, hipMemcpyAsync(ns, d_ns, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ni, d_ni, sizeof(int32_t) * N, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(C, d_C, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(med, d_med, sizeof(float) * N, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  return GRS_OK;
}

grs_status grs_forward_comp(grs_ctx *c, const double *pos, const double *q, int64_t N, int32_t *ns, int32_t *ni,
                            double *C, float *med, int32_t method) {
  return forward_comp_impl(c, pos, q, N, false, ns, ni, C, med, false, method);
}

grs_status grs_forward_comp_device(grs_ctx *c, const double *pos, const double *d_q, int64_t N, int32_t *d_ns,
                                   int32_t *d_ni, double *d_C, float *d_med, int32_t method) {
  return forward_comp_impl(c, pos, d_q, N, true, d_ns, d_ni, d_C, d_med, true, method);
}

/* Layout of the allocation-free scratch (byte offsets, 256-aligned):
 *   0      voxel position (3 doubles)
 *   256    omega-solve bounds of the voxel (GRS_VB doubles)
 *   512    S*U*B of a chunk (9 doubles per orientation)
 *   ...    valid flags of a chunk (int32 per orientation)
 *   keys   per-wave median key lists, nrot*H uint32 each
 * Every region is written before it is read within one call (position and
 * bounds first, then per chunk the S*U*B rows, then the kernel, whose waves
 * fill their key lists before the median reads them), so nothing has to be
 * cleared between calls. */
static const int64_t kNaPos = 0, kNaVb = 256, kNaSub = 512;
static const int64_t kNaValid = kNaSub + 72 * GRS_NOALLOC_CHUNK;
static const int64_t kNaKeys = ((kNaValid + 4 * GRS_NOALLOC_CHUNK) + 255) & ~(int64_t)255;
static_assert(GRS_VB * 8 <= kNaSub - kNaVb, "voxel bounds overflow their slot");
static_assert(kNaKeys < GRS_NOALLOC_SCRATCH_BYTES, "scratch too small for the chunk buffers");

grs_status grs_forward_comp_device_noalloc(grs_ctx *c, const double *pos, const double *d_q, int64_t N,
                                           int32_t *d_ns, int32_t *d_ni, double *d_C, float *d_med) {
  if (!c) return GRS_ERR_ARG;
  if (N < 0 || !pos || (N > 0 && (!d_q || !d_ns || !d_ni || !d_C || !d_med))) return fail(c, GRS_ERR_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (!c->noalloc) return fail(c, GRS_ERR_DEVICE, "allocation-free scratch unavailable: %s", c->noalloc_err);
  if (c->H > 240) return fail(c, GRS_ERR_ARG, "the fast method supports at most 240 reflections (got %d)", c->H);
  const int64_t stride = (int64_t)c->nrot * c->H;
  int64_t wmax = (GRS_NOALLOC_SCRATCH_BYTES - kNaKeys) / (4 * stride);
  const int64_t slots = grs_fast_wave_slots(c->device);
  if (wmax > slots) wmax = slots;
  wmax &= ~(int64_t)3;
  if (wmax < 4)
    return fail(c, GRS_ERR_ARG,
                "nrot*nrhkl = %lld events per evaluation: the %lld-byte allocation-free scratch holds fewer than 4 "
                "waves' key lists (at most %lld events)",
                (long long)stride, (long long)GRS_NOALLOC_SCRATCH_BYTES,
                (long long)((GRS_NOALLOC_SCRATCH_BYTES - kNaKeys) / 16));
  if (N == 0) return GRS_OK;
  unsigned char *base = c->noalloc;
  double *s_pos = reinterpret_cast<double *>(base + kNaPos);
  double *s_vb = reinterpret_cast<double *>(base + kNaVb);
  double *s_sub = reinterpret_cast<double *>(base + kNaSub);
  int32_t *s_valid = reinterpret_cast<int32_t *>(base + kNaValid);
  uint32_t *s_keys = reinterpret_cast<uint32_t *>(base + kNaKeys);
  /* the launch configuration the autotuner chose for this class, if any
   * (no tuning here: it records events) */
  tune_cfg kc = {8, GRS_MASK_BITS, 0};
  auto it = c->tunes.find(tune_key(c, 0));
  if (it != c->tunes.end() && it->second.done) kc = it->second.cfg;
  HIPCHK(c, hipMemcpyAsync(s_pos, pos, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_voxel_bounds(P<grs_dgeo>(c->d_geo), s_pos, 1, P<double>(c->d_cs), c->nrot, s_vb, c->st));
  if (c->timing) HIPCHK(c, timing_begin(c));
  for (int64_t n0 = 0; n0 < N; n0 += GRS_NOALLOC_CHUNK) {
    const int64_t n = N - n0 < GRS_NOALLOC_CHUNK ? N - n0 : GRS_NOALLOC_CHUNK;
    HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), d_q + 4 * n0, n, s_sub, s_valid, c->st));
    grs_eval_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = P<grs_dgeo>(c->d_geo);
    a.hkl = P<double>(c->d_hkl);
    a.cs = P<double>(c->d_cs);
    a.pos = s_pos;
    a.SUB = s_sub;
    a.valid = s_valid;
    a.keys = P<uint32_t>(c->d_keys);
    a.n_eval = n;
    a.map = GRS_MAP_ONE_VOXEL;
    a.scratch_stride = stride;
    a.scratch = s_keys;
    a.vb = s_vb;
    a.nrot = c->nrot;
    apply_cfg(&a, kc);
    a.out_ns = d_ns + n0;
    a.out_ni = d_ni + n0;
    a.out_C = d_C + n0;
    a.out_med = d_med + n0;
    int64_t waves = wmax;
    if (waves > n) waves = ((n + 3) / 4) * 4;
    HIPCHK(c, grs_launch_fast(&a, c->H, waves, kc.wpb, c->st));
  }
  if (c->timing) HIPCHK(c, timing_end(c));
  return GRS_OK;
}

/* voxel rows per device chunk (bounds the per-(voxel, omega) table) */
/* Voxels per launch of the grow / compete entries: device buffers per voxel
 * are a few dozen bytes (positions, bounds, outputs), so chunks are large. */
static int64_t voxel_chunk(const grs_ctx *c) {
  (void)c;
  return (int64_t)1 << 22;
}

grs_status grs_grow_indexed_region(grs_ctx *c, const double *vox, int64_t M, const double *U, const double *p1,
                                   double *out, int32_t method) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || (M > 0 && (!vox || !U || !p1 || !out))) return fail(c, GRS_ERR_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (M == 0) return GRS_OK;
  /* only voxels inside the mask are scored */
  std::vector<int64_t> idx;
  idx.reserve(M);
  for (int64_t i = 0; i < M; i++) {
    double *o = out + 7 * i;
    o[0] = vox[6 * i];
    o[1] = vox[6 * i + 1];
    o[2] = vox[6 * i + 2];
    o[3] = o[4] = o[5] = o[6] = 0.0;
    if (vox[6 * i + 4] > 0.0) idx.push_back(i);
  }
  ENSURE(c, c->d_U, sizeof(double) * 9);
  ENSURE(c, c->d_sub, sizeof(double) * 9);
  ENSURE(c, c->d_valid, sizeof(int32_t));
  ENSURE(c, c->d_p1, sizeof(double) * 14);
  HIPCHK(c, hipMemcpyAsync(c->d_U.p, U, sizeof(double) * 9, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_p1.p, p1, sizeof(double) * 14, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), 1, P<double>(c->d_sub),
                                  P<int32_t>(c->d_valid), c->st));
  const int64_t chunk = voxel_chunk(c);
  std::vector<double> rows;
  std::vector<int32_t> hns, hni;
  std::vector<double> hC;
  std::vector<float> hmed;
  for (int64_t b = 0; b < (int64_t)idx.size(); b += chunk) {
    int64_t m = (int64_t)idx.size() - b;
    if (m > chunk) m = chunk;
    rows.resize(3 * m);
    for (int64_t k = 0; k < m; k++)
      for (int j = 0; j < 3; j++) rows[3 * k + j] = vox[6 * idx[b + k] + j];
    ENSURE(c, c->d_rows, sizeof(double) * 3 * m);
    ENSURE(c, c->d_pos, sizeof(double) * 3 * m);
    ENSURE(c, c->d_ns, sizeof(int32_t) * m);
    ENSURE(c, c->d_ni, sizeof(int32_t) * m);
    ENSURE(c, c->d_C, sizeof(double) * m);
    ENSURE(c, c->d_med, sizeof(float) * m);
    HIPCHK(c, hipMemcpyAsync(c->d_rows.p, rows.data(), sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_voxel_pos(P<double>(c->d_rows), m, 3, P<double>(c->d_p1), P<double>(c->d_pos), c->st));
    s = run_eval(c, GRS_MAP_ONE_U, m, m, nullptr, nullptr, method, P<int32_t>(c->d_ns), P<int32_t>(c->d_ni),
                 P<double>(c->d_C), P<float>(c->d_med));
    if (s != GRS_OK) return s;
    hns.resize(m);
    hni.resize(m);
    hC.resize(m);
    hmed.resize(m);
    HIPCHK(c, hipMemcpyAsync(hns.data(), c->d_ns.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(hni.data(), c->d_ni.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(hC.data(), c->d_C.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(hmed.data(), c->d_med.p, sizeof(float) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int64_t k = 0; k < m; k++) {
      const double *v = vox + 6 * idx[b + k];
      double *o = out + 7 * idx[b + k];
      double Cn = hC[k], dm = (double)hmed[k];
      double lim = nanmin2(v[5], p1[13]);
      bool flag = (Cn > v[3] && dm <= lim) && (Cn >= p1[11] * (1.0 - p1[12]));
      o[3] = (double)hns[k];
      o[4] = (double)hni[k];
      o[5] = dm;
      o[6] = flag ? 1.0 : 0.0;
    }
  }
  return GRS_OK;
}

grs_status grs_forward_comp_compete(grs_ctx *c, const double *vox, int64_t M, const double *id_all, int64_t n_id,
                                    const double *euler, int64_t G, const double *p1, double *out,
                                    int32_t method) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || n_id < 0 || G < 0 || (M > 0 && (!vox || !p1 || !out)) || (n_id > 0 && !id_all) ||
      (G > 0 && !euler))
    return fail(c, GRS_ERR_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (M == 0) return GRS_OK;
  for (int64_t j = 0; j < n_id; j++) {
    double g = id_all[j];
    if (!(g >= 1.0 && g <= (double)G && g == std::floor(g)))
      return fail(c, GRS_ERR_ARG, "id_all[%lld] = %g is not a grain id in 1..%lld", (long long)j, g, (long long)G);
  }
  for (int64_t i = 0; i < M; i++) {
    double a = vox[5 * i + 3], b = vox[5 * i + 4];
    if (!(a == std::floor(a) && b == std::floor(b) && a >= 1.0 && b <= (double)n_id && b >= a - 1.0))
      return fail(c, GRS_ERR_ARG, "voxel row %lld has candidate range [%g, %g] outside 1..%lld", (long long)i, a, b,
                  (long long)n_id);
  }
  ENSURE(c, c->d_eu, sizeof(double) * 3 * (G > 0 ? G : 1));
  ENSURE(c, c->d_sub, sizeof(double) * 9 * (G > 0 ? G : 1));
  ENSURE(c, c->d_valid, sizeof(int32_t) * (G > 0 ? G : 1));
  ENSURE(c, c->d_p1, sizeof(double) * 11);
  if (G > 0) {
    HIPCHK(c, hipMemcpyAsync(c->d_eu.p, euler, sizeof(double) * 3 * G, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_sub_from_euler(P<grs_dgeo>(c->d_geo), P<double>(c->d_eu), G, P<double>(c->d_sub),
                                        P<int32_t>(c->d_valid), c->st));
  }
  HIPCHK(c, hipMemcpyAsync(c->d_p1.p, p1, sizeof(double) * 11, hipMemcpyHostToDevice, c->st));
  const int64_t chunk = voxel_chunk(c);
  std::vector<double> rows;
  std::vector<int32_t> evox, eu;
  std::vector<int64_t> rs;
  std::vector<int64_t> best;
  std::vector<int32_t> hni;
  std::vector<double> hC;
  std::vector<float> hmed;
  for (int64_t b = 0; b < M; b += chunk) {
    int64_t m = M - b;
    if (m > chunk) m = chunk;
    rows.resize(3 * m);
    rs.assign(m + 1, 0);
    evox.clear();
    eu.clear();
    for (int64_t k = 0; k < m; k++) {
      const double *v = vox + 5 * (b + k);
      rows[3 * k] = v[0];
      rows[3 * k + 1] = v[1];
      rows[3 * k + 2] = v[2];
      int64_t lo = (int64_t)v[3], hi = (int64_t)v[4];
      for (int64_t j = lo; j <= hi; j++) {
        evox.push_back((int32_t)k);
        eu.push_back((int32_t)id_all[j - 1] - 1);
      }
      rs[k + 1] = (int64_t)evox.size();
    }
    int64_t E = (int64_t)evox.size();
    ENSURE(c, c->d_rows, sizeof(double) * 3 * m);
    ENSURE(c, c->d_pos, sizeof(double) * 3 * m);
    ENSURE(c, c->d_rowstart, sizeof(int64_t) * (m + 1));
    ENSURE(c, c->d_best, sizeof(int64_t) * m);
    ENSURE(c, c->d_evox, sizeof(int32_t) * (E > 0 ? E : 1));
    ENSURE(c, c->d_eu_idx, sizeof(int32_t) * (E > 0 ? E : 1));
    ENSURE(c, c->d_ns, sizeof(int32_t) * (E > 0 ? E : 1));
    ENSURE(c, c->d_ni, sizeof(int32_t) * (E > 0 ? E : 1));
    ENSURE(c, c->d_C, sizeof(double) * (E > 0 ? E : 1));
    ENSURE(c, c->d_med, sizeof(float) * (E > 0 ? E : 1));
    HIPCHK(c, hipMemcpyAsync(c->d_rows.p, rows.data(), sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_rowstart.p, rs.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, c->st));
    if (E > 0) {
      HIPCHK(c, hipMemcpyAsync(c->d_evox.p, evox.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(c->d_eu_idx.p, eu.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, c->st));
    }
    HIPCHK(c, grs_launch_voxel_pos(P<double>(c->d_rows), m, 3, P<double>(c->d_p1), P<double>(c->d_pos), c->st));
    s = run_eval(c, GRS_MAP_PAIRS, E, m, P<int32_t>(c->d_evox), P<int32_t>(c->d_eu_idx), method,
                 P<int32_t>(c->d_ns), P<int32_t>(c->d_ni), P<double>(c->d_C), P<float>(c->d_med));
    if (s != GRS_OK) return s;
    HIPCHK(c, grs_launch_compete_reduce(P<int64_t>(c->d_rowstart), m, P<int32_t>(c->d_ns), P<int32_t>(c->d_ni),
                                        P<double>(c->d_C), P<float>(c->d_med), P<int64_t>(c->d_best), c->st));
    best.resize(m);
    hni.resize(E > 0 ? E : 1);
    hC.resize(E > 0 ? E : 1);
    hmed.resize(E > 0 ? E : 1);
    HIPCHK(c, hipMemcpyAsync(best.data(), c->d_best.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->st));
    if (E > 0) {
      HIPCHK(c, hipMemcpyAsync(hni.data(), c->d_ni.p, sizeof(int32_t) * E, hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipMemcpyAsync(hC.data(), c->d_C.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipMemcpyAsync(hmed.data(), c->d_med.p, sizeof(float) * E, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int64_t k = 0; k < m; k++) {
      double *o = out + 8 * (b + k);
      o[0] = (double)(b + k + 1);
      int64_t j = best[k];
      if (j < 0) {
        for (int t = 1; t < 8; t++) o[t] = 0.0;
        continue;
      }
      int32_t gi = eu[j];
      o[1] = euler[3 * gi];
      o[2] = euler[3 * gi + 1];
      o[3] = euler[3 * gi + 2];
      o[4] = (double)(gi + 1);
      o[5] = hC[j];
      o[6] = (double)hni[j];
      o[7] = (double)hmed[j];
    }
  }
  return GRS_OK;
}

/* hit mask (1 bit per pixel) of the current stack and threshold */
static grs_status ensure_bits(grs_ctx *c) {
  if (c->bits_valid) return GRS_OK;
  const int64_t n = c->snz * c->sny * c->snrot;
  ENSURE(c, c->d_bits, sizeof(uint32_t) * (size_t)((n + 31) / 32));
  HIPCHK(c, grs_launch_pack_hits(P<uint32_t>(c->d_keys), n, c->hgeo.hit_key, c->hgeo.has_hit, P<uint32_t>(c->d_bits),
                                 c->st));
  /* the same mask in 32 x 32 pixel tiles, when its bit numbers fit 32 bits */
  const int64_t tw = ((c->snz + 31) / 32) * ((c->sny + 31) / 32) * c->snrot * 32;
  c->tbits_ok = tw * 32 <= (int64_t)UINT32_MAX + 1;
  if (c->tbits_ok) {
    ENSURE(c, c->d_tbits, sizeof(uint32_t) * (size_t)tw);
    HIPCHK(c, grs_launch_pack_hits_tiled(P<uint32_t>(c->d_keys), c->snz, c->sny, c->snrot, c->hgeo.hit_key,
                                         c->hgeo.has_hit, P<uint32_t>(c->d_tbits), c->st));
  }
  c->bits_valid = true;
  return GRS_OK;
}

grs_status grs_compete_csr(grs_ctx *c, const double *pos, int64_t M, const int64_t *row_ptr, const double *cand_q,
                           int64_t K, int32_t nr, double gridsize, int64_t *best, double *best_q, int32_t *ns,
                           int32_t *ni, double *C, float *med) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || K < 0 || nr < 0 || nr > 40 || !row_ptr || (M > 0 && (!pos || !best || !best_q || !ns || !ni || !C || !med)) ||
      (K > 0 && !cand_q))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  if (row_ptr[0] != 0 || row_ptr[M] != K) return fail(c, GRS_ERR_ARG, "row_ptr must start at 0 and end at K");
  for (int64_t r = 0; r < M; r++)
    if (row_ptr[r + 1] < row_ptr[r]) return fail(c, GRS_ERR_ARG, "row_ptr must be non-decreasing");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if ((int64_t)c->nrot * c->H > GRS_KEY_MAX_EVENTS)
    return fail(c, GRS_ERR_ARG, "nrot*nrhkl = %lld exceeds %d events, the exact range of the packed argmax key",
                (long long)c->nrot * c->H, GRS_KEY_MAX_EVENTS);
  if (c->H > 240) return fail(c, GRS_ERR_ARG, "at most 240 reflections");
  if (M == 0) return GRS_OK;
  s = ensure_bits(c);
  if (s != GRS_OK) return s;
  const int64_t S = (int64_t)(nr + 1) * (nr + 1) * (nr + 1);
  const int64_t chunk = voxel_chunk(c);
  const int64_t stride = (int64_t)c->nrot * c->H;
  std::vector<int64_t> rp, eo;
  for (int64_t r0 = 0; r0 < M; r0 += chunk) {
    const int64_t m = M - r0 < chunk ? M - r0 : chunk;
    const int64_t c0 = row_ptr[r0], kc = row_ptr[r0 + m] - c0;
    rp.resize(m + 1);
    eo.resize(m + 1);
    for (int64_t r = 0; r <= m; r++) {
      rp[r] = row_ptr[r0 + r] - c0;
      eo[r] = rp[r] * S;
    }
    const int64_t E = eo[m];
    ENSURE(c, c->d_pos, sizeof(double) * 3 * m);
    ENSURE(c, c->d_rowptr, sizeof(int64_t) * (m + 1));
    ENSURE(c, c->d_roweoff, sizeof(int64_t) * (m + 1));
    ENSURE(c, c->d_candq, sizeof(double) * 4 * (kc > 0 ? kc : 1));
    ENSURE(c, c->d_rowbest, sizeof(unsigned long long) * m);
    ENSURE(c, c->d_elist, sizeof(int64_t) * m);
    ENSURE(c, c->d_ns, sizeof(int32_t) * m);
    ENSURE(c, c->d_ni, sizeof(int32_t) * m);
    ENSURE(c, c->d_C, sizeof(double) * m);
    ENSURE(c, c->d_med, sizeof(float) * m);
    HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos + 3 * r0, sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_rowptr.p, rp.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_roweoff.p, eo.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, c->st));
    if (kc > 0)
      HIPCHK(c, hipMemcpyAsync(c->d_candq.p, cand_q + 4 * c0, sizeof(double) * 4 * kc, hipMemcpyHostToDevice, c->st));
    ENSURE(c, c->d_vb, sizeof(double) * GRS_VB * (size_t)m);
    ENSURE(c, c->d_bestidx, sizeof(int64_t) * m);
    ENSURE(c, c->d_bestq, sizeof(double) * 4 * m);
    int64_t waves = grs_fast_wave_slots(c->device);
    const int64_t w1 = waves > E ? ((E + 3) / 4) * 4 : waves;
    const int64_t w2 = waves > m ? ((m + 3) / 4) * 4 : waves;
    ENSURE(c, c->d_scratch, sizeof(uint32_t) * (size_t)(w2 > 0 ? w2 : 1) * stride);
    grs_eval_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = P<grs_dgeo>(c->d_geo);
    a.hkl = P<double>(c->d_hkl);
    a.cs = P<double>(c->d_cs);
    a.vb = P<double>(c->d_vb);
    a.pos = P<double>(c->d_pos);
    a.keys = P<uint32_t>(c->d_keys);
    a.hit_bits = P<uint32_t>(c->d_bits);
    a.hit_tbits = P<uint32_t>(c->d_tbits);
    a.scratch_stride = stride;
    a.scratch = P<uint32_t>(c->d_scratch);
    a.csr.row_ptr = P<int64_t>(c->d_rowptr);
    a.csr.row_eoff = P<int64_t>(c->d_roweoff);
    a.csr.cand_q = P<double>(c->d_candq);
    a.csr.nrows = m;
    a.csr.nr = nr;
    a.csr.S = (int32_t)S;
    a.csr.gridsize = gridsize;
    a.csr.row_best = P<unsigned long long>(c->d_rowbest);
    a.csr.elist = P<int64_t>(c->d_elist);
    a.out_ns = P<int32_t>(c->d_ns);
    a.out_ni = P<int32_t>(c->d_ni);
    a.out_C = P<double>(c->d_C);
    a.out_med = P<float>(c->d_med);
    /* the whole chunk on the device: clear keys, bounds, counts-only scoring
     * of every expanded candidate, winners, median re-scoring of the winners,
     * winner index and quaternion */
    auto enqueue = [&]() -> grs_status {
      HIPCHK(c, grs_launch_fill_u64(P<unsigned long long>(c->d_rowbest), m, 0ull, c->st));
      HIPCHK(c, grs_launch_voxel_bounds(P<grs_dgeo>(c->d_geo), P<double>(c->d_pos), m, P<double>(c->d_cs), c->nrot,
                                        P<double>(c->d_vb), c->st));
      grs_status ls;
      if (E > 0) {
        grs_eval_args a1 = a;
        a1.n_eval = E;
        a1.map = GRS_MAP_CSR;
        if ((ls = launch_tuned(c, a1, 1, w1)) != GRS_OK) return ls;
      }
      HIPCHK(c, grs_launch_winner_list(P<unsigned long long>(c->d_rowbest), P<int64_t>(c->d_roweoff), m,
                                       P<int64_t>(c->d_elist), c->st));
      grs_eval_args a2 = a;
      a2.n_eval = m;
      a2.map = GRS_MAP_CSR_LIST;
      if ((ls = launch_tuned(c, a2, 0, w2)) != GRS_OK) return ls;
      HIPCHK(c, grs_launch_winner_out(P<int64_t>(c->d_elist), P<int64_t>(c->d_roweoff), P<int64_t>(c->d_rowptr),
                                      P<double>(c->d_candq), nr, gridsize, m, P<int64_t>(c->d_bestidx),
                                      P<double>(c->d_bestq), c->st));
      return GRS_OK;
    };
    if (c->timing) HIPCHK(c, timing_begin(c));
    /* a chunk that first tunes a launch configuration runs directly: the
     * tuning times launches, which a captured graph cannot */
    const bool tuning = tune_pending(c, 1, E, w1) || tune_pending(c, 0, m, w2);
    if (c->graphs && !tuning) {
      double gs = gridsize;
      uint64_t gbits;
      std::memcpy(&gbits, &gs, sizeof(gbits));
      std::vector<uint64_t> key = {(uint64_t)m, (uint64_t)E, (uint64_t)S, (uint64_t)nr, gbits, (uint64_t)c->H,
                                   (uint64_t)c->nrot, (uint64_t)w1, (uint64_t)w2, (uint64_t)stride,
                                   tune_code(c, 1), tune_code(c, 0), (uint64_t)c->autotune};
      for (const DevBuf *bp : {&c->d_geo, &c->d_hkl, &c->d_cs, &c->d_vb, &c->d_pos, &c->d_keys, &c->d_bits,
                               &c->d_tbits, &c->d_scratch, &c->d_rowptr, &c->d_roweoff, &c->d_candq, &c->d_rowbest, &c->d_elist,
                               &c->d_ns, &c->d_ni, &c->d_C, &c->d_med, &c->d_bestidx, &c->d_bestq})
        key.push_back((uint64_t)(uintptr_t)bp->p);
      if (!c->g_exec || key != c->g_key) {
        if (c->g_exec) {
          HIPCHK(c, hipGraphExecDestroy(c->g_exec));
          c->g_exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        grs_status es = enqueue();
        hipError_t e2 = hipStreamEndCapture(c->st, &graph);
        if (es != GRS_OK || e2 != hipSuccess) {
          if (graph) hipGraphDestroy(graph);
          if (es != GRS_OK) return es;
          return fail(c, GRS_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e2));
        }
        hipError_t e = hipGraphInstantiate(&c->g_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
          c->g_exec = nullptr;
          return fail(c, GRS_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        c->g_key = key;
        c->g_captures += 1;
      } else {
        c->g_replays += 1;
      }
      HIPCHK(c, hipGraphLaunch(c->g_exec, c->st));
    } else {
      grs_status es = enqueue();
      if (es != GRS_OK) return es;
    }
    if (c->timing) HIPCHK(c, timing_end(c));
    HIPCHK(c, hipMemcpyAsync(best + r0, c->d_bestidx.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(best_q + 4 * r0, c->d_bestq.p, sizeof(double) * 4 * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ns + r0, c->d_ns.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(ni + r0, c->d_ni.p, sizeof(int32_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(C + r0, c->d_C.p, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(med + r0, c->d_med.p, sizeof(float) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  return GRS_OK;
}

/* ---- spot-centroid residual and its adjoint ------------------------------ */

grs_status grs_set_spots(grs_ctx *c, int64_t n, const double *hkl, const double *omega, const double *meas,
                         const int32_t *grain, int64_t G) {
  if (!c) return GRS_ERR_ARG;
  if (n < 0 || G < 0 || (n > 0 && (!hkl || !omega || !meas || !grain))) return fail(c, GRS_ERR_ARG, "null argument");
  std::vector<int64_t> goff(G + 1, 0);
  for (int64_t i = 0; i < n; i++) {
    if (grain[i] < 0 || grain[i] >= G) return fail(c, GRS_ERR_ARG, "spot %lld has grain %d outside 0..%lld",
                                                   (long long)i, grain[i], (long long)(G - 1));
    if (i > 0 && grain[i] < grain[i - 1]) return fail(c, GRS_ERR_ARG, "spots must be sorted by grain");
    goff[grain[i] + 1]++;
  }
  for (int64_t g = 0; g < G; g++) goff[g + 1] += goff[g];
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->f_hkl, sizeof(double) * 3 * n);
  ENSURE(c, c->f_om, sizeof(double) * n);
  ENSURE(c, c->f_meas, sizeof(double) * 2 * n);
  ENSURE(c, c->f_grain, sizeof(int32_t) * n);
  ENSURE(c, c->f_goff, sizeof(int64_t) * (G + 1));
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(c->f_hkl.p, hkl, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->f_om.p, omega, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->f_meas.p, meas, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->f_grain.p, grain, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(c->f_goff.p, goff.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->f_n = n;
  c->f_G = G;
  return GRS_OK;
}

/* RotDet = RotZ*RotY*RotX (get_det_R.m) and its derivatives per degree. */
static void det_rot_and_derivs(const double *t, double *R, double dR[3][9]) {
  const double d2r = GRS_PI / 180.0;
  double cx = std::cos(t[0] * d2r), sx = std::sin(t[0] * d2r);
  double cy = std::cos(t[1] * d2r), sy = std::sin(t[1] * d2r);
  double cz = std::cos(t[2] * d2r), sz = std::sin(t[2] * d2r);
  double X[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx}, dX[9] = {0, 0, 0, 0, -sx * d2r, -cx * d2r, 0, cx * d2r, -sx * d2r};
  double Y[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy}, dY[9] = {-sy * d2r, 0, cy * d2r, 0, 0, 0, -cy * d2r, 0, -sy * d2r};
  double Z[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1}, dZ[9] = {-sz * d2r, -cz * d2r, 0, cz * d2r, -sz * d2r, 0, 0, 0, 0};
  auto mm = [](const double *A, const double *B, double *C) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
  };
  double ZY[9], T[9];
  mm(Z, Y, ZY);
  mm(ZY, X, R);
  mm(ZY, dX, dR[0]);
  mm(Z, dY, T);
  mm(T, X, dR[1]);
  mm(dZ, Y, T);
  mm(T, X, dR[2]);
}

/* the constants of the residual kernels (grs_fitc in grs_fit.hip) for geometry geo7 */
struct host_fitc {
  double S[9], B[9];
  double P0y, P0z, RAO, py, pz, dety0, detz0;
  double Lss, Lsd, dy00, dz00;
  double R[9];
  double dR[3][9];
};
static_assert(sizeof(host_fitc) == 8 * (18 + 7 + 4 + 9 + 27), "layout");

static void fill_fitc(const grs_ctx *c, const double *geo7, host_fitc *f) {
  std::memcpy(f->S, c->hgeo.S, sizeof(f->S));
  std::memcpy(f->B, c->hgeo.B, sizeof(f->B));
  f->P0y = c->hgeo.P0y;
  f->P0z = c->hgeo.P0z;
  f->RAO = c->hgeo.RAO;
  f->py = c->hgeo.py;
  f->pz = c->hgeo.pz;
  f->dety0 = c->hgeo.dety0;
  f->detz0 = c->hgeo.detz0;
  f->Lss = geo7[0];
  f->Lsd = geo7[1];
  f->dy00 = geo7[2];
  f->dz00 = geo7[3];
  det_rot_and_derivs(geo7 + 4, f->R, f->dR);
}

grs_status grs_spot_residual(grs_ctx *c, const double *geo7, const double *grains6, double *resid,
                             const double *upstream, double *grad7, double *grad6, double *jac) {
  if (!c) return GRS_ERR_ARG;
  if (!c->geo_set) return fail(c, GRS_ERR_STATE, "geometry not set");
  const int64_t n = c->f_n, G = c->f_G;
  if (!geo7 || (G > 0 && !grains6) || (n > 0 && !resid)) return fail(c, GRS_ERR_ARG, "null argument");
  if ((grad7 == nullptr) != (grad6 == nullptr)) return fail(c, GRS_ERR_ARG, "grad7 and grad6 go together");
  HIPCHK(c, hipSetDevice(c->device));
  if (grs_fitc_size() != sizeof(host_fitc)) return fail(c, GRS_ERR_ARG, "internal layout mismatch");
  host_fitc f;
  fill_fitc(c, geo7, &f);
  if (n == 0) {
    if (grad7) {
      for (int k = 0; k < 7; k++) grad7[k] = 0.0;
      for (int64_t k = 0; k < 6 * G; k++) grad6[k] = 0.0;
    }
    return GRS_OK;
  }
  ENSURE(c, c->f_gr6, sizeof(double) * 6 * (G > 0 ? G : 1));
  ENSURE(c, c->f_res, sizeof(double) * 2 * n);
  HIPCHK(c, hipMemcpyAsync(c->f_gr6.p, grains6, sizeof(double) * 6 * G, hipMemcpyHostToDevice, c->st));
  double *d_up = nullptr, *d_contrib = nullptr, *d_jac = nullptr;
  if (grad7) {
    ENSURE(c, c->f_contrib, sizeof(double) * 13 * n);
    d_contrib = P<double>(c->f_contrib);
    if (upstream) {
      ENSURE(c, c->f_up, sizeof(double) * 2 * n);
      HIPCHK(c, hipMemcpyAsync(c->f_up.p, upstream, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->st));
      d_up = P<double>(c->f_up);
    }
  }
  if (jac) {
    ENSURE(c, c->f_jac, sizeof(double) * 26 * n);
    d_jac = P<double>(c->f_jac);
  }
  if (c->timing) HIPCHK(c, timing_begin(c));
  HIPCHK(c, grs_launch_spot_residual(&f, n, P<double>(c->f_hkl), P<double>(c->f_om), P<double>(c->f_meas),
                                     P<int32_t>(c->f_grain), P<double>(c->f_gr6), d_up, P<double>(c->f_res),
                                     d_contrib, d_jac, c->st));
  if (grad7) {
    ENSURE(c, c->f_out, sizeof(double) * (7 + 6 * G));
    HIPCHK(c, grs_launch_grad_reduce(d_contrib, n, P<int64_t>(c->f_goff), G, P<double>(c->f_out),
                                     P<double>(c->f_out) + 7, c->st));
  }
  if (c->timing) HIPCHK(c, timing_end(c));
  HIPCHK(c, hipMemcpyAsync(resid, c->f_res.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->st));
  if (grad7) {
    HIPCHK(c, hipMemcpyAsync(grad7, c->f_out.p, sizeof(double) * 7, hipMemcpyDeviceToHost, c->st));
    if (G > 0)
      HIPCHK(c, hipMemcpyAsync(grad6, P<double>(c->f_out) + 7, sizeof(double) * 6 * G, hipMemcpyDeviceToHost, c->st));
  }
  if (jac) HIPCHK(c, hipMemcpyAsync(jac, c->f_jac.p, sizeof(double) * 26 * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_spot_normal(grs_ctx *c, const double *geo7, const double *grains6, double delta, double *Agg,
                           double *bg, double *Akk, double *Agk, double *bk, double *sums) {
  if (!c) return GRS_ERR_ARG;
  if (!c->geo_set) return fail(c, GRS_ERR_STATE, "geometry not set");
  const int64_t n = c->f_n, G = c->f_G;
  if (!geo7 || !Agg || !bg || !sums || (G > 0 && (!grains6 || !Akk || !Agk || !bk)))
    return fail(c, GRS_ERR_ARG, "null argument");
  if (!(delta >= 0.0) || !std::isfinite(delta)) return fail(c, GRS_ERR_ARG, "delta must be finite and >= 0");
  HIPCHK(c, hipSetDevice(c->device));
  if (grs_fitc_size() != sizeof(host_fitc)) return fail(c, GRS_ERR_ARG, "internal layout mismatch");
  host_fitc f;
  fill_fitc(c, geo7, &f);
  double out37[37] = {0};
  std::vector<double> pg((size_t)69 * (G > 0 ? G : 1), 0.0);
  if (n > 0 && G > 0) {
    ENSURE(c, c->f_gr6, sizeof(double) * 6 * G);
    ENSURE(c, c->f_ne, sizeof(double) * (69 * G + 37 * G + 37));
    HIPCHK(c, hipMemcpyAsync(c->f_gr6.p, grains6, sizeof(double) * 6 * G, hipMemcpyHostToDevice, c->st));
    double *d_pg = P<double>(c->f_ne), *d_part = d_pg + 69 * G, *d_out = d_part + 37 * G;
    if (c->timing) HIPCHK(c, timing_begin(c));
    HIPCHK(c, grs_launch_spot_normal(&f, G, P<double>(c->f_hkl), P<double>(c->f_om), P<double>(c->f_meas),
                                     P<int64_t>(c->f_goff), P<double>(c->f_gr6), delta, d_pg, d_part, d_out, c->st));
    if (c->timing) HIPCHK(c, timing_end(c));
    HIPCHK(c, hipMemcpyAsync(pg.data(), d_pg, sizeof(double) * 69 * G, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(out37, d_out, sizeof(double) * 37, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  /* expand the upper triangles */
  int k = 69;
  for (int a = 0; a < 7; a++)
    for (int b = a; b < 7; b++, k++) Agg[7 * a + b] = Agg[7 * b + a] = out37[k - 69];
  for (int a = 0; a < 7; a++) bg[a] = out37[28 + a];
  sums[0] = out37[35];
  sums[1] = out37[36];
  for (int64_t g = 0; g < G; g++) {
    const double *src = pg.data() + 69 * g;
    double *A = Akk + 36 * g;
    int m = 0;
    for (int a = 0; a < 6; a++)
      for (int b = a; b < 6; b++, m++) A[6 * a + b] = A[6 * b + a] = src[m];
    for (int a = 0; a < 7; a++)
      for (int b = 0; b < 6; b++, m++) Agk[42 * g + 6 * a + b] = src[m];
    for (int b = 0; b < 6; b++, m++) bk[6 * g + b] = src[m];
  }
  return GRS_OK;
}

/* ---- legacy single-precision operators ---------------------------------- */

static std::mutex g_legacy_mu;
static grs_ctx *g_legacy = nullptr;

static grs_ctx *legacy_ctx() {
  if (!g_legacy) {
    if (grs_create(0, &g_legacy) != GRS_OK) {
      std::snprintf(g_legacy_err, sizeof(g_legacy_err), "no HIP device available");
      g_legacy = nullptr;
    }
  }
  return g_legacy;
}

static grs_status legacy_setup(grs_ctx *c, const float *S, const float *B, const float *hkl, int64_t n_hkl,
                               const float *proj, int64_t n_proj, const float *rot, int64_t n_rot,
                               const float *RotDet, const float *param, int64_t n_param) {
  if (!S || !B || !hkl || !proj || !rot || !RotDet || !param) return fail(c, GRS_ERR_ARG, "null argument");
  if (n_param != 20) return fail(c, GRS_ERR_ARG, "param must hold 20 values (got %lld)", (long long)n_param);
  int64_t H = (int64_t)param[0], nrot = (int64_t)param[1];
  if ((double)H != (double)param[0] || (double)nrot != (double)param[1] || H < 1 || nrot < 1)
    return fail(c, GRS_ERR_ARG, "param(1:2) must be positive integers");
  if (n_hkl != 3 * H) return fail(c, GRS_ERR_ARG, "hkl holds %lld values, expected 3*nr_hkl = %lld",
                                  (long long)n_hkl, (long long)(3 * H));
  if (n_rot != nrot) return fail(c, GRS_ERR_ARG, "rot_angles holds %lld values, expected %lld", (long long)n_rot,
                                 (long long)nrot);
  double ysz = param[10], zsz = param[11];
  int64_t ny = (int64_t)ysz, nz = (int64_t)zsz;
  if ((double)ny != ysz || (double)nz != zsz || ny < 1 || nz < 1)
    return fail(c, GRS_ERR_ARG, "detector size must be positive integers");
  if (n_proj != ny * nz * nrot)
    return fail(c, GRS_ERR_ARG, "projection stack holds %lld values, expected detzsize*detysize*nr_rot = %lld",
                (long long)n_proj, (long long)(ny * nz * nrot));
  grs_geometry g;
  std::memset(&g, 0, sizeof(g));
  for (int k = 0; k < 9; k++) {
    g.S[k] = S[k];
    g.B[k] = B[k];
    g.RotDet[k] = RotDet[k];
  }
  g.Lsam2sou = param[2];
  g.Lsam2det = param[3];
  g.P0y = param[4];
  g.P0z = param[5];
  g.dety00 = param[6];
  g.detz00 = param[7];
  g.pixelysize = param[8];
  g.pixelzsize = param[9];
  g.detysize = ysz;
  g.detzsize = zsz;
  g.BeamStopY[0] = param[12];
  g.BeamStopY[1] = param[13];
  g.BeamStopZ[0] = param[14];
  g.BeamStopZ[1] = param[15];
  g.thetamax = param[16];
  g.lambda_min = param[17];
  g.lambda_max = param[18];
  g.RotAxisOffset = param[19];
  g.dety0 = ysz / 2.0 + 0.5;
  g.detz0 = zsz / 2.0 + 0.5;
  g.minEucDis = 0.5 * ((g.pixelysize + g.pixelzsize) / 2.0);
  std::vector<double> h(3 * H), r(nrot);
  for (int64_t k = 0; k < 3 * H; k++) h[k] = hkl[k];
  for (int64_t k = 0; k < nrot; k++) r[k] = rot[k];
  grs_status s = grs_set_geometry(c, &g, h.data(), H, r.data(), nrot);
  if (s != GRS_OK) return s;
  return grs_set_distance_stack(c, proj, nz, ny, nrot, 0);
}

static grs_status legacy_ret(grs_ctx *c, grs_status s) {
  if (s != GRS_OK && c) std::snprintf(g_legacy_err, sizeof(g_legacy_err), "%s", c->err);
  return s;
}

grs_status grs_cuda_forward_comp(const float *pos, const float *q, int64_t n_q, const float *S, const float *B,
                                 const float *hkl, int64_t n_hkl, const float *proj, int64_t n_proj,
                                 const float *rot, int64_t n_rot, const float *RotDet, const float *param,
                                 int64_t n_param, float *out) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  grs_ctx *c = legacy_ctx();
  if (!c) return GRS_ERR_DEVICE;
  if (!pos || !q || !out || n_q < 0 || n_q % 4) return legacy_ret(c, fail(c, GRS_ERR_ARG, "q must hold 4*N values"));
  grs_status s = legacy_setup(c, S, B, hkl, n_hkl, proj, n_proj, rot, n_rot, RotDet, param, n_param);
  if (s != GRS_OK) return legacy_ret(c, s);
  int64_t N = n_q / 4;
  std::vector<double> qd(n_q), pd = {pos[0], pos[1], pos[2]};
  for (int64_t k = 0; k < n_q; k++) qd[k] = q[k];
  std::vector<int32_t> ns(N), ni(N);
  std::vector<double> C(N);
  std::vector<float> med(N);
  s = grs_forward_comp(c, pd.data(), qd.data(), N, ns.data(), ni.data(), C.data(), med.data(), GRS_METHOD_FAST);
  if (s != GRS_OK) return legacy_ret(c, s);
  for (int64_t i = 0; i < N; i++) {
    out[4 * i] = (float)ns[i];
    out[4 * i + 1] = (float)ni[i];
    out[4 * i + 2] = (float)C[i];
    out[4 * i + 3] = med[i];
  }
  return GRS_OK;
}

grs_status grs_cuda_grow_indexed_region(const float *vox, int64_t n_vox, const float *U, const float *S,
                                        const float *B, const float *hkl, int64_t n_hkl, const float *proj,
                                        int64_t n_proj, const float *rot, int64_t n_rot, const float *RotDet,
                                        const float *param, int64_t n_param, const float *param1,
                                        int64_t n_param1, float *out) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  grs_ctx *c = legacy_ctx();
  if (!c) return GRS_ERR_DEVICE;
  if (!vox || !U || !param1 || !out || n_vox < 0 || n_vox % 6)
    return legacy_ret(c, fail(c, GRS_ERR_ARG, "voxel list must hold 6*M values"));
  if (n_param1 != 14) return legacy_ret(c, fail(c, GRS_ERR_ARG, "param1 must hold 14 values"));
  grs_status s = legacy_setup(c, S, B, hkl, n_hkl, proj, n_proj, rot, n_rot, RotDet, param, n_param);
  if (s != GRS_OK) return legacy_ret(c, s);
  int64_t M = n_vox / 6;
  std::vector<double> v(n_vox), Ud(9), p1(14), o(7 * M);
  for (int64_t k = 0; k < n_vox; k++) v[k] = vox[k];
  for (int k = 0; k < 9; k++) Ud[k] = U[k];
  for (int k = 0; k < 14; k++) p1[k] = param1[k];
  s = grs_grow_indexed_region(c, v.data(), M, Ud.data(), p1.data(), o.data(), GRS_METHOD_FAST);
  if (s != GRS_OK) return legacy_ret(c, s);
  for (int64_t k = 0; k < 7 * M; k++) out[k] = (float)o[k];
  return GRS_OK;
}

grs_status grs_cuda_forward_comp_compete(const float *vox, int64_t n_vox, const float *id_all, int64_t n_id,
                                         const float *euler, int64_t n_euler, const float *S, const float *B,
                                         const float *hkl, int64_t n_hkl, const float *proj, int64_t n_proj,
                                         const float *rot, int64_t n_rot, const float *RotDet,
                                         const float *param, int64_t n_param, const float *param1,
                                         int64_t n_param1, float *out) {
  std::lock_guard<std::mutex> lk(g_legacy_mu);
  grs_ctx *c = legacy_ctx();
  if (!c) return GRS_ERR_DEVICE;
  if (!vox || !param1 || !out || n_vox < 0 || n_vox % 5 || n_euler < 0 || n_euler % 3 || n_id < 0)
    return legacy_ret(c, fail(c, GRS_ERR_ARG, "voxel list must hold 5*M values and Euler 3*G values"));
  if (n_param1 != 11) return legacy_ret(c, fail(c, GRS_ERR_ARG, "param1 must hold 11 values"));
  grs_status s = legacy_setup(c, S, B, hkl, n_hkl, proj, n_proj, rot, n_rot, RotDet, param, n_param);
  if (s != GRS_OK) return legacy_ret(c, s);
  int64_t M = n_vox / 5, G = n_euler / 3;
  std::vector<double> v(n_vox), ids(n_id > 0 ? n_id : 1), eu(n_euler > 0 ? n_euler : 1), p1(11), o(8 * M);
  for (int64_t k = 0; k < n_vox; k++) v[k] = vox[k];
  for (int64_t k = 0; k < n_id; k++) ids[k] = id_all[k];
  for (int64_t k = 0; k < n_euler; k++) eu[k] = euler[k];
  for (int k = 0; k < 11; k++) p1[k] = param1[k];
  s = grs_forward_comp_compete(c, v.data(), M, ids.data(), n_id, eu.data(), G, p1.data(), o.data(),
                               GRS_METHOD_FAST);
  if (s != GRS_OK) return legacy_ret(c, s);
  for (int64_t k = 0; k < 8 * M; k++) out[k] = (float)o[k];
  return GRS_OK;
}

/* ---- orientation fit --------------------------------------------------------- */

/* M fits, fit m at pos[3m..] from x0[3m..] (the body of grs_fit_or_local) */
static grs_status fit_batch(grs_ctx *c, const double *pos, const double *x0, int64_t M, double half_width,
                            int32_t max_fun_evals, int32_t max_iter, double tol_fun, double tol_x, double *x_out,
                            double *fval, int32_t *n_fun_evals, int32_t *n_iter) {
  grs_status s = GRS_OK;
  if (M > INT32_MAX / 4) return fail(c, GRS_ERR_ARG, "too many fits");
  std::vector<grs_nm> nm((size_t)M);
  for (int64_t m = 0; m < M; m++) {
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      /* fit_OR_constrain_with_local_comp_cuda.m:124-125 */
      const double a = x0[3 * m + k] - half_width, b = x0[3 * m + k] + half_width;
      lo[k] = grs_nm_max(a, -1.0);
      hi[k] = (b != b) ? 1.0 : (b < 1.0 ? b : 1.0);
    }
    nm[m].max_fe = max_fun_evals;
    nm[m].max_it = max_iter;
    nm[m].tol_f = tol_fun;
    nm[m].tol_x = tol_x;
    nm[m].start(3, x0 + 3 * m, lo, hi);
  }
  ENSURE(c, c->d_pos, sizeof(double) * 3 * M);
  HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos, sizeof(double) * 3 * M, hipMemcpyHostToDevice, c->st));
  /* the positions stay fixed during the fits: their omega-solve bounds once */
  ENSURE(c, c->d_vb, sizeof(double) * GRS_VB * (size_t)M);
  HIPCHK(c, grs_launch_voxel_bounds(P<grs_dgeo>(c->d_geo), P<double>(c->d_pos), M, P<double>(c->d_cs), c->nrot,
                                    P<double>(c->d_vb), c->st));
  /* each step: [q (4E doubles) | eval voxel (E) | eval orientation (E)]
   * uploaded in one copy from pinned memory, [ns (E) | ni (E)] back in one.
   * A launch of E evaluations takes about the time of one while every
   * evaluation has a SIMD to itself; within that many, a fit's requested
   * reflection goes out with the three points its value can lead to
   * (grs_nm::follow_ups), and the one the method then requests is fed from
   * that launch's value instead of a launch of its own.  The method's
   * sequence of points, values and counts is unchanged. */
  const int64_t spec_max = grs_fast_wave_slots(c->device) / 16;
  std::vector<int64_t> first;
  std::vector<int32_t> npv, nspec;
  std::vector<double> fvals;
  std::vector<double> su((size_t)M * 3 * GRS_NM_NMAX);
  for (;;) {
    first.assign((size_t)M + 1, 0);
    npv.assign((size_t)M, 0);
    nspec.assign((size_t)M, 0);
    int64_t E = 0, nrefl = 0;
    for (int64_t m = 0; m < M; m++) {
      if (nm[m].done()) continue;
      E += nm[m].npend;
      nrefl += nm[m].st == grs_nm::REFLECT ? 1 : 0;
    }
    if (E == 0) break;
    const bool spec = E + 3 * nrefl <= spec_max;
    E = 0;
    for (int64_t m = 0; m < M; m++) {
      first[m] = E;
      if (nm[m].done()) continue;
      npv[m] = nm[m].npend;
      if (spec) nspec[m] = nm[m].follow_ups(reinterpret_cast<double (*)[GRS_NM_NMAX]>(&su[(size_t)m * 3 * GRS_NM_NMAX]));
      E += npv[m] + nspec[m];
    }
    first[M] = E;
    const size_t in_bytes = (sizeof(double) * 4 + sizeof(int32_t) * 2) * (size_t)E;
    const size_t cnt_bytes = sizeof(int32_t) * 2 * (size_t)E;
    if (c->h_pin_cap < in_bytes + cnt_bytes) {
      if (c->h_pin) HIPCHK(c, hipHostFree(c->h_pin));
      c->h_pin = nullptr;
      c->h_pin_cap = 0;
      const size_t want = 2 * (in_bytes + cnt_bytes);
      const hipError_t e = hipHostMalloc(&c->h_pin, want, hipHostMallocDefault);
      if (e != hipSuccess) {
        c->h_pin = nullptr;
        (void)hipGetLastError();
        return fail(c, GRS_ERR_NOMEM, "hipHostMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
      }
      c->h_pin_cap = want;
    }
    double *hq = static_cast<double *>(c->h_pin);
    int32_t *hvox = reinterpret_cast<int32_t *>(hq + 4 * E), *heu = hvox + E;
    int32_t *hcnt = reinterpret_cast<int32_t *>(static_cast<char *>(c->h_pin) + in_bytes);
    for (int64_t m = 0; m < M; m++) {
      if (nm[m].done()) continue;
      for (int p = 0; p < npv[m] + nspec[m]; p++) {
        const int64_t e = first[m] + p;
        double x[3];
        nm[m].xform(p < npv[m] ? nm[m].pend[p] : &su[((size_t)m * 3 + (size_t)(p - npv[m])) * GRS_NM_NMAX], x);
        hq[4 * e] = 0.0; /* q0 is recomputed from the vector part (fit_OR_fun_local.m:9) */
        hq[4 * e + 1] = x[0];
        hq[4 * e + 2] = x[1];
        hq[4 * e + 3] = x[2];
        hvox[e] = (int32_t)m;
        heu[e] = (int32_t)e;
      }
    }
    ENSURE(c, c->f_in, in_bytes);
    ENSURE(c, c->f_cnt, cnt_bytes);
    ENSURE(c, c->d_sub, sizeof(double) * 9 * E);
    ENSURE(c, c->d_valid, sizeof(int32_t) * E);
    ENSURE(c, c->d_C, sizeof(double) * E);
    ENSURE(c, c->d_med, sizeof(float) * E);
    HIPCHK(c, hipMemcpyAsync(c->f_in.p, hq, in_bytes, hipMemcpyHostToDevice, c->st));
    double *dq = P<double>(c->f_in);
    int32_t *dvox = reinterpret_cast<int32_t *>(dq + 4 * E), *deu = dvox + E;
    int32_t *dns = P<int32_t>(c->f_cnt), *dni = dns + E;
    HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), dq, E, P<double>(c->d_sub), P<int32_t>(c->d_valid),
                                    c->st));
    s = run_eval(c, GRS_MAP_PAIRS, E, M, dvox, deu, GRS_METHOD_FAST, dns, dni, P<double>(c->d_C),
                 P<float>(c->d_med), true);
    if (s != GRS_OK) return s;
    HIPCHK(c, hipMemcpyAsync(hcnt, dns, cnt_bytes, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    fvals.resize((size_t)E);
    for (int64_t e = 0; e < E; e++) /* fit_OR_fun_local.m:98; 0/0 (no event, or a skipped row) is NaN */
      fvals[e] = 1.0 - (double)hcnt[E + e] / (double)hcnt[e];
    for (int64_t m = 0; m < M; m++) {
      if (first[m + 1] == first[m]) continue;
      nm[m].feed(&fvals[first[m]]);
      /* an expansion or contraction found among the follow-ups (bit for bit) */
      while (nspec[m] > 0 && !nm[m].done() && nm[m].npend == 1 &&
             (nm[m].st == grs_nm::EXPAND || nm[m].st == grs_nm::CONTRACT_OUT || nm[m].st == grs_nm::CONTRACT_IN)) {
        int hit = -1;
        for (int s2 = 0; s2 < nspec[m] && hit < 0; s2++)
          if (std::memcmp(nm[m].pend[0], &su[((size_t)m * 3 + (size_t)s2) * GRS_NM_NMAX],
                          sizeof(double) * (size_t)nm[m].n) == 0)
            hit = s2;
        if (hit < 0) break;
        nm[m].feed(&fvals[first[m] + npv[m] + hit]);
      }
    }
  }
  for (int64_t m = 0; m < M; m++) {
    nm[m].result(x_out + 3 * m, fval + m);
    n_fun_evals[m] = nm[m].nf;
    n_iter[m] = nm[m].it;
  }
  return GRS_OK;
}

grs_status grs_fit_or_local(grs_ctx *c, const double *pos, const double *x0, int64_t M, double half_width,
                            int32_t max_fun_evals, int32_t max_iter, double tol_fun, double tol_x, double *x_out,
                            double *fval, int32_t *n_fun_evals, int32_t *n_iter) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || (M > 0 && (!pos || !x0 || !x_out || !fval || !n_fun_evals || !n_iter)) || !(half_width > 0.0) ||
      max_fun_evals < 1 || max_iter < 1 || !(tol_fun >= 0.0) || !(tol_x >= 0.0))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (M == 0) return GRS_OK;
  return fit_batch(c, pos, x0, M, half_width, max_fun_evals, max_iter, tol_fun, tol_x, x_out, fval, n_fun_evals,
                   n_iter);
}

/* ---- seed indexing ---------------------------------------------------------- */

grs_status grs_fit_orientations(const double *x, int64_t M, double *out10, double *U, int32_t *ok) {
  if (M < 0 || (M > 0 && (!x || !out10 || !U || !ok))) return GRS_ERR_ARG;
  for (int64_t m = 0; m < M; m++) ok[m] = grs_fit_orientation(x + 3 * m, out10 + 10 * m, U + 9 * m);
  return GRS_OK;
}

grs_status grs_set_orientation_set(grs_ctx *c, const double *q, int64_t N) {
  if (!c) return GRS_ERR_ARG;
  if (N < 1 || N > 0x7FFFFFFFll || !q) return fail(c, GRS_ERR_ARG, "need 1 .. 2^31-1 orientations");
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->o_q, sizeof(double) * 4 * (size_t)N);
  ENSURE(c, c->o_sub, sizeof(double) * 9 * (size_t)N);
  ENSURE(c, c->o_valid, sizeof(int32_t) * (size_t)N);
  HIPCHK(c, hipMemcpyAsync(c->o_q.p, q, sizeof(double) * 4 * (size_t)N, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->o_n = N;
  c->o_gen = 0; /* S*U*B computed on first use under the current geometry */
  return GRS_OK;
}

/* seed positions per device pass: the pipeline keeps per seed the grid
 * scores of the orientation set and Nr_seed*cap local rows */
static int64_t index_chunk(int64_t N, int64_t K, int64_t cap) {
  const int64_t Rcap = K * cap;
  const int64_t per = N * 16 + K * (48 + 24 * GRS_LS_MAX_GRID + 32 * cap + 64) + Rcap * 132;
  const int64_t budget = (int64_t)6 << 30;
  int64_t m = budget / (per > 0 ? per : 1);
  return m < 1 ? 1 : m;
}

grs_status grs_index_seeds(grs_ctx *c, const double *pos, int64_t n_pos, int32_t nr_seed, int32_t nr, double gridsize,
                           int32_t n_fit, double half_width, int32_t max_fun_evals, int32_t max_iter, double tol_fun,
                           double tol_x, double *indexing, int32_t *candidate, double *fit_info, int32_t *seed_idx,
                           double *seed_C) {
  if (!c) return GRS_ERR_ARG;
  if (n_pos < 0 || (n_pos > 0 && (!pos || !indexing || !candidate)) || nr_seed < 1 || nr < 0 || nr > 11 ||
      !(gridsize > 0.0) || !std::isfinite(gridsize) || n_fit < 1 || !(half_width > 0.0) || max_fun_evals < 1 ||
      max_iter < 1 || !(tol_fun >= 0.0) || !(tol_x >= 0.0))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (c->o_n < 1) return fail(c, GRS_ERR_STATE, "orientation set not set");
  if ((int64_t)c->nrot * c->H > GRS_KEY_MAX_EVENTS)
    return fail(c, GRS_ERR_ARG, "nrot*nrhkl = %lld exceeds %d events, the exact range of the packed sort key",
                (long long)c->nrot * c->H, GRS_KEY_MAX_EVENTS);
  if (c->H > 240) return fail(c, GRS_ERR_ARG, "at most 240 reflections");
  const int64_t N = c->o_n;
  const int64_t K = nr_seed < N ? nr_seed : N;
  if (K > GRS_TOPK_MAX) return fail(c, GRS_ERR_ARG, "Nr_seed above %d", GRS_TOPK_MAX);
  const int32_t nfit = n_fit < K ? n_fit : (int32_t)K;
  const int32_t cap_axis = nr + 1;
  const int64_t cap = (int64_t)cap_axis * cap_axis * cap_axis + 1;
  const int64_t Rcap = K * cap;
  if (n_pos == 0) return GRS_OK;
  if (c->o_gen != c->geo_gen) {
    HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), P<double>(c->o_q), N, P<double>(c->o_sub),
                                    P<int32_t>(c->o_valid), c->st));
    c->o_gen = c->geo_gen;
  }
  s = ensure_bits(c);
  if (s != GRS_OK) return s;
  const int64_t stride = (int64_t)c->nrot * c->H;
  const int64_t chunk = index_chunk(N, K, cap);
  for (int64_t p0 = 0; p0 < n_pos; p0 += chunk) {
    const int64_t m = n_pos - p0 < chunk ? n_pos - p0 : chunk;
    const int64_t E1 = m * N, E2 = m * Rcap;
    ENSURE(c, c->x_pos, sizeof(double) * 3 * m);
    ENSURE(c, c->x_vb, sizeof(double) * GRS_VB * m);
    ENSURE(c, c->x_key1, sizeof(unsigned long long) * E1);
    ENSURE(c, c->x_ns1, sizeof(int32_t) * E1);
    ENSURE(c, c->x_ni1, sizeof(int32_t) * E1);
    ENSURE(c, c->x_top, sizeof(unsigned long long) * m * K);
    ENSURE(c, c->x_cen, sizeof(double) * 4 * m * K);
    ENSURE(c, c->x_grid, sizeof(double) * 3 * GRS_LS_MAX_GRID * m * K);
    ENSURE(c, c->x_gn, sizeof(int32_t) * 3 * m * K);
    ENSURE(c, c->x_ls, sizeof(double) * 4 * m * K * cap);
    ENSURE(c, c->x_lsn, sizeof(int64_t) * m * K);
    ENSURE(c, c->x_owner, sizeof(int32_t) * E2);
    ENSURE(c, c->x_Rmax, sizeof(unsigned long long) * m);
    ENSURE(c, c->x_q2, sizeof(double) * 4 * E2);
    ENSURE(c, c->x_sub2, sizeof(double) * 9 * E2);
    ENSURE(c, c->x_valid2, sizeof(int32_t) * E2);
    ENSURE(c, c->x_key2, sizeof(unsigned long long) * E2);
    ENSURE(c, c->x_ns2, sizeof(int32_t) * E2);
    ENSURE(c, c->x_ni2, sizeof(int32_t) * E2);
    ENSURE(c, c->x_gcount, sizeof(uint32_t) * m * K);
    ENSURE(c, c->x_gkey, sizeof(unsigned long long) * m * K);
    ENSURE(c, c->x_grow, sizeof(int64_t) * m * K);
    ENSURE(c, c->x_fitkey, sizeof(unsigned long long) * m * nfit);
    ENSURE(c, c->x_x0, sizeof(double) * 3 * m * nfit);
    ENSURE(c, c->x_fgroup, sizeof(int32_t) * m * nfit);
    ENSURE(c, c->x_fns, sizeof(int32_t) * m * nfit);
    ENSURE(c, c->x_fni, sizeof(int32_t) * m * nfit);
    ENSURE(c, c->x_sidx, sizeof(int32_t) * m * K);
    ENSURE(c, c->x_sC, sizeof(double) * m * K);
    ENSURE(c, c->x_bad, sizeof(uint32_t));
    const int64_t slots = grs_fast_wave_slots(c->device);
    const int64_t w1 = slots > E1 ? ((E1 + 3) / 4) * 4 : slots;
    const int64_t w2 = slots > E2 ? ((E2 + 3) / 4) * 4 : slots;
    ENSURE(c, c->d_scratch, sizeof(uint32_t) * (size_t)(w1 > w2 ? w1 : w2) * stride);
    HIPCHK(c, hipMemcpyAsync(c->x_pos.p, pos + 3 * p0, sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->st));
    grs_eval_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = P<grs_dgeo>(c->d_geo);
    a.hkl = P<double>(c->d_hkl);
    a.cs = P<double>(c->d_cs);
    a.vb = P<double>(c->x_vb);
    a.pos = P<double>(c->x_pos);
    a.keys = P<uint32_t>(c->d_keys);
    a.hit_bits = P<uint32_t>(c->d_bits);
    a.hit_tbits = P<uint32_t>(c->d_tbits);
    a.scratch_stride = stride;
    a.scratch = P<uint32_t>(c->d_scratch);
    a.map = GRS_MAP_GRID;
    auto enqueue = [&]() -> grs_status {
      HIPCHK(c, grs_launch_fill_u32(P<uint32_t>(c->x_owner), E2, 0u, c->st));
      HIPCHK(c, grs_launch_fill_u32(P<uint32_t>(c->x_gcount), m * K, 0u, c->st));
      HIPCHK(c, grs_launch_fill_u32(P<uint32_t>(c->x_bad), 1, 0u, c->st));
      HIPCHK(c, grs_launch_fill_u64(P<unsigned long long>(c->x_Rmax), m, 0ull, c->st));
      HIPCHK(c, grs_launch_voxel_bounds(P<grs_dgeo>(c->d_geo), P<double>(c->x_pos), m, P<double>(c->d_cs), c->nrot,
                                        P<double>(c->x_vb), c->st));
      /* ForwardCalc_OR_space_comp_cuda over the orientation set */
      grs_eval_args a1 = a;
      a1.SUB = P<double>(c->o_sub);
      a1.valid = P<int32_t>(c->o_valid);
      a1.n_eval = E1;
      a1.n_u = N;
      a1.u_stride = 0;
      a1.out_key = P<unsigned long long>(c->x_key1);
      a1.out_ns = P<int32_t>(c->x_ns1);
      a1.out_ni = P<int32_t>(c->x_ni1);
      grs_status ls = launch_tuned(c, a1, 2, w1);
      if (ls != GRS_OK) return ls;
      /* sortrows(...,3,'descend'), rows 1:Nr_seed */
      HIPCHK(c, grs_launch_topk(P<unsigned long long>(c->x_key1), m, N, N, (int32_t)K,
                                P<unsigned long long>(c->x_top), c->st));
      HIPCHK(c, grs_launch_seed_out(P<unsigned long long>(c->x_top), P<int32_t>(c->x_ns1), P<int32_t>(c->x_ni1),
                                    P<int32_t>(c->o_valid), m, K, N, P<int32_t>(c->x_sidx), P<double>(c->x_sC),
                                    c->st));
      /* local sampling around every seed, laid out as q_all */
      HIPCHK(c, grs_launch_gather_centres(P<double>(c->o_q), P<unsigned long long>(c->x_top), m * K,
                                          P<double>(c->x_cen), c->st));
      HIPCHK(c, grs_launch_ls_grids(P<double>(c->x_cen), m * K, nr, gridsize, cap_axis, P<double>(c->x_grid),
                                    P<int32_t>(c->x_gn), P<uint32_t>(c->x_bad), c->st));
      HIPCHK(c, grs_launch_local_sampling(P<double>(c->x_cen), P<double>(c->x_grid), P<int32_t>(c->x_gn), m * K,
                                          P<double>(c->x_ls), cap, P<int64_t>(c->x_lsn), P<uint32_t>(c->x_bad),
                                          c->st));
      HIPCHK(c, grs_launch_qall(P<double>(c->x_ls), P<int64_t>(c->x_lsn), m, K, cap, P<int32_t>(c->x_owner),
                                P<unsigned long long>(c->x_Rmax), P<double>(c->x_q2), P<uint32_t>(c->x_gcount),
                                c->st));
      HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), P<double>(c->x_q2), E2, P<double>(c->x_sub2),
                                      P<int32_t>(c->x_valid2), c->st));
      /* cuda_forward_comp over q_all */
      grs_eval_args a2 = a;
      a2.SUB = P<double>(c->x_sub2);
      a2.valid = P<int32_t>(c->x_valid2);
      a2.n_eval = E2;
      a2.n_u = Rcap;
      a2.u_stride = Rcap;
      a2.out_key = P<unsigned long long>(c->x_key2);
      a2.out_ns = P<int32_t>(c->x_ns2);
      a2.out_ni = P<int32_t>(c->x_ni2);
      if ((ls = launch_tuned(c, a2, 2, w2)) != GRS_OK) return ls;
      /* per-seed maxima, sortrows(Output_max,3,'descend'), rows to fit */
      HIPCHK(c, grs_launch_group_max(P<unsigned long long>(c->x_key2), P<int64_t>(c->x_lsn),
                                     P<uint32_t>(c->x_gcount), m, K, Rcap, P<unsigned long long>(c->x_gkey),
                                     P<int64_t>(c->x_grow), c->st));
      HIPCHK(c, grs_launch_topk(P<unsigned long long>(c->x_gkey), m, K, K, nfit, P<unsigned long long>(c->x_fitkey),
                                c->st));
      HIPCHK(c, grs_launch_fit_starts(P<unsigned long long>(c->x_fitkey), P<int64_t>(c->x_grow), P<double>(c->x_q2),
                                      P<int32_t>(c->x_owner), P<int32_t>(c->x_ns2), P<int32_t>(c->x_ni2), m, K, nfit,
                                      Rcap, P<double>(c->x_x0), P<int32_t>(c->x_fgroup), P<int32_t>(c->x_fns),
                                      P<int32_t>(c->x_fni), c->st));
      return GRS_OK;
    };
    const bool tuning = tune_pending(c, 2, E1, w1) || tune_pending(c, 2, E2, w2);
    if (c->graphs && !tuning) {
      double gs = gridsize;
      uint64_t gbits;
      std::memcpy(&gbits, &gs, sizeof(gbits));
      std::vector<uint64_t> key = {(uint64_t)m, (uint64_t)N, (uint64_t)K, (uint64_t)nr, gbits, (uint64_t)nfit,
                                   (uint64_t)c->H, (uint64_t)c->nrot, (uint64_t)w1, (uint64_t)w2, (uint64_t)stride,
                                   tune_code(c, 2), (uint64_t)c->autotune};
      for (const DevBuf *bp :
           {&c->d_geo, &c->d_hkl, &c->d_cs, &c->d_keys, &c->d_bits, &c->d_tbits, &c->d_scratch, &c->o_q, &c->o_sub, &c->o_valid,
            &c->x_pos, &c->x_vb, &c->x_key1, &c->x_ns1, &c->x_ni1, &c->x_top, &c->x_cen, &c->x_grid, &c->x_gn,
            &c->x_ls, &c->x_lsn, &c->x_owner, &c->x_Rmax, &c->x_q2, &c->x_sub2, &c->x_valid2, &c->x_key2, &c->x_ns2,
            &c->x_ni2, &c->x_gcount, &c->x_gkey, &c->x_grow, &c->x_fitkey, &c->x_x0, &c->x_fgroup, &c->x_fns,
            &c->x_fni, &c->x_sidx, &c->x_sC, &c->x_bad})
        key.push_back((uint64_t)(uintptr_t)bp->p);
      auto it = c->gi_cache.find(key);
      hipGraphExec_t exec = nullptr;
      if (it == c->gi_cache.end()) {
        if (c->gi_cache.size() >= 4) {
          for (auto &kv : c->gi_cache) HIPCHK(c, hipGraphExecDestroy(kv.second));
          c->gi_cache.clear();
        }
        hipGraph_t graph = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        grs_status es = enqueue();
        hipError_t e2 = hipStreamEndCapture(c->st, &graph);
        if (es != GRS_OK || e2 != hipSuccess) {
          if (graph) hipGraphDestroy(graph);
          if (es != GRS_OK) return es;
          return fail(c, GRS_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e2));
        }
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(c, GRS_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
        c->gi_cache[key] = exec;
        c->g_captures += 1;
      } else {
        exec = it->second;
        c->g_replays += 1;
      }
      HIPCHK(c, hipGraphLaunch(exec, c->st));
    } else {
      grs_status es = enqueue();
      if (es != GRS_OK) return es;
    }
    const int64_t F = m * nfit;
    std::vector<double> x0((size_t)F * 3), fpos((size_t)F * 3);
    std::vector<int32_t> fgroup((size_t)F), fns((size_t)F), fni((size_t)F);
    uint32_t bad = 0;
    HIPCHK(c, hipMemcpyAsync(x0.data(), c->x_x0.p, sizeof(double) * 3 * F, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(fgroup.data(), c->x_fgroup.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(&bad, c->x_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
    std::vector<int32_t> sidx;
    std::vector<double> sC;
    if (seed_idx || seed_C) {
      sidx.resize((size_t)m * K);
      sC.resize((size_t)m * K);
      HIPCHK(c, hipMemcpyAsync(sidx.data(), c->x_sidx.p, sizeof(int32_t) * m * K, hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipMemcpyAsync(sC.data(), c->x_sC.p, sizeof(double) * m * K, hipMemcpyDeviceToHost, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (bad & 1u) return fail(c, GRS_ERR_ARG, "a local-sampling grid point has q1^2+q2^2+q3^2 > 1");
    if (bad & 2u) return fail(c, GRS_ERR_ARG, "a local-sampling grid has more than nr+1 values");
    for (int64_t p = 0; p < m; p++) {
      for (int64_t r = 0; r < nr_seed; r++) {
        if (seed_idx) seed_idx[(p0 + p) * nr_seed + r] = r < K ? sidx[(size_t)(p * K + r)] : -1;
        if (seed_C) seed_C[(p0 + p) * nr_seed + r] = r < K ? sC[(size_t)(p * K + r)] : std::nan("");
      }
      for (int32_t f = 0; f < nfit; f++)
        for (int k = 0; k < 3; k++) fpos[(size_t)(p * nfit + f) * 3 + k] = pos[3 * (p0 + p) + k];
    }
    /* fminsearchbnd from every start (fit_OR_constrain_with_local_comp_cuda.m:115-126) */
    std::vector<double> fx((size_t)F * 3), fval((size_t)F);
    std::vector<int32_t> fnf((size_t)F), fit_((size_t)F);
    s = fit_batch(c, fpos.data(), x0.data(), F, half_width, max_fun_evals, max_iter, tol_fun, tol_x, fx.data(),
                  fval.data(), fnf.data(), fit_.data());
    if (s != GRS_OK) return s;
    /* Euler angles, Rodrigues vector, quaternion, U (:133-148) and the
     * verification counts with that U (:151-153) */
    std::vector<double> o10((size_t)F * 10), U((size_t)F * 9);
    std::vector<int32_t> ok((size_t)F), evox((size_t)F), eu((size_t)F), vns((size_t)F), vni((size_t)F);
    std::vector<double> vC((size_t)F);
    std::vector<float> vmed((size_t)F);
    for (int64_t f = 0; f < F; f++) {
      ok[f] = grs_fit_orientation(&fx[(size_t)f * 3], &o10[(size_t)f * 10], &U[(size_t)f * 9]);
      evox[f] = (int32_t)f;
      eu[f] = (int32_t)f;
    }
    ENSURE(c, c->d_pos, sizeof(double) * 3 * F);
    ENSURE(c, c->d_U, sizeof(double) * 9 * F);
    ENSURE(c, c->d_sub, sizeof(double) * 9 * F);
    ENSURE(c, c->d_valid, sizeof(int32_t) * F);
    ENSURE(c, c->d_evox, sizeof(int32_t) * F);
    ENSURE(c, c->d_eu_idx, sizeof(int32_t) * F);
    ENSURE(c, c->d_ns, sizeof(int32_t) * F);
    ENSURE(c, c->d_ni, sizeof(int32_t) * F);
    ENSURE(c, c->d_C, sizeof(double) * F);
    ENSURE(c, c->d_med, sizeof(float) * F);
    HIPCHK(c, hipMemcpyAsync(c->d_pos.p, fpos.data(), sizeof(double) * 3 * F, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_U.p, U.data(), sizeof(double) * 9 * F, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_evox.p, evox.data(), sizeof(int32_t) * F, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_eu_idx.p, eu.data(), sizeof(int32_t) * F, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), F, P<double>(c->d_sub),
                                    P<int32_t>(c->d_valid), c->st));
    s = run_eval(c, GRS_MAP_PAIRS, F, F, P<int32_t>(c->d_evox), P<int32_t>(c->d_eu_idx), GRS_METHOD_FAST,
                 P<int32_t>(c->d_ns), P<int32_t>(c->d_ni), P<double>(c->d_C), P<float>(c->d_med));
    if (s != GRS_OK) return s;
    HIPCHK(c, hipMemcpyAsync(vns.data(), c->d_ns.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(vni.data(), c->d_ni.p, sizeof(int32_t) * F, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipMemcpyAsync(vmed.data(), c->d_med.p, sizeof(float) * F, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int64_t p = 0; p < m; p++) {
      /* [~,ind] = max(indexing_result(:,7)) (:161): NaN ignored, first maximum */
      int32_t best = 0;
      double bestC = std::nan("");
      for (int32_t f = 0; f < nfit; f++) {
        const int64_t i = p * nfit + f;
        const double Cf = ok[i] ? (double)vni[i] / (double)vns[i] : std::nan("");
        vC[i] = Cf;
        if (Cf == Cf && (bestC != bestC || Cf > bestC)) {
          best = f;
          bestC = Cf;
        }
      }
      const int64_t i = p * nfit + best;
      double *row = indexing + (p0 + p) * 18;
      row[0] = (double)(p0 + p + 1);
      for (int k = 0; k < 3; k++) row[1 + k] = pos[3 * (p0 + p) + k];
      if (ok[i]) {
        row[4] = (double)vns[i];
        row[5] = (double)vni[i];
        row[6] = vC[i];
        row[7] = (double)vmed[i];
      } else {
        row[4] = row[5] = row[6] = row[7] = std::nan("");
      }
      for (int k = 0; k < 10; k++) row[8 + k] = o10[(size_t)i * 10 + k];
      candidate[p0 + p] = fgroup[i];
      if (fit_info) {
        double *fi = fit_info + (p0 + p) * 9;
        for (int k = 0; k < 3; k++) fi[k] = x0[(size_t)i * 3 + k];
        for (int k = 0; k < 3; k++) fi[3 + k] = fx[(size_t)i * 3 + k];
        fi[6] = fval[i];
        fi[7] = (double)fnf[i];
        fi[8] = (double)fit_[i];
      }
    }
  }
  return GRS_OK;
}

/* ---- spot pairing ----------------------------------------------------------- */

grs_status grs_set_frame_spots(grs_ctx *c, const int64_t *frame_ptr, int64_t nframes, const double *spots,
                               int64_t n) {
  if (!c) return GRS_ERR_ARG;
  if (nframes < 1 || !frame_ptr || n < 0 || (n > 0 && !spots)) return fail(c, GRS_ERR_ARG, "invalid argument");
  if (frame_ptr[0] != 0 || frame_ptr[nframes] != n) return fail(c, GRS_ERR_ARG, "frame_ptr must run from 0 to n");
  for (int64_t i = 0; i < nframes; i++)
    if (frame_ptr[i + 1] < frame_ptr[i]) return fail(c, GRS_ERR_ARG, "frame_ptr must be non-decreasing");
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->p_fptr, sizeof(int64_t) * (size_t)(nframes + 1));
  ENSURE(c, c->p_fsp, sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  HIPCHK(c, hipMemcpyAsync(c->p_fptr.p, frame_ptr, sizeof(int64_t) * (size_t)(nframes + 1), hipMemcpyHostToDevice,
                           c->st));
  if (n > 0)
    HIPCHK(c, hipMemcpyAsync(c->p_fsp.p, spots, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->p_nframes = nframes;
  c->p_n = n;
  c->p_maxframe = 0;
  for (int64_t i = 0; i < nframes; i++)
    c->p_maxframe = frame_ptr[i + 1] - frame_ptr[i] > c->p_maxframe ? frame_ptr[i + 1] - frame_ptr[i] : c->p_maxframe;
  return GRS_OK;
}

grs_status grs_pair_spots(grs_ctx *c, const double *pos, const double *U, int64_t M, int32_t *ns, int32_t *ni,
                          float *med, int64_t *row_ptr, double *rows, int64_t cap) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || !row_ptr || (M > 0 && (!pos || !U || !ns || !ni || !med)) || cap < 0)
    return fail(c, GRS_ERR_ARG, "invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (c->p_nframes != c->nrot)
    return fail(c, GRS_ERR_STATE, "experimental spots of %d projections needed (grs_set_frame_spots)", c->nrot);
  row_ptr[0] = 0;
  if (M == 0) return GRS_OK;
  /* Nr_simu, Nr_intersect, dis_median: the scoring kernel */
  std::vector<int32_t> ident((size_t)M);
  for (int64_t m = 0; m < M; m++) ident[m] = (int32_t)m;
  ENSURE(c, c->d_pos, sizeof(double) * 3 * M);
  ENSURE(c, c->d_U, sizeof(double) * 9 * M);
  ENSURE(c, c->d_sub, sizeof(double) * 9 * M);
  ENSURE(c, c->d_valid, sizeof(int32_t) * M);
  ENSURE(c, c->d_evox, sizeof(int32_t) * M);
  ENSURE(c, c->d_eu_idx, sizeof(int32_t) * M);
  ENSURE(c, c->d_ns, sizeof(int32_t) * M);
  ENSURE(c, c->d_ni, sizeof(int32_t) * M);
  ENSURE(c, c->d_C, sizeof(double) * M);
  ENSURE(c, c->d_med, sizeof(float) * M);
  HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos, sizeof(double) * 3 * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_U.p, U, sizeof(double) * 9 * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_evox.p, ident.data(), sizeof(int32_t) * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_eu_idx.p, ident.data(), sizeof(int32_t) * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), M, P<double>(c->d_sub),
                                  P<int32_t>(c->d_valid), c->st));
  s = run_eval(c, GRS_MAP_PAIRS, M, M, P<int32_t>(c->d_evox), P<int32_t>(c->d_eu_idx), GRS_METHOD_FAST,
               P<int32_t>(c->d_ns), P<int32_t>(c->d_ni), P<double>(c->d_C), P<float>(c->d_med));
  if (s != GRS_OK) return s;
  HIPCHK(c, hipMemcpyAsync(ns, c->d_ns.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(ni, c->d_ni.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(med, c->d_med.p, sizeof(float) * M, hipMemcpyDeviceToHost, c->st));
  /* the pairs, in chunks of orientations */
  const int64_t E = (int64_t)c->nrot * c->H;
  int64_t chunk = ((int64_t)1 << 26) / E;
  if (chunk < 1) chunk = 1;
  std::vector<int64_t> cnt;
  for (int64_t m0 = 0; m0 < M; m0 += chunk) {
    const int64_t m = M - m0 < chunk ? M - m0 : chunk;
    const int64_t me = m * E;
    ENSURE(c, c->p_flag, sizeof(int32_t) * me);
    ENSURE(c, c->p_sid, sizeof(int64_t) * me);
    ENSURE(c, c->p_lam, sizeof(double) * me);
    ENSURE(c, c->p_dyz, sizeof(int32_t) * 2 * me);
    ENSURE(c, c->p_dis, sizeof(float) * me);
    ENSURE(c, c->p_off, sizeof(int32_t) * me);
    ENSURE(c, c->p_cnt, sizeof(int64_t) * (m + 1));
    ENSURE(c, c->p_rowptr, sizeof(int64_t) * (m + 1));
    HIPCHK(c, grs_launch_pair_events(P<grs_dgeo>(c->d_geo), P<double>(c->d_hkl), P<double>(c->d_cs),
                                     P<uint32_t>(c->d_keys), P<double>(c->d_pos) + 3 * m0,
                                     P<double>(c->d_sub) + 9 * m0, m, E, P<int64_t>(c->p_fptr), P<double>(c->p_fsp),
                                     P<int32_t>(c->p_flag), P<int64_t>(c->p_sid), P<double>(c->p_lam),
                                     P<int32_t>(c->p_dyz), P<float>(c->p_dis), c->st));
    HIPCHK(c, grs_launch_pair_scan(P<int32_t>(c->p_flag), m, E, P<int32_t>(c->p_off), P<int64_t>(c->p_cnt), c->st));
    cnt.resize((size_t)m);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), c->p_cnt.p, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::vector<int64_t> rp((size_t)m + 1, 0);
    for (int64_t k = 0; k < m; k++) {
      rp[k + 1] = rp[k] + cnt[k];
      row_ptr[m0 + k + 1] = row_ptr[m0 + k] + cnt[k];
    }
    if (!rows || rp[m] == 0) continue;
    if (row_ptr[m0 + m] > cap)
      return fail(c, GRS_ERR_ARG, "%lld pair rows do not fit the capacity %lld", (long long)row_ptr[m0 + m],
                  (long long)cap);
    ENSURE(c, c->p_rows, sizeof(double) * 23 * rp[m]);
    HIPCHK(c, hipMemcpyAsync(c->p_rowptr.p, rp.data(), sizeof(int64_t) * (m + 1), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_pair_rows(P<grs_dgeo>(c->d_geo), P<double>(c->d_hkl), P<double>(c->d_rot),
                                   P<double>(c->d_pos) + 3 * m0, P<double>(c->d_U) + 9 * m0, m, E,
                                   P<double>(c->p_fsp), P<int32_t>(c->p_flag), P<int64_t>(c->p_sid),
                                   P<double>(c->p_lam), P<int32_t>(c->p_dyz), P<float>(c->p_dis),
                                   P<int32_t>(c->p_off), P<int64_t>(c->p_rowptr), P<double>(c->p_rows), c->st));
    HIPCHK(c, hipMemcpyAsync(rows + 23 * row_ptr[m0], c->p_rows.p, sizeof(double) * 23 * rp[m],
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

/* ---- region growth with resident volumes -------------------------------------- */

grs_status grs_set_volume(grs_ctx *c, int32_t which, const double *v, int64_t nx, int64_t ny, int64_t nz) {
  if (!c) return GRS_ERR_ARG;
  if (which < 0 || which > 3 || !v || nx < 1 || ny < 1 || nz < 1) return fail(c, GRS_ERR_ARG, "invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  if (nx != c->v_dim[0] || ny != c->v_dim[1] || nz != c->v_dim[2]) {
    for (int k = 0; k < 4; k++) c->v_set[k] = false;
    c->v_dim[0] = nx;
    c->v_dim[1] = ny;
    c->v_dim[2] = nz;
  }
  const size_t n = (size_t)(nx * ny * nz);
  ENSURE(c, c->v_vol[which], sizeof(double) * n);
  HIPCHK(c, hipMemcpyAsync(c->v_vol[which].p, v, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->v_set[which] = true;
  return GRS_OK;
}

grs_status grs_get_volume(grs_ctx *c, int32_t which, double *v, int64_t n) {
  if (!c) return GRS_ERR_ARG;
  if (which < 0 || which > 3 || !v) return fail(c, GRS_ERR_ARG, "invalid argument");
  if (!c->v_set[which]) return fail(c, GRS_ERR_STATE, "volume %d not set", which);
  if (n != c->v_dim[0] * c->v_dim[1] * c->v_dim[2]) return fail(c, GRS_ERR_ARG, "size mismatch");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(v, c->v_vol[which].p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_get_volume_at(grs_ctx *c, int32_t which, const int64_t *lin, int64_t n, double *out) {
  if (!c) return GRS_ERR_ARG;
  if (which < 0 || which > 3 || n < 0 || (n > 0 && (!lin || !out))) return fail(c, GRS_ERR_ARG, "invalid argument");
  if (!c->v_set[which]) return fail(c, GRS_ERR_STATE, "volume %d not set", which);
  const int64_t total = c->v_dim[0] * c->v_dim[1] * c->v_dim[2];
  for (int64_t i = 0; i < n; i++)
    if (lin[i] < 0 || lin[i] >= total)
      return fail(c, GRS_ERR_ARG, "linear index %lld outside the volume (%lld voxels)", (long long)lin[i],
                  (long long)total);
  if (n == 0) return GRS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->v_lin, sizeof(int64_t) * (size_t)n);
  ENSURE(c, c->v_out, sizeof(double) * (size_t)n);
  HIPCHK(c, hipMemcpyAsync(c->v_lin.p, lin, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_gather_f64(P<double>(c->v_vol[which]), P<int64_t>(c->v_lin), n, P<double>(c->v_out), c->st));
  HIPCHK(c, hipMemcpyAsync(out, c->v_out.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_set_volume_at(grs_ctx *c, int32_t which, const int64_t *lin, int64_t n, const double *values) {
  if (!c) return GRS_ERR_ARG;
  if (which < 0 || which > 3 || n < 0 || (n > 0 && (!lin || !values))) return fail(c, GRS_ERR_ARG, "invalid argument");
  if (!c->v_set[which]) return fail(c, GRS_ERR_STATE, "volume %d not set", which);
  const int64_t total = c->v_dim[0] * c->v_dim[1] * c->v_dim[2];
  for (int64_t i = 0; i < n; i++)
    if (lin[i] < 0 || lin[i] >= total)
      return fail(c, GRS_ERR_ARG, "linear index %lld outside the volume (%lld voxels)", (long long)lin[i],
                  (long long)total);
  {
    /* a repeated index would leave the written value to the hardware's order */
    std::vector<int64_t> sorted(lin, lin + n);
    std::sort(sorted.begin(), sorted.end());
    for (int64_t i = 1; i < n; i++)
      if (sorted[i] == sorted[i - 1]) return fail(c, GRS_ERR_ARG, "linear index %lld repeated", (long long)sorted[i]);
  }
  if (n == 0) return GRS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->v_lin, sizeof(int64_t) * (size_t)n);
  ENSURE(c, c->v_out, sizeof(double) * (size_t)n);
  HIPCHK(c, hipMemcpyAsync(c->v_lin.p, lin, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->v_out.p, values, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_scatter_f64(P<double>(c->v_vol[which]), P<int64_t>(c->v_lin), P<double>(c->v_out), n, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_grow_region(grs_ctx *c, const int32_t *seed, double indexed_comp, const double *U, double spot_L,
                           double drop_off, double maxD, double maxDmedian, const double *p1, double *center,
                           int32_t *replace_center, int64_t *n_indexed, int32_t *indexed, int64_t cap,
                           int64_t *n_checked) {
  GRS_NO_CONTRACT
  if (!c) return GRS_ERR_ARG;
  if (!seed || !U || !p1 || !center || !replace_center || !n_indexed || cap < 0 || (cap > 0 && !indexed))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  for (int k = 0; k < 4; k++)
    if (!c->v_set[k]) return fail(c, GRS_ERR_STATE, "volume %d not set (grs_set_volume)", k);
  const int64_t nx = c->v_dim[0], ny = c->v_dim[1], nz = c->v_dim[2];
  const int64_t dim[3] = {nx, ny, nz};
  for (int k = 0; k < 3; k++)
    if (seed[k] < 1 || seed[k] > dim[k]) return fail(c, GRS_ERR_ARG, "seed outside the volume");
  /* grow_indexed_region_cuda.m:44-66 */
  const grs_dgeo &g = c->hgeo;
  const double boxLd = __builtin_ceil((spot_L * ((g.py + g.pz) / 2.0)) / ((g.Lsd / g.Lss) * p1[4]));
  if (!(boxLd >= 0.0 && boxLd < 4096.0)) return fail(c, GRS_ERR_ARG, "bounding box half width %g", boxLd);
  const int64_t boxL = (int64_t)boxLd;
  int64_t lo[3], hi[3], bdim[3];
  double cp[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = seed[k] - boxL > 0 ? seed[k] - boxL : 1;
    hi[k] = seed[k] + boxL < dim[k] ? seed[k] + boxL : dim[k];
    bdim[k] = hi[k] - lo[k] + 1;
    cp[k] = grs_round((double)(hi[k] - lo[k])) / 2.0 + (double)lo[k];
  }
  const int64_t nbox = bdim[0] * bdim[1] * bdim[2];
  const int64_t seed_lin = (seed[0] - 1) + nx * ((seed[1] - 1) + ny * (seed[2] - 1));
  /* I(x,y,z) = indexed_comp, also the seed's completeness on return */
  HIPCHK(c, hipMemcpyAsync(P<double>(c->v_vol[0]) + seed_lin, &indexed_comp, sizeof(double), hipMemcpyHostToDevice,
                           c->st));
  ENSURE(c, c->r_rows, sizeof(double) * 6 * nbox);
  ENSURE(c, c->r_counts, sizeof(uint32_t) * 4);
  HIPCHK(c, grs_launch_fill_u32(P<uint32_t>(c->r_counts), 4, 0u, c->st));
  HIPCHK(c, grs_launch_region_list(P<double>(c->v_vol[0]), P<double>(c->v_vol[1]), P<double>(c->v_vol[2]), nx, ny,
                                   lo, bdim, cp, boxLd, seed_lin, indexed_comp, P<double>(c->r_rows),
                                   P<uint32_t>(c->r_counts), c->st));
  uint32_t cnt[4] = {0, 0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(cnt, c->r_counts.p, sizeof(uint32_t) * 2, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const int64_t M = cnt[0];
  if (n_checked) *n_checked = cnt[1];
  /* cuda_grow_indexed_region on the listed voxels, then the growth rule and
   * the write-back */
  double p14[14];
  for (int k = 0; k < 11; k++) p14[k] = p1[k];
  p14[11] = indexed_comp;
  p14[12] = drop_off;
  p14[13] = maxDmedian;
  std::vector<int64_t> acc_lin;
  std::vector<double> acc_C;
  if (M > 0) {
    ENSURE(c, c->d_p1, sizeof(double) * 14);
    ENSURE(c, c->d_U, sizeof(double) * 9);
    ENSURE(c, c->d_sub, sizeof(double) * 9);
    ENSURE(c, c->d_valid, sizeof(int32_t));
    ENSURE(c, c->d_pos, sizeof(double) * 3 * M);
    ENSURE(c, c->d_ns, sizeof(int32_t) * M);
    ENSURE(c, c->d_ni, sizeof(int32_t) * M);
    ENSURE(c, c->d_C, sizeof(double) * M);
    ENSURE(c, c->d_med, sizeof(float) * M);
    ENSURE(c, c->r_acc_lin, sizeof(int64_t) * M);
    ENSURE(c, c->r_acc_C, sizeof(double) * M);
    HIPCHK(c, hipMemcpyAsync(c->d_p1.p, p14, sizeof(double) * 14, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_U.p, U, sizeof(double) * 9, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), 1, P<double>(c->d_sub),
                                    P<int32_t>(c->d_valid), c->st));
    HIPCHK(c, grs_launch_voxel_pos(P<double>(c->r_rows), M, 6, P<double>(c->d_p1), P<double>(c->d_pos), c->st));
    s = run_eval(c, GRS_MAP_ONE_U, M, M, nullptr, nullptr, GRS_METHOD_FAST, P<int32_t>(c->d_ns), P<int32_t>(c->d_ni),
                 P<double>(c->d_C), P<float>(c->d_med));
    if (s != GRS_OK) return s;
    HIPCHK(c, grs_launch_region_accept(P<double>(c->r_rows), M, P<int32_t>(c->d_ns), P<int32_t>(c->d_ni),
                                       P<double>(c->d_C), P<float>(c->d_med), P<double>(c->d_p1), nx, ny,
                                       P<double>(c->v_vol[0]), P<double>(c->v_vol[2]), P<double>(c->v_vol[3]),
                                       P<int64_t>(c->r_acc_lin), P<double>(c->r_acc_C),
                                       P<uint32_t>(c->r_counts) + 2, c->st));
    HIPCHK(c, hipMemcpyAsync(cnt + 2, P<uint32_t>(c->r_counts) + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    acc_lin.resize(cnt[2]);
    acc_C.resize(cnt[2]);
    if (cnt[2] > 0) {
      HIPCHK(c, hipMemcpyAsync(acc_lin.data(), c->r_acc_lin.p, sizeof(int64_t) * cnt[2], hipMemcpyDeviceToHost,
                               c->st));
      HIPCHK(c, hipMemcpyAsync(acc_C.data(), c->r_acc_C.p, sizeof(double) * cnt[2], hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
    }
  }
  /* J = seed and accepted voxels in find() order with their completeness
   * after the write-back (:130-151) */
  std::vector<std::pair<int64_t, double>> J;
  J.reserve(acc_lin.size() + 1);
  bool seed_acc = false;
  for (size_t k = 0; k < acc_lin.size(); k++) {
    J.push_back({acc_lin[k], acc_C[k]});
    seed_acc = seed_acc || acc_lin[k] == seed_lin;
  }
  if (!seed_acc) J.push_back({seed_lin, indexed_comp});
  std::sort(J.begin(), J.end());
  const double sx = seed[0], sy = seed[1], sz = seed[2];
  double sum = indexed_comp;
  double sp[3] = {indexed_comp * sx, indexed_comp * sy, indexed_comp * sz};
  const double c0[3] = {sp[0] / sum, sp[1] / sum, sp[2] / sum};
  if (J.size() > 1) {
    for (const auto &j : J) {
      const int64_t x = j.first % nx + 1, y = (j.first / nx) % ny + 1, z = j.first / (nx * ny) + 1;
      sum = sum + j.second;
      sp[0] = sp[0] + j.second * (double)x;
      sp[1] = sp[1] + j.second * (double)y;
      sp[2] = sp[2] + j.second * (double)z;
    }
  }
  const double c1[3] = {sp[0] / sum, sp[1] / sum, sp[2] / sum};
  const double d0 = c1[0] - c0[0], d1 = c1[1] - c0[1], d2 = c1[2] - c0[2];
  const double dist = grs_sqrt((d0 * d0 + d1 * d1) + d2 * d2);
  const double alt = (maxD / 10000.0) * (double)J.size();
  const double thr = maxD != maxD ? alt : (alt != alt ? maxD : (maxD > alt ? maxD : alt));
  if (dist > thr) {
    for (int k = 0; k < 3; k++) center[k] = grs_round(c1[k]);
    *replace_center = 1;
  } else {
    for (int k = 0; k < 3; k++) center[k] = c0[k];
    *replace_center = 0;
  }
  *n_indexed = (int64_t)J.size();
  for (int64_t k = 0; k < (int64_t)J.size() && k < cap; k++) {
    indexed[3 * k] = (int32_t)(J[k].first % nx + 1);
    indexed[3 * k + 1] = (int32_t)((J[k].first / nx) % ny + 1);
    indexed[3 * k + 2] = (int32_t)(J[k].first / (nx * ny) + 1);
  }
  return GRS_OK;
}

/* ---- grain centre-of-mass fit ------------------------------------------------------ */

grs_status grs_fit_grain_com(grs_ctx *c, const double *rows, const int64_t *row_ptr, int64_t M, int32_t max_fun_evals,
                             int32_t max_iter, double tol_fun, double tol_x, double *x_out, double *fval,
                             int32_t *n_fun_evals, int32_t *n_iter) {
  GRS_NO_CONTRACT
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || !row_ptr || (M > 0 && (!rows || !x_out || !fval || !n_fun_evals || !n_iter)) || max_fun_evals < 1 ||
      max_iter < 1 || !(tol_fun >= 0.0) || !(tol_x >= 0.0))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  if (M == 0) return GRS_OK;
  if (row_ptr[0] != 0) return fail(c, GRS_ERR_ARG, "row_ptr must start at 0");
  for (int64_t m = 0; m < M; m++) {
    const int64_t n = row_ptr[m + 1] - row_ptr[m];
    if (n < 1) return fail(c, GRS_ERR_ARG, "fit %lld has no pair rows", (long long)m);
    /* (point, pair) work items of one step are counted in 32 bits */
    if (n > 0x7fffffff / (GRS_NM_NMAX + 1)) return fail(c, GRS_ERR_ARG, "fit %lld has too many pair rows", (long long)m);
  }
  if (M > 0x7fffffff) return fail(c, GRS_ERR_ARG, "too many fits in one call");
  /* each iteration feeds at most three requests: this bounds every wave of k_com_nm */
  if (max_iter > 100000) return fail(c, GRS_ERR_ARG, "max_iter above 100000");
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  const int64_t total = row_ptr[M];
  /* fit_grain_COM.m:7,17-23: x0 = [position, u2euler_corr(U)] from the first
   * row, bounds x0 -/+ 0.1 mm and the angles -/+ 3 degrees within [0, 360] */
  std::vector<grs_nm> nm((size_t)M);
  for (int64_t m = 0; m < M; m++) {
    const double *r = rows + 23 * row_ptr[m];
    double x0[6], lo[6], hi[6];
    x0[0] = r[9];
    x0[1] = r[10];
    x0[2] = r[11];
    grs_u2euler_corr(r, x0 + 3);
    for (int k = 0; k < 3; k++) {
      lo[k] = x0[k] - 0.1;
      hi[k] = x0[k] + 0.1;
    }
    for (int k = 3; k < 6; k++) {
      lo[k] = grs_nm_max(x0[k] - 3.0, 0.0);
      const double u = x0[k] + 3.0;
      hi[k] = u != u ? 360.0 : (u < 360.0 ? u : 360.0);
    }
    nm[m].max_fe = max_fun_evals;
    nm[m].max_it = max_iter;
    nm[m].tol_f = tol_fun;
    nm[m].tol_x = tol_x;
    nm[m].start(6, x0, lo, hi);
  }
  ENSURE(c, c->m_rows, sizeof(double) * 23 * (size_t)total);
  ENSURE(c, c->m_rptr, sizeof(int64_t) * (size_t)(M + 1));
  ENSURE(c, c->m_nm, sizeof(grs_nm) * (size_t)M);
  ENSURE(c, c->m_cs, sizeof(double) * 2 * (size_t)total);
  ENSURE(c, c->m_dis, sizeof(double) * (GRS_NM_NMAX + 1) * (size_t)total);
  HIPCHK(c, hipMemcpyAsync(c->m_rows.p, rows, sizeof(double) * 23 * (size_t)total, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->m_rptr.p, row_ptr, sizeof(int64_t) * (size_t)(M + 1), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->m_nm.p, nm.data(), sizeof(grs_nm) * (size_t)M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_com_nm(P<grs_dgeo>(c->d_geo), P<double>(c->m_rows), P<int64_t>(c->m_rptr), M, total,
                              P<double>(c->m_cs), P<double>(c->m_dis), P<grs_nm>(c->m_nm), c->st));
  HIPCHK(c, hipMemcpyAsync(nm.data(), c->m_nm.p, sizeof(grs_nm) * (size_t)M, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int64_t m = 0; m < M; m++) {
    nm[m].result(x_out + 6 * m, fval + m);
    n_fun_evals[m] = nm[m].nf;
    n_iter[m] = nm[m].it;
  }
  return GRS_OK;
}

grs_status grs_euler_to_rod_quat(const double *euler_deg, int64_t M, double *rod_quat, int32_t *ok) {
  if (M < 0 || (M > 0 && (!euler_deg || !rod_quat || !ok))) return GRS_ERR_ARG;
  for (int64_t m = 0; m < M; m++) {
    double *o = rod_quat + 7 * m;
    ok[m] = grs_angle2rod_zxz_deg(euler_deg + 3 * m, o);
    grs_rod2quat(o, o + 3);
  }
  return GRS_OK;
}

grs_status grs_euler2u_deg(const double *euler_deg, int64_t M, double *U) {
  if (M < 0 || (M > 0 && (!euler_deg || !U))) return GRS_ERR_ARG;
  for (int64_t m = 0; m < M; m++) grs_U_from_euler_deg(euler_deg[3 * m], euler_deg[3 * m + 1], euler_deg[3 * m + 2],
                                                       U + 9 * m);
  return GRS_OK;
}

grs_status grs_u2euler(const double *U_colmajor, int64_t M, double *euler_deg) {
  if (M < 0 || (M > 0 && (!U_colmajor || !euler_deg))) return GRS_ERR_ARG;
  for (int64_t m = 0; m < M; m++) grs_u2euler_corr(U_colmajor + 9 * m, euler_deg + 3 * m);
  return GRS_OK;
}

/* ---- G-vector matching score ---------------------------------------------------- */

grs_status grs_forward_gt(grs_ctx *c, const double *pos, const double *q, int64_t N, const double *fam_sq,
                          const double *d_possible, int32_t nfam, int32_t semantics, double *out) {
  GRS_NO_CONTRACT
  if (!c) return GRS_ERR_ARG;
  if (N < 0 || !pos || (N > 0 && (!q || !out)) || nfam < 1 || !fam_sq || !d_possible)
    return fail(c, GRS_ERR_ARG, "invalid argument");
  if (semantics != GRS_GT_MATCH && semantics != GRS_GT_MATCH_FUN)
    return fail(c, GRS_ERR_ARG, "unknown Gt semantics %d", semantics);
  HIPCHK(c, hipSetDevice(c->device));
  grs_status s = check_ready(c);
  if (s != GRS_OK) return s;
  if (c->p_nframes != c->nrot)
    return fail(c, GRS_ERR_STATE, "experimental spots of %d projections needed (grs_set_frame_spots)", c->nrot);
  if (N == 0) return GRS_OK;
  const int32_t F = nfam < 4 ? nfam : 4;
  const grs_dgeo &g = c->hgeo;
  /* ForwardCalc_OR_space_Gt.m:129-133 = calc_Gt_match_fun.m:12-16 */
  const double v[3] = {g.Lsd, g.dety00 - g.RAO, g.detz00};
  /* calc_Gt_match shifts the position once, pos(:,2) - RotAxisOffset
   * (ForwardCalc_OR_space_Gt.m:135); calc_Gt_match_fun does not */
  const double pos_eff[3] = {pos[0], semantics == GRS_GT_MATCH ? pos[1] - g.RAO : pos[1], pos[2]};
  double d_tr[3];
  for (int r = 0; r < 3; r++) d_tr[r] = (g.R[r] * v[0] + g.R[3 + r] * v[1]) + g.R[6 + r] * v[2];
  const double Lsd_tr = d_tr[0] / g.R[0];
  const double dety00_tr = g.R[1] * Lsd_tr - d_tr[1];
  const double detz00_tr = g.R[2] * Lsd_tr - d_tr[2];
  const double L_tr = g.Lss + Lsd_tr;
  std::vector<int32_t> fam((size_t)c->H, -1);
  for (int32_t h = 0; h < c->H; h++) {
    const double *x = &c->hkl[3 * (size_t)h];
    const double s2 = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2];
    for (int32_t k = 0; k < F; k++)
      if (s2 == fam_sq[k]) {
        fam[h] = k;
        break;
      }
  }
  double dposs[4] = {0, 0, 0, 0}, glen[4] = {0, 0, 0, 0};
  for (int32_t k = 0; k < F; k++) {
    dposs[k] = d_possible[k];
    glen[k] = (2.0 * GRS_PI) / d_possible[k];
  }
  const int64_t nsp = c->p_n;
  const int64_t cap = (int64_t)c->nrot * c->H;
  int64_t waves = grs_gt_wave_slots(c->device);
  if (waves > N) waves = N;
  const int64_t wslots = ((waves + 3) / 4) * 4;
  ENSURE(c, c->t_pos, sizeof(double) * 3);
  ENSURE(c, c->t_proj, sizeof(double) * 18 * (size_t)c->nrot);
  ENSURE(c, c->t_units, sizeof(double) * 12 * (size_t)(nsp > 0 ? nsp : 1));
  ENSURE(c, c->t_fam, sizeof(int32_t) * (size_t)c->H);
  ENSURE(c, c->t_dposs, sizeof(double) * 4);
  ENSURE(c, c->t_glen, sizeof(double) * 4);
  ENSURE(c, c->t_scratch, sizeof(double) * (size_t)(wslots * cap));
  ENSURE(c, c->t_out, sizeof(double) * 4 * (size_t)N);
  ENSURE(c, c->d_q, sizeof(double) * 4 * N);
  ENSURE(c, c->d_sub, sizeof(double) * 9 * N);
  ENSURE(c, c->d_valid, sizeof(int32_t) * N);
  HIPCHK(c, hipMemcpyAsync(c->t_pos.p, pos_eff, sizeof(double) * 3, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->t_fam.p, fam.data(), sizeof(int32_t) * fam.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->t_dposs.p, dposs, sizeof(double) * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->t_glen.p, glen, sizeof(double) * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_q.p, q, sizeof(double) * 4 * N, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_sub_from_q(P<grs_dgeo>(c->d_geo), P<double>(c->d_q), N, P<double>(c->d_sub),
                                  P<int32_t>(c->d_valid), c->st));
  HIPCHK(c, grs_launch_gt_proj(P<grs_dgeo>(c->d_geo), P<double>(c->t_pos), P<double>(c->d_cs), c->nrot, L_tr,
                               P<double>(c->t_proj), c->st));
  HIPCHK(c, grs_launch_gt_spots(P<grs_dgeo>(c->d_geo), P<double>(c->t_proj), P<int64_t>(c->p_fptr), c->nrot,
                                P<double>(c->p_fsp), nsp, Lsd_tr, dety00_tr, detz00_tr, P<double>(c->t_dposs), F,
                                P<double>(c->t_units), c->st));
  /* spots sorted by unit-vector z per (projection, family) when every
   * projection fits the sort's LDS; otherwise every spot is visited */
  const bool sorted = nsp > 0 && c->p_maxframe <= grs_gt_sort_max();
  if (sorted) {
    ENSURE(c, c->t_skey, sizeof(double) * 4 * (size_t)nsp);
    ENSURE(c, c->t_su, sizeof(double) * 12 * (size_t)nsp);
    ENSURE(c, c->t_sorig, sizeof(int32_t) * 4 * (size_t)nsp);
    ENSURE(c, c->t_bst, sizeof(int32_t) * 4 * (GRS_GT_NB + 1) * (size_t)c->nrot);
    ENSURE(c, c->t_slen, sizeof(int32_t) * 4 * (size_t)c->nrot);
  }
  if (c->timing) HIPCHK(c, timing_begin(c));
  if (sorted)
    HIPCHK(c, grs_launch_gt_sort(P<int64_t>(c->p_fptr), c->nrot, P<double>(c->t_units), nsp, c->p_maxframe,
                                 P<double>(c->t_skey), P<double>(c->t_su), P<int32_t>(c->t_sorig),
                                 P<int32_t>(c->t_bst), P<int32_t>(c->t_slen), c->st));
  HIPCHK(c, grs_launch_gt_score(P<grs_dgeo>(c->d_geo), P<double>(c->d_hkl), P<int32_t>(c->t_fam),
                                P<double>(c->t_glen), P<double>(c->d_sub), P<int32_t>(c->d_valid), N,
                                P<double>(c->t_proj), P<int64_t>(c->p_fptr), P<double>(c->t_units),
                                sorted ? P<double>(c->t_skey) : nullptr, sorted ? P<double>(c->t_su) : nullptr,
                                sorted ? P<int32_t>(c->t_sorig) : nullptr, sorted ? P<int32_t>(c->t_bst) : nullptr,
                                sorted ? P<int32_t>(c->t_slen) : nullptr, nsp, semantics == GRS_GT_MATCH ? 1 : 0,
                                P<double>(c->t_scratch), cap, waves,
                                P<double>(c->t_out), c->st));
  if (c->timing) HIPCHK(c, timing_end(c));
  HIPCHK(c, hipMemcpyAsync(out, c->t_out.p, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

/* ---- candidate generation ------------------------------------------------- */

grs_status grs_local_sampling(grs_ctx *c, const double *centres, int64_t K, int32_t nr, double gridsize,
                              double *out_q, int64_t cap, int64_t *out_n) {
  if (!c) return GRS_ERR_ARG;
  if (K < 0 || nr < 0 || cap < 1 || (K > 0 && (!centres || !out_q || !out_n)) || !std::isfinite(gridsize))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  if (K == 0) return GRS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> grids((size_t)K * 3 * GRS_LS_MAX_GRID, 0.0);
  std::vector<int32_t> gn((size_t)K * 3);
  const double half = (gridsize * (double)nr) / 2.0;
  for (int64_t k = 0; k < K; k++) {
    int64_t rows = 1;
    for (int j = 0; j < 3; j++) {
      const double cj = centres[4 * k + 1 + j];
      const int64_t n = grs_colon(cj - half, gridsize, cj + half, &grids[(size_t)(3 * k + j) * GRS_LS_MAX_GRID],
                                     GRS_LS_MAX_GRID);
      if (n < 0) return fail(c, GRS_ERR_ARG, "centre %lld: grid of more than %d values", (long long)k, GRS_LS_MAX_GRID);
      gn[3 * k + j] = (int32_t)n;
      rows *= n;
    }
    if (1 + rows > GRS_LS_MAX_ROWS)
      return fail(c, GRS_ERR_ARG, "centre %lld: %lld rows exceed %d", (long long)k, (long long)(1 + rows),
                  GRS_LS_MAX_ROWS);
    if (1 + rows > cap)
      return fail(c, GRS_ERR_ARG, "centre %lld: up to %lld rows, capacity %lld", (long long)k, (long long)(1 + rows),
                  (long long)cap);
  }
  ENSURE(c, c->c_cen, sizeof(double) * 4 * K);
  ENSURE(c, c->c_grid, sizeof(double) * grids.size());
  ENSURE(c, c->c_gn, sizeof(int32_t) * gn.size());
  ENSURE(c, c->c_out, sizeof(double) * 4 * (size_t)K * cap);
  ENSURE(c, c->c_n, sizeof(int64_t) * K);
  ENSURE(c, c->d_bad, sizeof(uint32_t));
  HIPCHK(c, hipMemcpyAsync(c->c_cen.p, centres, sizeof(double) * 4 * K, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->c_grid.p, grids.data(), sizeof(double) * grids.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->c_gn.p, gn.data(), sizeof(int32_t) * gn.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_bad.p, 0, sizeof(uint32_t), c->st));
  HIPCHK(c, grs_launch_local_sampling(P<double>(c->c_cen), P<double>(c->c_grid), P<int32_t>(c->c_gn), K,
                                      P<double>(c->c_out), cap, P<int64_t>(c->c_n), P<uint32_t>(c->d_bad), c->st));
  uint32_t bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(out_n, c->c_n.p, sizeof(int64_t) * K, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(out_q, c->c_out.p, sizeof(double) * 4 * (size_t)K * cap, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (bad) return fail(c, GRS_ERR_ARG, "a grid point has q1^2+q2^2+q3^2 > 1");
  return GRS_OK;
}

grs_status grs_find_neighbor_ors(grs_ctx *c, const int32_t *grain_id, int64_t nx, int64_t ny, int64_t nz,
                                 const int32_t *vox, int64_t M, int32_t search_radius, int32_t nr_max,
                                 int32_t *id_all, int64_t *ind, int64_t *n_total) {
  if (!c) return GRS_ERR_ARG;
  if (nx < 1 || ny < 1 || nz < 1 || M < 0 || search_radius < 0 || nr_max < 1 || !grain_id || !n_total ||
      (M > 0 && (!vox || !id_all || !ind)))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  const int64_t nvol = nx * ny * nz;
  int32_t maxid = 0;
  for (int64_t i = 0; i < nvol; i++) maxid = grain_id[i] > maxid ? grain_id[i] : maxid;
  if (maxid <= 0) return fail(c, GRS_ERR_ARG, "the grain-id volume holds no positive id");
  if (maxid >= 32 * 16384) return fail(c, GRS_ERR_ARG, "grain ids above %d are not supported", 32 * 16384 - 1);
  const int64_t dims[3] = {nx, ny, nz};
  for (int64_t m = 0; m < M; m++)
    for (int j = 0; j < 3; j++)
      if (vox[3 * m + j] < 1 || vox[3 * m + j] > dims[j])
        return fail(c, GRS_ERR_ARG, "voxel %lld lies outside the volume", (long long)m);
  *n_total = 0;
  if (M == 0) return GRS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const int32_t W = (maxid >> 5) + 1;
  ENSURE(c, c->c_gid, sizeof(int32_t) * nvol);
  ENSURE(c, c->c_vox, sizeof(int32_t) * 3 * M);
  ENSURE(c, c->c_ids, sizeof(int32_t) * (size_t)M * nr_max);
  ENSURE(c, c->c_nout, sizeof(int32_t) * M);
  ENSURE(c, c->d_bad, sizeof(uint32_t));
  HIPCHK(c, hipMemcpyAsync(c->c_gid.p, grain_id, sizeof(int32_t) * nvol, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->c_vox.p, vox, sizeof(int32_t) * 3 * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_bad.p, 0, sizeof(uint32_t), c->st));
  HIPCHK(c, grs_launch_find_neighbors(P<int32_t>(c->c_gid), nx, ny, nz, P<int32_t>(c->c_vox), M, search_radius, nr_max,
                                      W, P<int32_t>(c->c_ids), P<int32_t>(c->c_nout), P<uint32_t>(c->d_bad), c->st));
  std::vector<int32_t> nout((size_t)M), ids((size_t)M * nr_max);
  uint32_t bad = 0;
  HIPCHK(c, hipMemcpyAsync(&bad, c->d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(nout.data(), c->c_nout.p, sizeof(int32_t) * M, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(ids.data(), c->c_ids.p, sizeof(int32_t) * ids.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (bad) return fail(c, GRS_ERR_STATE, "a search box grew to the whole volume without a positive id");
  /* pack one voxel after the other, 1-based [first last] (find_neighbor_ORs.m:26-29) */
  int64_t count = 0;
  for (int64_t m = 0; m < M; m++) {
    const int32_t n = nout[m];
    for (int32_t k = 0; k < n; k++) id_all[count + k] = ids[(size_t)m * nr_max + k];
    count += n;
    ind[2 * m] = count - n + 1;
    ind[2 * m + 1] = count;
  }
  *n_total = count;
  return GRS_OK;
}

/* ---- input preparation ---------------------------------------------------- */

grs_status grs_render_spots(grs_ctx *c, const double *pos, const int32_t *grain, int64_t M, const double *U,
                            int64_t G, int32_t clear) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || G < 0 || (M > 0 && (!pos || !grain || !U || G == 0))) return fail(c, GRS_ERR_ARG, "null argument");
  if (!c->geo_set) return fail(c, GRS_ERR_STATE, "geometry not set");
  HIPCHK(c, hipSetDevice(c->device));
  for (int64_t i = 0; i < M; i++)
    if (grain[i] < 0 || grain[i] >= G) return fail(c, GRS_ERR_ARG, "grain[%lld] = %d outside 0..%lld",
                                                   (long long)i, grain[i], (long long)(G - 1));
  const int64_t nz = (int64_t)c->hgeo.detzsize, ny = (int64_t)c->hgeo.detysize, nr = c->nrot;
  if ((double)nz != c->hgeo.detzsize || (double)ny != c->hgeo.detysize)
    return fail(c, GRS_ERR_ARG, "rendering needs an integer detector size");
  const int64_t n = nz * ny * nr;
  if (c->mnz != nz || c->mny != ny || c->mnrot != nr) clear = 1;
  ENSURE(c, c->d_mask, (size_t)n);
  c->mnz = nz;
  c->mny = ny;
  c->mnrot = nr;
  if (clear) HIPCHK(c, hipMemsetAsync(c->d_mask.p, 0, (size_t)n, c->st));
  /* the renderer indexes the mask with the detector size, so point the geometry at it */
  c->hgeo.nz = nz;
  c->hgeo.ny = ny;
  ENSURE(c, c->d_geo, sizeof(grs_dgeo));
  HIPCHK(c, hipMemcpyAsync(c->d_geo.p, &c->hgeo, sizeof(grs_dgeo), hipMemcpyHostToDevice, c->st));
  ENSURE(c, c->d_U, sizeof(double) * 9 * G);
  ENSURE(c, c->d_sub, sizeof(double) * 9 * G);
  ENSURE(c, c->d_valid, sizeof(int32_t) * G);
  HIPCHK(c, hipMemcpyAsync(c->d_U.p, U, sizeof(double) * 9 * G, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), G, P<double>(c->d_sub),
                                  P<int32_t>(c->d_valid), c->st));
  const int64_t chunk = (int64_t)1 << 22;
  for (int64_t b = 0; b < M; b += chunk) {
    int64_t m = M - b < chunk ? M - b : chunk;
    ENSURE(c, c->d_pos, sizeof(double) * 3 * m);
    ENSURE(c, c->d_grain, sizeof(int32_t) * m);
    HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos + 3 * b, sizeof(double) * 3 * m, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_grain.p, grain + b, sizeof(int32_t) * m, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, grs_launch_render(P<grs_dgeo>(c->d_geo), P<double>(c->d_pos), P<int32_t>(c->d_grain), m, c->nrot,
                                P<double>(c->d_sub), P<double>(c->d_hkl), P<double>(c->d_cs), P<uint8_t>(c->d_mask),
                                c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  /* restore the stack shape the scoring kernels index with */
  return upload_geo(c);
}

grs_status grs_simulate_spots(grs_ctx *c, const double *pos, const int32_t *grain, int64_t M, const double *U,
                              int64_t G, int64_t cap, int32_t *vox_out, int32_t *om_out, int32_t *h_out,
                              double *yz_out, int64_t *n_out) {
  if (!c) return GRS_ERR_ARG;
  if (M < 0 || G < 0 || cap < 0 || !n_out || (M > 0 && (!pos || !grain || !U || G == 0)) ||
      (cap > 0 && (!vox_out || !om_out || !h_out || !yz_out)))
    return fail(c, GRS_ERR_ARG, "invalid argument");
  if (!c->geo_set) return fail(c, GRS_ERR_STATE, "geometry not set");
  for (int64_t i = 0; i < M; i++)
    if (grain[i] < 0 || grain[i] >= G) return fail(c, GRS_ERR_ARG, "grain[%lld] out of range", (long long)i);
  HIPCHK(c, hipSetDevice(c->device));
  *n_out = 0;
  if (M == 0) return GRS_OK;
  ENSURE(c, c->d_U, sizeof(double) * 9 * G);
  ENSURE(c, c->d_sub, sizeof(double) * 9 * G);
  ENSURE(c, c->d_valid, sizeof(int32_t) * G);
  ENSURE(c, c->d_pos, sizeof(double) * 3 * M);
  ENSURE(c, c->d_grain, sizeof(int32_t) * M);
  ENSURE(c, c->d_bad, sizeof(unsigned long long));
  int64_t cc = cap > 0 ? cap : 1;
  ENSURE(c, c->d_evox, sizeof(int32_t) * cc);
  ENSURE(c, c->d_eu_idx, sizeof(int32_t) * cc);
  ENSURE(c, c->d_ns, sizeof(int32_t) * cc);
  ENSURE(c, c->d_y, sizeof(double) * 2 * cc);
  HIPCHK(c, hipMemcpyAsync(c->d_U.p, U, sizeof(double) * 9 * G, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_pos.p, pos, sizeof(double) * 3 * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->d_grain.p, grain, sizeof(int32_t) * M, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipMemsetAsync(c->d_bad.p, 0, sizeof(unsigned long long), c->st));
  HIPCHK(c, grs_launch_sub_from_U(P<grs_dgeo>(c->d_geo), P<double>(c->d_U), G, P<double>(c->d_sub),
                                  P<int32_t>(c->d_valid), c->st));
  HIPCHK(c, grs_launch_sim_events(P<grs_dgeo>(c->d_geo), P<double>(c->d_pos), P<int32_t>(c->d_grain), M, c->nrot,
                                  P<double>(c->d_sub), P<double>(c->d_hkl), P<double>(c->d_cs), cap,
                                  P<unsigned long long>(c->d_bad), P<int32_t>(c->d_evox), P<int32_t>(c->d_eu_idx),
                                  P<int32_t>(c->d_ns), P<double>(c->d_y), c->st));
  unsigned long long total = 0;
  HIPCHK(c, hipMemcpyAsync(&total, c->d_bad.p, sizeof(total), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  *n_out = (int64_t)total;
  int64_t k = (int64_t)total < cap ? (int64_t)total : cap;
  if (k == 0) return GRS_OK;
  std::vector<int32_t> v(k), o(k), h(k);
  std::vector<double> yz(2 * k);
  HIPCHK(c, hipMemcpyAsync(v.data(), c->d_evox.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(o.data(), c->d_eu_idx.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_ns.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(yz.data(), c->d_y.p, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  std::vector<int64_t> idx(k);
  for (int64_t i = 0; i < k; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    if (v[a] != v[b]) return v[a] < v[b];
    if (o[a] != o[b]) return o[a] < o[b];
    return h[a] < h[b];
  });
  for (int64_t i = 0; i < k; i++) {
    vox_out[i] = v[idx[i]];
    om_out[i] = o[idx[i]];
    h_out[i] = h[idx[i]];
    yz_out[2 * i] = yz[2 * idx[i]];
    yz_out[2 * i + 1] = yz[2 * idx[i] + 1];
  }
  return GRS_OK;
}

grs_status grs_set_binary_stack(grs_ctx *c, const uint8_t *mask, int64_t nz, int64_t ny, int64_t nrot,
                                int32_t on_device) {
  if (!c) return GRS_ERR_ARG;
  if (!mask || nz < 1 || ny < 1 || nrot < 1) return fail(c, GRS_ERR_ARG, "invalid mask shape");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t n = nz * ny * nrot;
  ENSURE(c, c->d_mask, (size_t)n);
  HIPCHK(c, hipMemcpyAsync(c->d_mask.p, mask, (size_t)n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->mnz = nz;
  c->mny = ny;
  c->mnrot = nrot;
  return GRS_OK;
}

grs_status grs_distance_from_mask(grs_ctx *c) {
  if (!c) return GRS_ERR_ARG;
  if (c->mnz < 1) return fail(c, GRS_ERR_STATE, "no mask: render or set one first");
  HIPCHK(c, hipSetDevice(c->device));
  const int64_t nz = c->mnz, ny = c->mny, nr = c->mnrot, n = nz * ny * nr;
  ENSURE(c, c->d_g1, sizeof(int32_t) * n);
  ENSURE(c, c->d_keys, sizeof(float) * n);
  HIPCHK(c, grs_launch_edt_cols(P<uint8_t>(c->d_mask), nz, ny * nr, P<int32_t>(c->d_g1), c->st));
  int64_t per_slice = nz * ny * (int64_t)(sizeof(int32_t) + sizeof(double));
  int64_t ns = ((int64_t)1536 << 20) / (per_slice > 0 ? per_slice : 1);
  if (ns < 1) ns = 1;
  if (ns > nr) ns = nr;
  ENSURE(c, c->d_edt_v, sizeof(int32_t) * (size_t)(nz * ny * ns));
  ENSURE(c, c->d_edt_z, sizeof(double) * (size_t)(nz * ny * ns));
  for (int64_t s0 = 0; s0 < nr; s0 += ns) {
    int64_t k = nr - s0 < ns ? nr - s0 : ns;
    HIPCHK(c, grs_launch_edt_rows(P<int32_t>(c->d_g1), nz, ny, s0, k, P<int32_t>(c->d_edt_v),
                                  P<double>(c->d_edt_z), P<float>(c->d_keys), c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->snz = nz;
  c->sny = ny;
  c->snrot = nr;
  c->stack_set = true;
  c->bits_valid = false;
  if (c->geo_set) return upload_geo(c);
  return GRS_OK;
}

grs_status grs_get_mask(grs_ctx *c, uint8_t *out, int64_t n) {
  if (!c) return GRS_ERR_ARG;
  if (!out || n != c->mnz * c->mny * c->mnrot) return fail(c, GRS_ERR_ARG, "output size does not match the mask");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, c->d_mask.p, (size_t)n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_get_distance_stack(grs_ctx *c, float *out, int64_t n) {
  if (!c) return GRS_ERR_ARG;
  if (!c->stack_set) return fail(c, GRS_ERR_STATE, "distance stack not set");
  if (!out || n != c->snz * c->sny * c->snrot) return fail(c, GRS_ERR_ARG, "output size does not match the stack");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, c->d_keys.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_device_alloc(grs_ctx *c, int64_t bytes, void **d_out) {
  if (!c || !d_out || bytes < 0) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  *d_out = nullptr;
  HIPCHK(c, hipMalloc(d_out, bytes > 0 ? (size_t)bytes : 8));
  return GRS_OK;
}

grs_status grs_device_free(grs_ctx *c, void *d) {
  if (!c) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (d) HIPCHK(c, hipFree(d));
  return GRS_OK;
}

grs_status grs_copy_to_device(grs_ctx *c, void *d_dst, const void *src, int64_t bytes) {
  if (!c || !d_dst || !src || bytes < 0) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(d_dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_device_copy(grs_ctx *c, void *d_dst, const void *d_src, int64_t bytes) {
  if (!c || !d_dst || !d_src || bytes < 0 || bytes % 4) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->timing) HIPCHK(c, timing_begin(c));
  HIPCHK(c, grs_launch_copy_u32(static_cast<const uint32_t *>(d_src), static_cast<uint32_t *>(d_dst), bytes / 4, c->st));
  if (c->timing) HIPCHK(c, timing_end(c));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_copy_to_host(grs_ctx *c, void *dst, const void *d_src, int64_t bytes) {
  if (!c || !dst || !d_src || bytes < 0) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

/* ---- measurement and diagnostics ---------------------------------------- */

grs_status grs_set_timing(grs_ctx *c, int32_t enable) {
  if (!c) return GRS_ERR_ARG;
  c->timing = enable != 0;
  return GRS_OK;
}

grs_status grs_set_autotune(grs_ctx *c, int32_t enable) {
  if (!c) return GRS_ERR_ARG;
  c->autotune = enable != 0;
  return GRS_OK;
}

grs_status grs_get_tuning(grs_ctx *c, int32_t kind, int32_t *waves_per_block, int32_t *mask_layout, int32_t *xcd_map,
                          int32_t *n_configs, double *ms_per_eval) {
  if (!c || kind < 0 || kind > 2 || !waves_per_block || !mask_layout || !xcd_map || !n_configs || !ms_per_eval)
    return GRS_ERR_ARG;
  auto it = c->tunes.find(tune_key(c, kind));
  if (it == c->tunes.end() || !it->second.done) return fail(c, GRS_ERR_STATE, "not tuned yet");
  *waves_per_block = it->second.cfg.wpb;
  *mask_layout = it->second.cfg.mask;
  *xcd_map = it->second.cfg.xcd;
  *n_configs = it->second.ncfg;
  for (int k = 0; k < GRS_TUNE_MAX; k++) ms_per_eval[k] = it->second.ms_per_eval[k];
  return GRS_OK;
}

grs_status grs_set_launch_config(grs_ctx *c, int32_t kind, int32_t waves_per_block, int32_t mask_layout,
                                 int32_t xcd_map) {
  if (!c || kind < 0 || kind > 2) return GRS_ERR_ARG;
  const uint64_t key = tune_key(c, kind);
  if (waves_per_block == 0) {
    c->tunes.erase(key);
    return GRS_OK;
  }
  if (!(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4 || waves_per_block == 8) ||
      mask_layout < 0 || mask_layout > 2 ||
      (kind == 0 && mask_layout != GRS_MASK_BITS) || xcd_map < 0 || xcd_map > 1)
    return fail(c, GRS_ERR_ARG, "invalid launch configuration");
  grs_tune_state t;
  t.done = true;
  t.cfg = tune_cfg{waves_per_block, mask_layout, xcd_map};
  t.ncfg = 0;
  c->tunes[key] = t;
  return GRS_OK;
}

grs_status grs_set_graphs(grs_ctx *c, int32_t enable) {
  if (!c) return GRS_ERR_ARG;
  c->graphs = enable != 0;
  return GRS_OK;
}

grs_status grs_graph_stats(grs_ctx *c, int64_t *captures, int64_t *replays) {
  if (!c || !captures || !replays) return GRS_ERR_ARG;
  *captures = c->g_captures;
  *replays = c->g_replays;
  return GRS_OK;
}

grs_status grs_set_audit(grs_ctx *c, int32_t enable) {
  if (!c) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (enable) {
    ENSURE(c, c->d_audit, sizeof(unsigned long long) * GRS_AUDIT_WORDS);
    HIPCHK(c, hipMemsetAsync(c->d_audit.p, 0, sizeof(unsigned long long) * GRS_AUDIT_WORDS, c->st));
  }
  c->audit = enable != 0;
  return GRS_OK;
}

grs_status grs_get_audit(grs_ctx *c, int64_t *counts, double *max_dev) {
  if (!c || !counts || !max_dev) return GRS_ERR_ARG;
  if (!c->d_audit.p) return fail(c, GRS_ERR_STATE, "audit never enabled");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long w[GRS_AUDIT_WORDS];
  HIPCHK(c, hipMemcpyAsync(w, c->d_audit.p, sizeof(w), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (int k = 0; k < 5; k++) counts[k] = (int64_t)w[k];
  counts[5] = (int64_t)w[6];
  std::memcpy(max_dev, &w[5], sizeof(double));
  return GRS_OK;
}

grs_status grs_kernel_time(grs_ctx *c, double *ms, int64_t *launches, int32_t reset) {
  if (!c) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, timing_harvest(c));
  if (ms) *ms = c->ms_total;
  if (launches) *launches = c->launches;
  if (reset) {
    c->ms_total = 0.0;
    c->launches = 0;
  }
  return GRS_OK;
}

grs_status grs_math_device(grs_ctx *c, int32_t fn, const double *x, double *y, int64_t n) {
  if (!c) return GRS_ERR_ARG;
  if (n < 0 || (n > 0 && (!x || !y))) return fail(c, GRS_ERR_ARG, "null argument");
  if (n == 0) return GRS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  ENSURE(c, c->d_x, sizeof(double) * n);
  ENSURE(c, c->d_y, sizeof(double) * n);
  HIPCHK(c, hipMemcpyAsync(c->d_x.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
  HIPCHK(c, grs_launch_math_eval(fn, P<double>(c->d_x), P<double>(c->d_y), n, c->st));
  HIPCHK(c, hipMemcpyAsync(y, c->d_y.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

grs_status grs_synchronize(grs_ctx *c) {
  if (!c) return GRS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return GRS_OK;
}

}  // extern "C"
