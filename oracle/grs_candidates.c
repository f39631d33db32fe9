/*
 * grs_candidates.c -- CPU restatement of GrainRecon's candidate generation.
 * TEST INFRASTRUCTURE ONLY, like grs_oracle.c: the checker for the engine's
 * grs_local_sampling / grs_find_neighbor_ors; the product never links it.
 *
 *   ori_local_sampling.m:12-34  -> oracle_local_sampling()
 *       (q1, q2, q3) grids bgrid/cgrid/dgrid = c - gridsize*nr/2 : gridsize :
 *       c + gridsize*nr/2 (:22-24), the centre row first (:26), then one row
 *       [sqrt(1-(b^2+c^2+d^2)) b c d] per grid point in i, j, k loop order
 *       (:27-33), then unique(q,'rows') (:34): rows sorted ascending,
 *       column 1 first, exact duplicates removed.
 *   find_neighbor_ORs.m:3-96    -> oracle_find_neighbor_ors()
 *       per voxel the box [p - r, p + r] clipped to the volume (:40-51),
 *       every GrainId > 0 inside it (:56-57), the box grown by 2 voxels per
 *       try while it holds none (:58-87), unique grain ids in ascending order
 *       (:89-94), at most nr_max of them (:13-19), packed one voxel after the
 *       other with 1-based [first last] positions (:26-29).
 *
 * The colon operator a:d:b is MATLAB's built-in; its algorithm is taken from
 * MathWorks' published description of it (the "colonop" demonstration: the
 * interval count n = round((b-a)/d), reduced by one when a+n*d overshoots b
 * by more than tol = 2*eps*max(|a|,|b|); the right end point snapped to b
 * when within tol; elements a+k*d from the left and last-k*d from the right,
 * the middle one (a+last)/2 when n is even).  No MATLAB is available to run,
 * so this part is parity unpinned; x^2 is x*x as everywhere in the oracle.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* a:d:b -> v (at most cap values); returns the element count, or -1 when the
 * count exceeds cap or an argument is not finite. */
int64_t oracle_colon(double a, double d, double b, double *v, int64_t cap) {
  if (!isfinite(a) || !isfinite(d) || !isfinite(b)) return -1;
  if (d == 0.0 || (a < b && d < 0.0) || (b < a && d > 0.0)) return 0;
  const double tol = 2.0 * DBL_EPSILON * fmax(fabs(a), fabs(b));
  const double sig = d > 0.0 ? 1.0 : -1.0;
  double n;
  if (a == floor(a) && d == 1.0) {
    n = floor(b) - a;
  } else if (a == floor(a) && d == floor(d)) {
    double q = floor(a / d);
    double r = a - q * d;
    n = floor((b - r) / d) - q;
  } else {
    n = round((b - a) / d);
    if (sig * ((a + n * d) - b) > tol) n = n - 1.0;
  }
  if (n + 1.0 > (double)cap) return -1;
  double last = a + n * d;
  if (sig * (last - b) > -tol) last = b;
  const int64_t N = (int64_t)n;
  for (int64_t k = 0; k <= N / 2; k++) v[k] = a + (double)k * d;
  for (int64_t k = 0; k <= N / 2; k++) v[N - k] = last - (double)k * d;
  if (N % 2 == 0) v[N / 2] = (a + last) / 2.0;
  return N + 1;
}

typedef struct {
  double q[4];
  int64_t idx;
} ls_row;

static int cmp_rows(const void *pa, const void *pb) {
  const ls_row *a = (const ls_row *)pa, *b = (const ls_row *)pb;
  for (int c = 0; c < 4; c++) {
    if (a->q[c] < b->q[c]) return -1;
    if (a->q[c] > b->q[c]) return 1;
  }
  return a->idx < b->idx ? -1 : (a->idx > b->idx ? 1 : 0);
}

/* centre: [q0 q1 q2 q3] as given (ori_local_sampling.m:16).  Writes the
 * sorted unique rows to out (4 per row, at most cap rows) and returns their
 * number, or -1 (capacity / non-finite grid / a row with b^2+c^2+d^2 > 1). */
int64_t oracle_local_sampling(const double *centre, int32_t nr, double gridsize, double *out, int64_t cap) {
  double half = (gridsize * (double)nr) / 2.0;
  double gb[256], gc[256], gd[256];
  int64_t nb = oracle_colon(centre[1] - half, gridsize, centre[1] + half, gb, 256);
  int64_t nc = oracle_colon(centre[2] - half, gridsize, centre[2] + half, gc, 256);
  int64_t nd = oracle_colon(centre[3] - half, gridsize, centre[3] + half, gd, 256);
  if (nb < 0 || nc < 0 || nd < 0) return -1;
  int64_t R = 1 + nb * nc * nd;
  ls_row *rows = (ls_row *)malloc(sizeof(ls_row) * (size_t)R);
  if (!rows) return -1;
  memcpy(rows[0].q, centre, 4 * sizeof(double));
  rows[0].idx = 0;
  int64_t r = 1;
  for (int64_t i = 0; i < nb; i++)
    for (int64_t j = 0; j < nc; j++)
      for (int64_t k = 0; k < nd; k++) {
        double ss = (gb[i] * gb[i] + gc[j] * gc[j]) + gd[k] * gd[k];
        if (ss > 1.0) {
          free(rows);
          return -1;
        }
        rows[r].q[0] = sqrt(1.0 - ss);
        rows[r].q[1] = gb[i];
        rows[r].q[2] = gc[j];
        rows[r].q[3] = gd[k];
        rows[r].idx = r;
        r++;
      }
  qsort(rows, (size_t)R, sizeof(ls_row), cmp_rows);
  int64_t n = 0;
  for (int64_t p = 0; p < R; p++) {
    if (p > 0 && rows[p].q[0] == rows[p - 1].q[0] && rows[p].q[1] == rows[p - 1].q[1] &&
        rows[p].q[2] == rows[p - 1].q[2] && rows[p].q[3] == rows[p - 1].q[3])
      continue;
    if (n >= cap) {
      free(rows);
      return -1;
    }
    memcpy(out + 4 * n, rows[p].q, 4 * sizeof(double));
    n++;
  }
  free(rows);
  return n;
}

/* gid: nx x ny x nz grain ids, x fastest (MATLAB column order).  vox: M rows
 * of 1-based [x y z].  id_all receives at most M*nr_max ids, ind M rows of
 * 1-based [first last].  Returns the number of ids written, or -1 when the
 * volume holds no positive id. */
int64_t oracle_find_neighbor_ors(const int32_t *gid, int64_t nx, int64_t ny, int64_t nz, const int32_t *vox,
                                 int64_t M, int32_t radius, int32_t nr_max, int32_t *id_all, int64_t *ind) {
  const int64_t dims[3] = {nx, ny, nz};
  int32_t maxid = 0;
  for (int64_t i = 0; i < nx * ny * nz; i++)
    if (gid[i] > maxid) maxid = gid[i];
  if (maxid <= 0) return -1;
  unsigned char *seen = (unsigned char *)calloc((size_t)maxid + 1, 1);
  if (!seen) return -1;
  int64_t count = 0;
  for (int64_t m = 0; m < M; m++) {
    int64_t lo[3], hi[3];
    int found = 0;
    for (int64_t tries = 0; !found; tries++) {
      const int64_t r = (int64_t)radius + 2 * tries;
      for (int j = 0; j < 3; j++) {
        const int64_t p = vox[3 * m + j];
        lo[j] = p - r > 0 ? p - r : 1;
        hi[j] = p + r < dims[j] ? p + r : dims[j];
      }
      memset(seen, 0, (size_t)maxid + 1);
      for (int64_t z = lo[2]; z <= hi[2]; z++)
        for (int64_t y = lo[1]; y <= hi[1]; y++)
          for (int64_t x = lo[0]; x <= hi[0]; x++) {
            int32_t id = gid[(x - 1) + (y - 1) * nx + (z - 1) * nx * ny];
            if (id > 0) {
              seen[id] = 1;
              found = 1;
            }
          }
    }
    int64_t nn = 0;
    for (int32_t id = 1; id <= maxid && nn < nr_max; id++)
      if (seen[id]) id_all[count + nn++] = id;
    count += nn;
    ind[2 * m] = count - nn + 1;
    ind[2 * m + 1] = count;
  }
  free(seen);
  return count;
}
