/*
 * grs_geofit.c -- the geometry-fit objective of geo_fitting.m:5-79 on the CPU
 * (TEST INFRASTRUCTURE and the CPU baseline of the geometry fit; the engine
 * never links it).  The grain-dependent terms of every matched spot are
 * computed by the caller (oracle/geo_fit.py): SW = Omega*S*pos (pos(2)
 * shifted by -RotAxisOffset, :32) and Gt = Omega*S*U*B*hkl.  For the 7
 * geometry values x = [Lsam2sou Lsam2det dety00 detz00 tilt_x tilt_y tilt_z]
 * (degrees, get_det_R.m = RotZ*RotY*RotX) this computes the spot of every
 * pair with the statements of :35-71 (P0y, P0z already shifted by
 * -RotAxisOffset, :19-20; RotAxisOffset not added back) and returns the mean
 * Euclidean distance between the rounded pixel (half away from zero) and the
 * measured centroid (:73-78).  OpenMP over spots; the partial sums of the
 * threads are added in thread order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double mround(double v) { return v >= 0.0 ? floor(v + 0.5) : -floor(-v + 0.5); }

double oracle_geo_fitting(const double *SW, const double *Gt, const double *meas, int64_t n, const double *x,
                          double P0y, double P0z, double py, double pz, double dety0, double detz0, double *sim,
                          int nthreads) {
  const double pi = 3.14159265358979311600;
  const double Lss = x[0], Lsd = x[1], dy00 = x[2], dz00 = x[3];
  const double ax = x[4] * pi / 180.0, ay = x[5] * pi / 180.0, az = x[6] * pi / 180.0;
  const double cx = cos(ax), sx = sin(ax), cy = cos(ay), sy = sin(ay), cz = cos(az), sz = sin(az);
  /* RotDet = RotZ * RotY * RotX */
  const double RX[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
  const double RY[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
  const double RZ[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
  double T[9], R[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      T[3 * i + j] = RZ[3 * i] * RY[j] + RZ[3 * i + 1] * RY[3 + j] + RZ[3 * i + 2] * RY[6 + j];
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      R[3 * i + j] = T[3 * i] * RX[j] + T[3 * i + 1] * RX[3 + j] + T[3 * i + 2] * RX[6 + j];
    }
  const double L = Lss + Lsd;
  double *part = (double *)calloc((size_t)(nthreads > 0 ? nthreads : 1), sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double acc = 0.0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      const double *sw = SW + 3 * i, *g = Gt + 3 * i;
      const double gp0 = Lss + sw[0], gp1 = sw[1] - P0y, gp2 = sw[2] - P0z;
      const double c1 = gp1 * L / gp0, c2 = gp2 * L / gp0;
      const double alpha = atan(sqrt(gp1 * gp1 + gp2 * gp2) / gp0);
      const double ng = sqrt(gp0 * gp0 + gp1 * gp1 + gp2 * gp2);
      const double glen = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      const double beta = acos((gp0 / ng) * (g[0] / glen) + (gp1 / ng) * (g[1] / glen) + (gp2 / ng) * (g[2] / glen));
      const double theta = beta - pi / 2.0;
      const double nv = sqrt(g[1] * g[1] + g[2] * g[2]);
      const double nc = sqrt(L * L + c1 * c1 + c2 * c2);
      const double phix = acos((g[1] / nv) * (c1 / nc) + (g[2] / nv) * (c2 / nc));
      const double phiy = phix - 2.0 * theta;
      const double L2 = (Lsd - sw[0]) / cos(alpha);
      const double diffvec = L2 * sin(2.0 * theta) / sin(phiy);
      const double y22 = c1 + diffvec * g[1] / nv, z22 = c2 + diffvec * g[2] / nv;
      const double k0 = Lsd - sw[0], k1 = y22 - sw[1], k2 = z22 - sw[2];
      const double nk = sqrt(k0 * k0 + k1 * k1 + k2 * k2);
      const double K0 = k0 / nk, K1 = k1 / nk, K2 = k2 / nk;
      const double t = (R[0] * (Lsd - sw[0]) + R[3] * (dy00 - sw[1]) + R[6] * (dz00 - sw[2])) /
                       (R[0] * K0 + R[3] * K1 + R[6] * K2);
      const double w0 = t * K0 + (sw[0] - Lsd), w1 = t * K1 + (sw[1] - dy00), w2 = t * K2 + (sw[2] - dz00);
      const double Y = R[1] * w0 + R[4] * w1 + R[7] * w2, Z = R[2] * w0 + R[5] * w1 + R[8] * w2;
      const double ry = -Y / py + dety0, rz = -Z / pz + detz0;
      if (sim) {
        sim[2 * i] = ry;
        sim[2 * i + 1] = rz;
      }
      const double ey = mround(ry) - meas[2 * i], ez = mround(rz) - meas[2 * i + 1];
      acc += sqrt(ey * ey + ez * ez);
    }
    part[tid] = acc;
  }
  double total = 0.0;
  for (int k = 0; k < (nthreads > 0 ? nthreads : 1); k++) total += part[k];
  free(part);
  return n > 0 ? total / (double)n : NAN;
}
