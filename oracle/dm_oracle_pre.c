/*
 * CPU oracle, depth preprocessing: bilateral filter and Gaussian pyramid
 * (include/image/pyramid/pyramid.cu:200-282, paths relative to the reference
 * root).  TEST INFRASTRUCTURE ONLY: used by the tests as the checker, never
 * by the product.  The rules are those of the HIP kernels: symmetric
 * (2r+1)^2 window, weight exp(-dd^2/(2 sd^2)) * exp(-|q-p|^2/(2 ss^2))
 * (K: the reference's weight form; its 1/(2 pi s^2) factor cancels),
 * pixels outside the image or without depth skipped, a pixel without depth
 * stays 0 (D); pyramid: binomial 5x5 around (2u, 2v) over pixels within
 * 3 sigma_depth of the centre (D, KinectFusion's rule).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifndef OR_REAL
#define OR_REAL float
#endif
typedef OR_REAL real;
#define OR_EXPORT __attribute__((visibility("default")))

OR_EXPORT void or_bilateral(const float *in, float *out, int rows, int cols, int r, double sigma_d, double sigma_s) {
  const real a = (real)1 / ((real)2 * (real)sigma_d * (real)sigma_d);
  const real b = (real)1 / ((real)2 * (real)sigma_s * (real)sigma_s);
#pragma omp parallel for schedule(static)
  for (int v = 0; v < rows; ++v)
    for (int u = 0; u < cols; ++u) {
      const real dc = (real)in[(size_t)v * cols + u];
      real res = 0;
      if (dc > 0) {
        real sw = 0, sd = 0;
        for (int j = -r; j <= r; ++j)
          for (int i = -r; i <= r; ++i) {
            const int x = u + i, y = v + j;
            if (x < 0 || y < 0 || x >= cols || y >= rows) continue;
            const real d = (real)in[(size_t)y * cols + x];
            if (!(d > 0)) continue;
            const real dd = d - dc;
            const real w = (real)exp(-(double)((dd * dd) * a)) * (real)exp(-(double)((real)(i * i + j * j) * b));
            sw = sw + w;
            sd = sd + w * d;
          }
        res = sd / sw;
      }
      out[(size_t)v * cols + u] = (float)res;
    }
}

OR_EXPORT void or_pyr_down(const float *in, int rows, int cols, double sigma_d, float *out) {
  const int orows = rows >> 1, ocols = cols >> 1;
  const real thr = (real)3 * (real)sigma_d;
  static const real bw[5] = {1, 4, 6, 4, 1};
  for (int v = 0; v < orows; ++v)
    for (int u = 0; u < ocols; ++u) {
      const int cu = 2 * u, cv = 2 * v;
      const real dc = (real)in[(size_t)cv * cols + cu];
      real res = 0;
      if (dc > 0) {
        real sw = 0, sd = 0;
        for (int j = -2; j <= 2; ++j)
          for (int i = -2; i <= 2; ++i) {
            const int x = cu + i, y = cv + j;
            if (x < 0 || y < 0 || x >= cols || y >= rows) continue;
            const real d = (real)in[(size_t)y * cols + x];
            if (!(d > 0) || !((real)fabs((double)(d - dc)) < thr)) continue;
            const real w = bw[i + 2] * bw[j + 2];
            sw = sw + w;
            sd = sd + w * d;
          }
        res = sd / sw;
      }
      out[(size_t)v * ocols + u] = (float)res;
    }
}
