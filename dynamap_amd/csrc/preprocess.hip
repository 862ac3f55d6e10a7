// Depth preprocessing before tracking: bilateral filter and Gaussian
// pyramid (include/image/pyramid/pyramid.cu:200-282, bilateralFilterKernel /
// applyBilateralFilter; include/image/kernel2D/kernel2D.cu:14-48).
//
// Bilateral filter: out(p) = sum_q w(p,q) d(q) / sum_q w(p,q) over the
// (2r+1)^2 window, w = g(d(q) - d(p), sigma_depth) * g(|p - q|, sigma_space),
// g(x, s) = exp(-x^2 / (2 s^2)).  K: the weight form.  D: the window is
// symmetric (the reference loops over a 2r-wide window offset by -r), pixels
// outside the image and pixels without depth are skipped (the reference reads
// past the image and mixes zeros in), and a pixel without depth stays 0.
// The normalising constant 1/(2 pi s^2) of the reference's gaussian() cancels
// in the ratio and is left out.
//
// Pyramid level l+1 from level l: pixel (u, v) averages the 5x5
// neighbourhood of (2u, 2v) with binomial weights [1 4 6 4 1] x [1 4 6 4 1],
// over the pixels with depth within 3 sigma_depth of the centre pixel's
// (KinectFusion's rule); a centre without depth gives 0.  D: replaces the
// reference's fixed-offset nearest subsample into a 640x480 canvas
// (pyramid.cu:48-82).
//
// Workgroups of 16x16 pixels stage their tile plus the halo in LDS.
#include <algorithm>
#include "dm_device.h"
#include "dm_scan.h"
#include "dm_session.h"

namespace dm {

constexpr int kBilatMaxR = 7;
constexpr int kTile = 16;

__global__ __launch_bounds__(256) void k_bilateral(const float *__restrict__ in, float *__restrict__ out, int rows,
                                                  int cols, int r, float inv2sd2, float inv2ss2) {
  __shared__ float tile[(kTile + 2 * kBilatMaxR) * (kTile + 2 * kBilatMaxR)];
  const int W = kTile + 2 * r;
  const int u0 = blockIdx.x * kTile - r, v0 = blockIdx.y * kTile - r;
  for (int k = threadIdx.x; k < W * W; k += 256) {
    const int uu = u0 + k % W, vv = v0 + k / W;
    tile[k] = (uu >= 0 && vv >= 0 && uu < cols && vv < rows) ? in[(size_t)vv * cols + uu] : 0.f;
  }
  __syncthreads();
  const int tu = threadIdx.x & 15, tv = threadIdx.x >> 4;
  const int u = blockIdx.x * kTile + tu, v = blockIdx.y * kTile + tv;
  if (u >= cols || v >= rows) return;
  const float dc = tile[(tv + r) * W + tu + r];
  float res = 0.f;
  if (dc > 0.f) {
    float sw = 0.f, sd = 0.f;
    for (int j = -r; j <= r; ++j)
      for (int i = -r; i <= r; ++i) {
        const float d = tile[(tv + r + j) * W + tu + r + i];
        if (!(d > 0.f)) continue;  // also outside the image (staged as 0)
        const float dd = d - dc;
        const float w = expf(-(dd * dd) * inv2sd2) * expf(-(float)(i * i + j * j) * inv2ss2);
        sw = sw + w;
        sd = sd + w * d;
      }
    res = sd / sw;
  }
  out[(size_t)v * cols + u] = res;
}

__global__ __launch_bounds__(256) void k_pyr_down(const float *__restrict__ in, int rows, int cols,
                                                 float *__restrict__ out, int orows, int ocols, float thr) {
  const int u = blockIdx.x * kTile + (threadIdx.x & 15), v = blockIdx.y * kTile + (threadIdx.x >> 4);
  if (u >= ocols || v >= orows) return;
  const int cu = 2 * u, cv = 2 * v;
  const float dc = in[(size_t)cv * cols + cu];
  float res = 0.f;
  if (dc > 0.f) {
    const float b[5] = {1.f, 4.f, 6.f, 4.f, 1.f};
    float sw = 0.f, sd = 0.f;
    for (int j = -2; j <= 2; ++j)
      for (int i = -2; i <= 2; ++i) {
        const int x = cu + i, y = cv + j;
        if (x < 0 || y < 0 || x >= cols || y >= rows) continue;
        const float d = in[(size_t)y * cols + x];
        if (!(d > 0.f) || !(fabsf(d - dc) < thr)) continue;
        const float w = b[i + 2] * b[j + 2];
        sw = sw + w;
        sd = sd + w * d;
      }
    res = sd / sw;
  }
  out[(size_t)v * ocols + u] = res;
}

void launch_bilateral(const float *in, float *out, int rows, int cols, int r, float sigma_d, float sigma_s,
                      hipStream_t st) {
  const dim3 grid((unsigned)((cols + kTile - 1) / kTile), (unsigned)((rows + kTile - 1) / kTile));
  hipLaunchKernelGGL(k_bilateral, grid, dim3(256), 0, st, in, out, rows, cols, r, 1.0f / (2.0f * sigma_d * sigma_d),
                     1.0f / (2.0f * sigma_s * sigma_s));
}

}  // namespace dm

using namespace dm;

extern "C" {

int dm_depth_bilateral(const float *host_in, float *host_out, uint32_t rows, uint32_t cols, int32_t radius,
                       float sigma_depth, float sigma_space) {
  if (!host_in || !host_out || rows == 0 || cols == 0 || radius < 0 || radius > kBilatMaxR ||
      !(sigma_depth > 0.f) || !(sigma_space > 0.f)) {
    set_error("dm_depth_bilateral: invalid argument (0 <= radius <= 7, sigmas > 0)");
    return DM_E_INVALID;
  }
  const size_t P = (size_t)rows * cols;
  DevBuf din, dout;
  if (int rc = dev_alloc(din, P * 4)) return rc;
  if (int rc = dev_alloc(dout, P * 4)) return rc;
  DM_HIP(hipMemcpy(din.p, host_in, P * 4, hipMemcpyHostToDevice));
  launch_bilateral(din.as<float>(), dout.as<float>(), (int)rows, (int)cols, radius, sigma_depth, sigma_space, 0);
  DM_HIP(hipGetLastError());
  DM_HIP(hipMemcpy(host_out, dout.p, P * 4, hipMemcpyDeviceToHost));
  return DM_OK;
}

int dm_depth_pyramid(const float *host_in, uint32_t rows, uint32_t cols, float sigma_depth, int32_t levels,
                     float *host_out) {
  if (!host_in || !host_out || rows == 0 || cols == 0 || levels < 2 || levels > 8 || !(sigma_depth > 0.f)) {
    set_error("dm_depth_pyramid: invalid argument (2 <= levels <= 8, sigma_depth > 0)");
    return DM_E_INVALID;
  }
  size_t total = 0;
  for (int l = 0; l < levels; ++l) total += (size_t)(rows >> l) * (cols >> l);
  DevBuf d;
  if (int rc = dev_alloc(d, total * 4)) return rc;
  float *base = d.as<float>();
  DM_HIP(hipMemcpy(base, host_in, (size_t)rows * cols * 4, hipMemcpyHostToDevice));
  size_t off = 0;
  for (int l = 1; l < levels; ++l) {
    const int r0 = (int)(rows >> (l - 1)), c0 = (int)(cols >> (l - 1)), r1 = (int)(rows >> l), c1 = (int)(cols >> l);
    float *src = base + off;
    float *dst = src + (size_t)r0 * c0;
    if (r1 > 0 && c1 > 0) {
      const dim3 grid((unsigned)((c1 + kTile - 1) / kTile), (unsigned)((r1 + kTile - 1) / kTile));
      hipLaunchKernelGGL(k_pyr_down, grid, dim3(256), 0, 0, src, r0, c0, dst, r1, c1, 3.0f * sigma_depth);
    }
    off += (size_t)r0 * c0;
  }
  DM_HIP(hipGetLastError());
  DM_HIP(hipMemcpy(host_out, base, total * 4, hipMemcpyDeviceToHost));
  return DM_OK;
}

}  // extern "C"
