// Live-frame maps and projective frame-to-model ICP.
//
// Reference: include/image/sensor/rgbdImage.cu (creatPointCloudKernel :22-38,
// vertextoNormal :126-145), include/image/pyramid/pyramid.cu (downsampling
// :48-82), include/solver/gnSolver/gnSolver.cu (buildNLSKernel :25-196,
// solve :317-438).
//
// Per iteration the reduction kernel evaluates one Gauss-Newton linearisation
// per pixel and reduces the 30 sums (21 of J J^T, 6 of J r, counts) with a
// wavefront shuffle tree plus an LDS stage into one record per workgroup; the
// solve kernel adds the records in FP64 in a fixed order, solves the damped
// 6x6 system with an FP64 LDL^T and composes the SE(3) increment on the
// device, so a frame's iterations run back to back without a host round trip.
#include "dm_device.h"
#include "dm_solve.h"

namespace dm {

__device__ __forceinline__ Intr level_intr(const DmView &v, int l) {
  const float s = (float)(1 << l);
  return Intr{v.rows >> l, v.cols >> l, v.fx / s, v.fy / s, v.cx / s, v.cy / s};
}

__device__ __forceinline__ float intensity_rgb(const uint8_t *rgb, int idx) {
  if (!rgb) return 0.f;
  return (0.2126f * (float)rgb[3 * idx] + 0.7152f * (float)rgb[3 * idx + 1] + 0.0722f * (float)rgb[3 * idx + 2]) *
         (1.0f / 255.0f);
}

// depth of level-l pixel j, taken from the level-0 frame at (2^l u, 2^l v)
__device__ __forceinline__ float level_depth(const DmView &v, int l, int j, int cols_l) {
  const int vv = j / cols_l, uu = j - vv * cols_l;
  const float *d = v.depth_track ? v.depth_track : v.depth;
  return d[(size_t)(vv << l) * v.cols + (uu << l)];
}

// Vertex of level-l pixel j.  D: evaluated in FP64 and rounded to float, so
// the live-frame maps equal the FP64 oracle's (depth <= 0 gives (0,0,0)).
__device__ __forceinline__ f3 level_vertex(const DmView &v, int l, int j, const Intr &s) {
  const float d = level_depth(v, l, j, s.cols);
  if (!(d > 0.0f)) return mk(0.f, 0.f, 0.f);
  const int vv = j / s.cols, uu = j - vv * s.cols;
  const double dd = (double)d;
  return mk((float)((((double)uu - (double)s.cx) * dd) / (double)s.fx),
            (float)((((double)vv - (double)s.cy) * dd) / (double)s.fy), d);
}

// Source maps of all pyramid levels straight from the level-0 depth:
// vertex (creatPointCloudKernel), intensity, and the normal of
// vertextoNormal (K: cross(V[i+2]-V[i-2], V[i+W]-V[i-W]) with out-of-range
// neighbours replaced by the pixel itself; D: NaN when a neighbour has no
// depth, evaluated in FP64).  D: level l is the true (rows>>l) x (cols>>l)
// nearest subsample.
__global__ __launch_bounds__(256) void k_source_maps(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int P0 = v.rows * v.cols;
  const int P1 = (v.rows >> 1) * (v.cols >> 1);
  const int P2 = (v.rows >> 2) * (v.cols >> 2);
  int t = blockIdx.x * 256 + threadIdx.x;
  int l;
  if (t < P0) l = 0;
  else if (t < P0 + P1) { l = 1; t -= P0; }
  else if (t < P0 + P1 + P2) { l = 2; t -= P0 + P1; }
  else return;
  const Intr s = level_intr(v, l);
  const int size = s.rows * s.cols;
  const f3 c = level_vertex(v, l, t, s);
  const int vv = t / s.cols, uu = t - vv * s.cols;
  const float inten = intensity_rgb(v.rgb, (vv << l) * v.cols + (uu << l));
  v.svert[l][t] = make_float4(c.x, c.y, c.z, inten);
  const int li = t - 2 < 0 ? t : t - 2;
  const int ri = t + 2 >= size ? t : t + 2;
  const int ui = t - s.cols <= 0 ? t : t - s.cols;
  const int di = t + s.cols >= size ? t : t + s.cols;
  const f3 L = level_vertex(v, l, li, s), R = level_vertex(v, l, ri, s), U = level_vertex(v, l, ui, s),
           D = level_vertex(v, l, di, s);
  float4 n;
  if (L.z == 0.f || R.z == 0.f || U.z == 0.f || D.z == 0.f || c.z == 0.f) {
    const float qn = __int_as_float(0x7fc00000);
    n = make_float4(qn, qn, qn, 0.f);
  } else {
    const double ax = (double)R.x - (double)L.x, ay = (double)R.y - (double)L.y, az = (double)R.z - (double)L.z;
    const double bx = (double)D.x - (double)U.x, by = (double)D.y - (double)U.y, bz = (double)D.z - (double)U.z;
    const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const double nn = sqrt(cx * cx + cy * cy + cz * cz);
    n = make_float4((float)(cx / nn), (float)(cy / nn), (float)(cz / nn), 0.f);
  }
  v.snorm[l][t] = n;
}

__device__ __forceinline__ bool bad3(float4 n) {
  return !(fabsf(n.x) <= 3.4e38f) || !(fabsf(n.y) <= 3.4e38f) || !(fabsf(n.z) <= 3.4e38f) ||
         (n.x == 0.f && n.y == 0.f && n.z == 0.f);
}

constexpr int kVals = 30;

// Association and gates of one pixel (buildNLSKernel, gnSolver.cu:55-195).
// K: Project() association, the |cos| > cos(pi/10) normal gate, r =
// (p - q).n, J = [p x n; n].  D: source normal rotated by Tinc, pixel bounds
// check, |p - q| <= dist_max, depth gate [icp_dmin, icp_dmax].  The
// transform, projection and gates are in FP64 (the pixel choice and the
// inclusion decisions follow the FP64 geometry); residual and Jacobian are
// FP32.
struct IcpAssoc {
  f3 p, pq, n;    // transformed source point, p - q, model normal
  float qw, vw;   // model and source intensities
  int pu, pv, j;  // associated model pixel
};
__device__ __forceinline__ bool icp_assoc(const DmView &v, int level, const Intr &s, int i, IcpAssoc &a) {
  const float4 V = v.svert[level][i];
  const float4 N = v.snorm[level][i];
  const float4 *tv = v.tvert[level];
  const float4 *tn = v.tnorm[level];
  bool ok = V.z >= v.icp_dmin && V.z <= v.icp_dmax && !bad3(N);
  const double *Td = v.posed + PD_TINC;
  double px = 0.0, py = 0.0, pz = 0.0;
  int pu = -1, pv = -1, j = 0;
  float4 Q = make_float4(0.f, 0.f, 0.f, 0.f), NT = Q;
  if (ok) {
    const double X = V.x, Y = V.y, Z = V.z;
    px = Td[0] * X + Td[1] * Y + Td[2] * Z + Td[3];
    py = Td[4] * X + Td[5] * Y + Td[6] * Z + Td[7];
    pz = Td[8] * X + Td[9] * Y + Td[10] * Z + Td[11];
    pu = to_i32d(round(((double)s.fx * px) / pz + (double)s.cx));
    pv = to_i32d(round(((double)s.fy * py) / pz + (double)s.cy));
    ok = !(pu < 0 || pv < 0 || pu >= s.cols || pv >= s.rows);
  }
  if (ok) {
    j = pv * s.cols + pu;
    Q = tv[j];
    NT = tn[j];
    ok = Q.z >= v.icp_dmin && Q.z <= v.icp_dmax && !bad3(NT);
  }
  if (ok) {
    const double X = N.x, Y = N.y, Z = N.z;
    const double nrx = Td[0] * X + Td[1] * Y + Td[2] * Z;
    const double nry = Td[4] * X + Td[5] * Y + Td[6] * Z;
    const double nrz = Td[8] * X + Td[9] * Y + Td[10] * Z;
    const double ct = (double)NT.x * nrx + (double)NT.y * nry + (double)NT.z * nrz;
    const double dqx = px - (double)Q.x, dqy = py - (double)Q.y, dqz = pz - (double)Q.z;
    ok = fabs(ct) > 0.95105651629515357 && sqrt(dqx * dqx + dqy * dqy + dqz * dqz) <= (double)v.dist_max;
    a.p = mk((float)px, (float)py, (float)pz);
    a.pq = mk((float)dqx, (float)dqy, (float)dqz);
  }
  a.n = mk(NT.x, NT.y, NT.z);
  a.qw = Q.w;
  a.vw = V.w;
  a.pu = pu;
  a.pv = pv;
  a.j = j;
  return ok;
}

// Per-pixel terms of one Gauss-Newton linearisation at pyramid level `level`
// accumulated into vals (upper JtJ 21, Jtr 6, geometric count, r^2,
// photometric count); the photometric term adds w_photo Jp Jp^T and
// w_photo Jp rp.
__device__ __forceinline__ void icp_pixel(const DmView &v, int level, const Intr &s, int i, float *vals) {
  IcpAssoc as;
  if (!icp_assoc(v, level, s, i, as)) return;
  const float4 *tv = v.tvert[level];
  const f3 p = as.p, n = as.n;
  float pvl[kVals];
  const float r = dot(as.pq, n);
  const f3 c = cross(p, n);
  const float J[6] = {c.x, c.y, c.z, n.x, n.y, n.z};
  int k = 0;
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) pvl[k++] = J[a] * J[b];
#pragma unroll
  for (int a = 0; a < 6; ++a) pvl[21 + a] = J[a] * r;
  pvl[27] = 1.f;
  pvl[28] = r * r;
  pvl[29] = 0.f;
  if (v.w_photo > 0.f && as.pu >= 1 && as.pv >= 1 && as.pu < s.cols - 1 && as.pv < s.rows - 1) {
    const int j = as.j;
    const float gx = (tv[j + 1].w - tv[j - 1].w) * 0.5f;
    const float gy = (tv[j + s.cols].w - tv[j - s.cols].w) * 0.5f;
    const float rp = as.qw - as.vw;
    const float iz = 1.0f / p.z;
    const f3 G = mk(gx * s.fx * iz, gy * s.fy * iz, -(gx * s.fx * p.x + gy * s.fy * p.y) * iz * iz);
    const f3 cg = cross(p, G);
    const float Jp[6] = {cg.x, cg.y, cg.z, G.x, G.y, G.z};
    const float wp = v.w_photo;
    int kk = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) { pvl[kk] = pvl[kk] + wp * Jp[a] * Jp[b]; ++kk; }
#pragma unroll
    for (int a = 0; a < 6; ++a) pvl[21 + a] = pvl[21 + a] + wp * Jp[a] * rp;
    pvl[29] = 1.f;
  }
#pragma unroll
  for (int q = 0; q < kVals; ++q) vals[q] = vals[q] + pvl[q];
}

// The feature row of one pixel for the matrix-core reduction: geometric
// [J, r, 1] in columns 0-7 and photometric sqrt(w_photo) [Jp, rp] with a 1 in
// columns 8-15 (zeros where a term does not apply).  F^T F summed over the
// pixels holds J^T J and J^T r of both terms, r^2 and both counts.
__device__ __forceinline__ void icp_features(const DmView &v, int level, const Intr &s, int i, float swp,
                                             float *f) {
#pragma unroll
  for (int k = 0; k < 16; ++k) f[k] = 0.f;
  IcpAssoc as;
  if (!icp_assoc(v, level, s, i, as)) return;
  const float4 *tv = v.tvert[level];
  const f3 p = as.p, n = as.n;
  const float r = dot(as.pq, n);
  const f3 c = cross(p, n);
  f[0] = c.x; f[1] = c.y; f[2] = c.z; f[3] = n.x; f[4] = n.y; f[5] = n.z; f[6] = r; f[7] = 1.f;
  if (v.w_photo > 0.f && as.pu >= 1 && as.pv >= 1 && as.pu < s.cols - 1 && as.pv < s.rows - 1) {
    const int j = as.j;
    const float gx = (tv[j + 1].w - tv[j - 1].w) * 0.5f;
    const float gy = (tv[j + s.cols].w - tv[j - s.cols].w) * 0.5f;
    const float rp = as.qw - as.vw;
    const float iz = 1.0f / p.z;
    const f3 G = mk(gx * s.fx * iz, gy * s.fy * iz, -(gx * s.fx * p.x + gy * s.fy * p.y) * iz * iz);
    const f3 cg = cross(p, G);
    f[8] = swp * cg.x; f[9] = swp * cg.y; f[10] = swp * cg.z;
    f[11] = swp * G.x; f[12] = swp * G.y; f[13] = swp * G.z;
    f[14] = swp * rp; f[15] = 1.f;
  }
}

// The damped Gauss-Newton step from the summed records: Tinc <- exp(xi)
// Tinc.  On the session's last iteration of the frame apply the reference's
// guard (gnSolver.cu:426-428) and set the frame pose T = Tprev * Tinc.
__device__ void icp_step_device(const DmView &v, int level, int it, const double *sums) {
  double A[36], b[6], xi[6], E[16];
  sym6_from_upper(sums, A);
  for (int k = 0; k < 6; ++k) b[k] = sums[21 + k];
  icp_step(A, b, it, (double)v.lambda, xi);
  se3_exp(xi, E);
  double *Tinc = v.posed + PD_TINC;
  mat4_mul(E, Tinc, Tinc);
  for (int k = 0; k < 6; ++k) v.posed[PD_ACC + k] += xi[k];
  // gnSolver.cu:413-415: the level ends once the summed twist changed by at
  // most min_increment since the last step that did not end a level
  const bool stop = level_converged(v.posed + PD_ACC, v.posed + PD_PREVACC, (double)v.min_inc);
  if (stop) v.ctr[C_ICP_STOP] = level + 1;
  bool last = it == v.iters[level] || stop;
  for (int l = 0; l < level; ++l) last = last && v.iters[l] == 0;
  if (last) {
    const double *acc = v.posed + PD_ACC;
    const double ssum = acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5];
    if (ssum != ssum || isinf(ssum) || ssum >= (double)v.max_inc) {
      for (int k = 0; k < 16; ++k) Tinc[k] = (k % 5 == 0) ? 1.0 : 0.0;
      atomicAdd(&v.ctr[C_ICP_FAIL], 1);
    }
    double *T = v.posed + PD_T;
    mat4_mul(v.posed + PD_TPREV, Tinc, T);
    for (int k = 0; k < 16; ++k) v.posef[PF_T + k] = (float)T[k];
    rigid_inverse_ff(v.posef + PF_T, v.posef + PF_TINV);
  }
  for (int k = 0; k < 16; ++k) v.posef[PF_TINC + k] = (float)Tinc[k];
}

// One linearisation: every thread accumulates 4 pixels (strided by 256 for
// coalescing) in FP32; a wavefront tree (wave_sum: DPP row permutes, then
// two shuffles) and an LDS stage reduce them to one record per workgroup.
__global__ __launch_bounds__(256) void k_icp_reduce(const DmView *__restrict__ views, const int *__restrict__ active,
                                                     int level, int it) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (v.ctr[C_HAS_MODEL] == 0 || it > v.iters[level] || v.ctr[C_ICP_STOP] == level + 1) return;
  const Intr s = level_intr(v, level);
  const int P = s.rows * s.cols;
  const int nwg = icp_wg_count(P);
  if ((int)blockIdx.x >= nwg) return;
  float vals[kVals];
#pragma unroll
  for (int k = 0; k < kVals; ++k) vals[k] = 0.f;
#pragma unroll
  for (int q = 0; q < kIcpPix / 256; ++q) {
    const int i = blockIdx.x * kIcpPix + q * 256 + threadIdx.x;
    if (i < P) icp_pixel(v, level, s, i, vals);
  }
  // wavefront tree (wave_sum) then LDS across the four waves
#pragma unroll
  for (int k = 0; k < kVals; ++k) vals[k] = wave_sum(vals[k]);
  __shared__ float red[4][kVals];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kVals; ++k) red[wave][k] = vals[k];
  }
  __syncthreads();
  if (threadIdx.x < kVals) {
    const int k = threadIdx.x;
    v.partials[(size_t)blockIdx.x * kIcpVals + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  }
}

// The same linearisation on the matrix cores: each pixel's 16-float feature
// row (icp_features) is staged through LDS and v_mfma_f32_16x16x4f32
// accumulates F^T F over 4 rows per instruction (lane l supplies feature l %
// 16 of row l / 16 as both operands; the 16x16 result holds column l % 16,
// rows 4 (l / 16) .. + 3 in lane l).  The four waves' tiles are added in a
// fixed order and the record takes J^T J = G + P (the geometric and
// photometric 6x6 blocks), J^T r, the counts and r^2 from the tile.  Same
// grid and record layout as k_icp_reduce; the sums differ from it by
// rounding only (w_photo enters as sqrt(w_photo)^2).
typedef float icp_f4 __attribute__((ext_vector_type(4)));
constexpr int kFeatPitch = 17;  // LDS row pitch (floats): rows of a wave fall on distinct banks
__global__ __launch_bounds__(256) void k_icp_reduce_mfma(const DmView *__restrict__ views,
                                                          const int *__restrict__ active, int level, int it) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (v.ctr[C_HAS_MODEL] == 0 || it > v.iters[level] || v.ctr[C_ICP_STOP] == level + 1) return;
  const Intr s = level_intr(v, level);
  const int P = s.rows * s.cols;
  const int nwg = icp_wg_count(P);
  if ((int)blockIdx.x >= nwg) return;
  __shared__ float feat[4][64 * kFeatPitch];
  __shared__ float tile[4][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float *F = feat[wave];
  const float swp = v.w_photo > 0.f ? sqrtf(v.w_photo) : 0.f;
  icp_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int q = 0; q < kIcpPix / 256; ++q) {
    const int i = blockIdx.x * kIcpPix + q * 256 + threadIdx.x;
    float f[16];
    if (i < P) {
      icp_features(v, level, s, i, swp, f);
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) f[k] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) F[lane * kFeatPitch + k] = f[k];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const float a = F[(4 * m + (lane >> 4)) * kFeatPitch + (lane & 15)];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, a, acc, 0, 0, 0);
    }
    __syncthreads();  // the rows are rewritten by the next group of pixels
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) tile[wave][(4 * (lane >> 4) + e) * 16 + (lane & 15)] = acc[e];
  __syncthreads();
  if (threadIdx.x < kVals) {
    const int k = threadIdx.x;
    auto T = [&](int r, int c) {
      const int o = r * 16 + c;
      return (tile[0][o] + tile[1][o]) + (tile[2][o] + tile[3][o]);
    };
    float val;
    if (k < 21) {
      int a = 0, kk = k;
      while (kk >= 6 - a) { kk -= 6 - a; ++a; }
      const int b = a + kk;
      val = T(a, b) + T(8 + a, 8 + b);
    } else if (k < 27) {
      val = T(k - 21, 6) + T(8 + (k - 21), 14);
    } else if (k == 27) {
      val = T(7, 7);
    } else if (k == 28) {
      val = T(6, 6);
    } else {
      val = T(15, 15);
    }
    v.partials[(size_t)blockIdx.x * kIcpVals + k] = val;
  }
}

// Sum the records in FP64 in a fixed order (thread t takes records t, t+256,
// ...; wave_sum's tree per wavefront; (w0 + w1) + (w2 + w3); host replica:
// icp_sum_records) and take the Gauss-Newton step (one workgroup per session).
__global__ __launch_bounds__(256) void k_icp_solve(const DmView *__restrict__ views, const int *__restrict__ active,
                                                    int level, int it) {
  const int slot_ = active[blockIdx.x];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (v.ctr[C_HAS_MODEL] == 0 || it > v.iters[level] || v.ctr[C_ICP_STOP] == level + 1) return;
  const int nwg = icp_wg_count((v.rows >> level) * (v.cols >> level));
  double acc[kVals];
#pragma unroll
  for (int k = 0; k < kVals; ++k) acc[k] = 0.0;
  for (int w = threadIdx.x; w < nwg; w += 256) {
    const float *rec = v.partials + (size_t)w * kIcpVals;
#pragma unroll
    for (int k = 0; k < kVals; ++k) acc[k] += (double)rec[k];
  }
#pragma unroll
  for (int k = 0; k < kVals; ++k) acc[k] = wave_sum(acc[k]);
  __shared__ double dred[4][kVals];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kVals; ++k) dred[wave][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double sums[kVals];
  for (int k = 0; k < kVals; ++k) sums[k] = (dred[0][k] + dred[1][k]) + (dred[2][k] + dred[3][k]);
  icp_step_device(v, level, it, sums);
}

// Start of a frame: Tprev <- T when a model exists, otherwise T <- Tinit;
// Tinc <- I, summed twist <- 0.
__global__ void k_pose_begin(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.x];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (threadIdx.x != 0) return;
  double *P = v.posed;
  if (v.ctr[C_HAS_MODEL]) {
    for (int k = 0; k < 16; ++k) P[PD_TPREV + k] = P[PD_T + k];
  } else {
    for (int k = 0; k < 16; ++k) P[PD_T + k] = P[PD_TINIT + k];
  }
  for (int k = 0; k < 16; ++k) {
    v.posef[PF_T + k] = (float)P[PD_T + k];
    v.posef[PF_TPREV + k] = (float)P[PD_TPREV + k];
    P[PD_TINC + k] = (k % 5 == 0) ? 1.0 : 0.0;
    v.posef[PF_TINC + k] = (k % 5 == 0) ? 1.f : 0.f;
  }
  rigid_inverse_ff(v.posef + PF_T, v.posef + PF_TINV);
  for (int k = 0; k < 6; ++k) P[PD_ACC + k] = 0.0;
  for (int k = 0; k < 6; ++k) P[PD_PREVACC + k] = 0.0;
  v.ctr[C_ICP_STOP] = 0;
  v.ctr[C_NVIS] = 0;
  v.ctr[C_NVISIT] = 0;
}

void launch_source_maps(const Launch &L) {
  if (L.nact <= 0) return;
  const int tot = L.max_pixels + L.max_pixels / 4 + L.max_pixels / 16 + 2;
  hipLaunchKernelGGL(k_source_maps, dim3((unsigned)((tot + 255) / 256), (unsigned)L.nact), dim3(256), 0, L.stream,
                     L.views, L.active);
}

void launch_icp_reduce(const Launch &L, int level, int it, bool solve) {
  if (L.nact <= 0) return;
  const int P = L.max_pixels >> (2 * level);
  if (L.icp_mfma)
    hipLaunchKernelGGL(k_icp_reduce_mfma, dim3((unsigned)icp_wg_count(P), (unsigned)L.nact), dim3(256), 0, L.stream,
                       L.views, L.active, level, it);
  else
    hipLaunchKernelGGL(k_icp_reduce, dim3((unsigned)icp_wg_count(P), (unsigned)L.nact), dim3(256), 0, L.stream,
                       L.views, L.active, level, it);
  if (solve)
    hipLaunchKernelGGL(k_icp_solve, dim3((unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active, level, it);
}

void icp_sum_records(const float *part, int nwg, double *sums) {
  for (int k = 0; k < kVals; ++k) {
    double lane[256];
    for (int t = 0; t < 256; ++t) {
      double a = 0.0;
      for (int w = t; w < nwg; w += 256) a += (double)part[(size_t)w * kIcpVals + k];
      lane[t] = a;
    }
    // wave_sum's steps: lane i adds lane src(i) for src = i ^ 1, i ^ 2, the
    // half-row mirror, the row mirror, i ^ 16, i ^ 32
    auto src = [](int step, int i) {
      switch (step) {
        case 0: return i ^ 1;
        case 1: return i ^ 2;
        case 2: return (i & ~7) | (7 - (i & 7));
        case 3: return (i & ~15) | (15 - (i & 15));
        case 4: return i ^ 16;
        default: return i ^ 32;
      }
    };
    double wv[4];
    for (int q = 0; q < 4; ++q) {
      double x[64];
      for (int i = 0; i < 64; ++i) x[i] = lane[64 * q + i];
      for (int step = 0; step < 6; ++step) {
        double y[64];
        for (int i = 0; i < 64; ++i) y[i] = x[i] + x[src(step, i)];
        for (int i = 0; i < 64; ++i) x[i] = y[i];
      }
      wv[q] = x[0];
    }
    sums[k] = (wv[0] + wv[1]) + (wv[2] + wv[3]);
  }
}

void launch_pose_begin(const Launch &L) {
  if (L.nact <= 0) return;
  hipLaunchKernelGGL(k_pose_begin, dim3((unsigned)L.nact), dim3(64), 0, L.stream, L.views, L.active);
}

// Copy the frame pose of every session of a batch into a log (dest[y] is the
// log row of list entry y).
__global__ void k_pose_log(const DmView *__restrict__ views, const int *__restrict__ active,
                           const int *__restrict__ dest, double *__restrict__ log) {
  const int slot_ = active[blockIdx.x];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (threadIdx.x < 16) log[(size_t)dest[blockIdx.x] * 16 + threadIdx.x] = v.posed[PD_T + threadIdx.x];
}

void launch_pose_log(const Launch &L, const int *dest, double *log) {
  if (L.nact <= 0) return;
  hipLaunchKernelGGL(k_pose_log, dim3((unsigned)L.nact), dim3(64), 0, L.stream, L.views, L.active, dest, log);
}


}  // namespace dm
