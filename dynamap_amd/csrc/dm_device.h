// Device-side math shared by the kernels.  The evaluation order of every
// expression follows the reference's float code (cited per helper); the
// library is compiled with -ffp-contract=off and correctly rounded division
// and square root, so FP32 results are reproducible on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dm_internal.h"

namespace dm {

// Load through a global-address-space pointer (a global_load instead of a
// flat one: pointers read from a DmView carry no address space).
template <class T> __device__ __forceinline__ T ldg(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// include/math/cuda_math.h:409-413
__device__ __forceinline__ float norm(f3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
// v / sqrt(dot): correctly rounded, no reciprocal square root
__device__ __forceinline__ f3 normalize(f3 v) {
  float n = sqrtf(dot(v, v));
  return mk(v.x / n, v.y / n, v.z / n);
}

// row-major 4x4 (float4x4 layout of include/math/utils_matrix.h:829-834)
struct M4 { float m[16]; };
__device__ __forceinline__ M4 load_m4(const float *p) {
  M4 r;
#pragma unroll
  for (int k = 0; k < 16; ++k) r.m[k] = p[k];
  return r;
}
// include/math/utils_matrix.h:923-930: ((m11 x + m12 y) + m13 z) + m14*1
__device__ __forceinline__ f3 xform(const M4 &T, f3 v) {
  return mk(T.m[0] * v.x + T.m[1] * v.y + T.m[2] * v.z + T.m[3],
            T.m[4] * v.x + T.m[5] * v.y + T.m[6] * v.z + T.m[7],
            T.m[8] * v.x + T.m[9] * v.y + T.m[10] * v.z + T.m[11]);
}
__device__ __forceinline__ f3 rot(const M4 &T, f3 v) {
  return mk(T.m[0] * v.x + T.m[1] * v.y + T.m[2] * v.z, T.m[4] * v.x + T.m[5] * v.y + T.m[6] * v.z,
            T.m[8] * v.x + T.m[9] * v.y + T.m[10] * v.z);
}

// float -> int32 with CUDA conversion semantics (truncate, saturate, NaN -> 0)
__device__ __forceinline__ int to_i32(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 0x7fffffff;
  if (x <= -2147483648.0f) return (int)0x80000000;
  return (int)x;
}

__device__ __forceinline__ int to_i32d(double x) {
  if (x != x) return 0;
  if (x >= 2147483648.0) return 0x7fffffff;
  if (x <= -2147483648.0) return (int)0x80000000;
  return (int)x;
}

struct Intr { int rows, cols; float fx, fy, cx, cy; };

// include/utils/utils.h:104-112 getPoint3d (scale 1)
__device__ __forceinline__ f3 get_point3d(int i, float depth, const Intr &s) {
  int v = i / s.cols;
  int u = i - s.cols * v;
  f3 p;
  p.z = depth;
  p.x = (((float)u - s.cx) * p.z) / s.fx;
  p.y = (((float)v - s.cy) * p.z) / s.fy;
  return p;
}
// getPoint3d of pixel (u, v) = (i % cols, i / cols), without the division
__device__ __forceinline__ f3 get_point3d_uv(int u, int v, float depth, const Intr &s) {
  f3 p;
  p.z = depth;
  p.x = (((float)u - s.cx) * p.z) / s.fx;
  p.y = (((float)v - s.cy) * p.z) / s.fy;
  return p;
}
// include/utils/utils.h:114-119 Project
__device__ __forceinline__ void project(f3 p, const Intr &s, int &u, int &v) {
  u = to_i32(roundf((s.fx * p.x) / p.z + s.cx));
  v = to_i32(roundf((s.fy * p.y) / p.z + s.cy));
}

// include/dataStructure/hashTable/hashTable.cu:152-163 (bucket index, i.e.
// the reference's value divided by bucket_size); int32 products wrap.
// For a power-of-two bucket count the C remainder plus the negative fix-up
// equals the low bits of the two's-complement hash, so no division is needed.
__device__ __forceinline__ int hash_bucket(int x, int y, int z, int num_buckets) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
  if ((num_buckets & (num_buckets - 1)) == 0) return (int)(h & (uint32_t)(num_buckets - 1));
  int r = (int)h % num_buckets;
  if (r < 0) r += num_buckets;
  return r;
}
__device__ __forceinline__ bool in_key_range(int x, int y, int z) {
  return x >= -kKeyRange && x < kKeyRange && y >= -kKeyRange && y < kKeyRange && z >= -kKeyRange &&
         z < kKeyRange;
}
// occupied flag in bit 63, 21 bits per coordinate
__device__ __forceinline__ uint64_t pack_key(int x, int y, int z) {
  const uint64_t M = (1u << 21) - 1u;
  return (1ull << 63) | (((uint64_t)(uint32_t)x & M) << 42) | (((uint64_t)(uint32_t)y & M) << 21) |
         ((uint64_t)(uint32_t)z & M);
}
__device__ __forceinline__ int sext21(uint64_t v) { return (int)((int32_t)((uint32_t)(v & 0x1fffff) << 11) >> 11); }
__device__ __forceinline__ void unpack_key(uint64_t k, int &x, int &y, int &z) {
  x = sext21(k >> 42); y = sext21(k >> 21); z = sext21(k);
}

// bucket record accessors
__device__ __forceinline__ const uint8_t *bucket_ptr(const DmView &v, int b) {
  return v.buckets + (size_t)b * (size_t)v.bstride;
}
__device__ __forceinline__ int bucket_count(const DmView &v, const uint8_t *bk) {
  return ldg((const int32_t *)(bk + 12 * v.bucket_size));
}

// FindHashEntry (hashTable.cu:136-150): pool index or -1.  Only the occupied
// prefix of the bucket is scanned.
__device__ __forceinline__ int find_block(const DmView &v, int x, int y, int z) {
  if (!in_key_range(x, y, z)) return -1;
  const uint64_t key = pack_key(x, y, z);
  const uint8_t *bk = bucket_ptr(v, hash_bucket(x, y, z, v.num_buckets));
  const int cnt = bucket_count(v, bk);
  const uint64_t *keys = (const uint64_t *)bk;
  for (int s = 0; s < cnt; ++s)
    if (ldg(keys + s) == key) return ldg((const int32_t *)(bk + 8 * v.bucket_size) + s);
  return -1;
}

// world -> global voxel (tsdf.cu:189-193): trunc(x/vs + sign(x)*0.5)
__device__ __forceinline__ float signf_(float x) { return (float)((x > 0) - (x < 0)); }
__device__ __forceinline__ int world_to_voxel(float x, float vs) { return to_i32(x / vs + signf_(x) * 0.5f); }
// floor(g / 8) and g mod 8 (in 0..7) of a two's-complement int
__device__ __forceinline__ int floordiv8(int g) { return g >> 3; }
__device__ __forceinline__ int mod8(int g) { return g & 7; }

// voxel word: low 32 bits sdf, then r, g, b, weight bytes
__device__ __forceinline__ float vox_sdf(uint64_t w) { return __uint_as_float((uint32_t)w); }
__device__ __forceinline__ uint32_t vox_rgbw(uint64_t w) { return (uint32_t)(w >> 32); }

// Small per-thread cache of the last block looked up.
struct BlockCache {
  int bx, by, bz, ptr;
  __device__ void init() { bx = 0x7fffffff; by = 0; bz = 0; ptr = -1; }
};

// Voxel word of global voxel (gx, gy, gz), 0 when its block is unallocated.
__device__ __forceinline__ uint64_t voxel_at(const DmView &v, int gx, int gy, int gz, BlockCache &c) {
  const int bx = floordiv8(gx), by = floordiv8(gy), bz = floordiv8(gz);
  if (bx != c.bx || by != c.by || bz != c.bz) {
    c.bx = bx; c.by = by; c.bz = bz;
    c.ptr = find_block(v, bx, by, bz);
  }
  if (c.ptr < 0) return 0ull;
  return ldg(v.vox + ((size_t)c.ptr * kBlockVox + mod8(gx) * 64 + mod8(gy) * 8 + mod8(gz)));
}

// world_to_voxel(x, vs) without a division in the common case: x * (1/vs)
// is within a few ulp of x / vs, so trunc(q + sign(x) * 0.5) is the same
// integer unless q + sign(x) * 0.5 lies within ~8 ulp of an integer; then
// (and for huge values) the exact expression is evaluated.  copysignf(0.5f,
// x) stands in for signf_(x) * 0.5f: they differ only at x = +-0, where both
// truncate to voxel 0.
__device__ __forceinline__ int world_to_voxel_fast(float x, float vs, float inv_vs) {
  const float t = x * inv_vs + copysignf(0.5f, x);
  const float tr = truncf(t);
  const float fr = fabsf(t - tr);
  const float tol = fabsf(t) * 1.0e-6f;
  if (!(fabsf(t) < 1.0e8f) || fr <= tol || fr >= 1.0f - tol) return world_to_voxel(x, vs);
  return (int)tr;
}

__device__ __forceinline__ uint32_t to_u8(float x) {
  if (!(x > 0.f)) return 0u;
  if (x >= 255.f) return 255u;
  return (uint32_t)x;
}

// ---- lane sums through DPP row permutes
// x moved across lanes by a DPP pattern
template <int kCtrl> __device__ __forceinline__ float dpp_f32(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xf, 0xf, false));
}
template <int kCtrl> __device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), kCtrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
// Sum over each row of 16 lanes, the same fixed tree in every lane: lane
// pairs (l, l ^ 1), (l, l ^ 2) (quad permutes), then the half-row and row
// mirrors (lane l with 7 - l and 15 - l within its 8 and 16 lanes); each
// step adds two equal-shaped partial sums, so every lane of the row ends
// with the same value.  Every lane of the row must be active.
template <class T> __device__ __forceinline__ T row_sum(T x) {
  if constexpr (sizeof(T) == 8) {
    x += dpp_f64<0xb1>(x);   // quad_perm [1, 0, 3, 2]
    x += dpp_f64<0x4e>(x);   // quad_perm [2, 3, 0, 1]
    x += dpp_f64<0x141>(x);  // row_half_mirror
    x += dpp_f64<0x140>(x);  // row_mirror
  } else {
    x += dpp_f32<0xb1>(x);
    x += dpp_f32<0x4e>(x);
    x += dpp_f32<0x141>(x);
    x += dpp_f32<0x140>(x);
  }
  return x;
}
// Sum over the wavefront (all 64 lanes active): the row sums, then lanes
// l ^ 16 and l ^ 32; every lane ends with the same value.
template <class T> __device__ __forceinline__ T wave_sum(T x) {
  x = row_sum(x);
  x += __shfl_xor(x, 16, 64);
  x += __shfl_xor(x, 32, 64);
  return x;
}

}  // namespace dm
