/*
 * CPU oracle for the DynaMap dense-mapping path.  TEST INFRASTRUCTURE ONLY:
 * nothing in the product library links, loads or calls this file; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark do, as the
 * checker.
 *
 * It is a plain-C restatement of the reference algorithm, written from the
 * reference sources read as text (file:line cited per function, paths relative
 * to the reference repository root).  It is compiled twice:
 *   -DOR_REAL=float   the FP32 transcription: the same IEEE operations in the
 *                     same order as the reference (and the HIP kernels), no
 *                     FMA contraction, correctly rounded / and sqrt.  Integer
 *                     outputs (allocated block set, hash occupancy, cube
 *                     indices) are defined by this build.
 *   -DOR_REAL=double  the FP64 oracle for real-valued outputs (TSDF, depth,
 *                     poses, vertices).
 *
 * Where the build deliberately diverges from the reference, the divergence is
 * marked "D:" with its reason; "K:" marks a reference quirk that is kept
 * because it defines the reference semantics.
 *
 * Pinning (see tests/): hash vs an independent Python formula and a
 * hand-computed table; DDA vs brute-force segment sampling; integration vs the
 * closed form of a fronto-parallel wall; trilinear sampling exact on a linear
 * field; raycast vs analytic ray-plane / ray-sphere depth; normal map vs a
 * plane's normal and a cubic surface's closed form; LDLT vs numpy.linalg.solve;
 * SE(3) exp vs scipy expm; ICP normal equations (geometric and photometric) vs
 * an FP64 per-pixel sum with finite-difference Jacobians, known-pose
 * convergence, photometric recovery of in-plane motion; MC tables vs a digest
 * of the reference's tables, MC vs sphere topology.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#ifndef OR_REAL
#define OR_REAL float
#endif
typedef OR_REAL real;

#define OR_EXPORT __attribute__((visibility("default")))

OR_EXPORT int or_real_size(void) { return (int)sizeof(real); }

/* OpenMP threads of the oracle's parallel loops (CPU-baseline timing). */
#include <omp.h>
OR_EXPORT void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
OR_EXPORT int or_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------ */
/* small vector helpers (evaluation order spelled out)                 */
/* ------------------------------------------------------------------ */
typedef struct { real x, y, z; } r3;
static inline r3 mk3(real x, real y, real z) { r3 v = {x, y, z}; return v; }
static inline r3 add3(r3 a, r3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline r3 sub3(r3 a, r3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline r3 scl3(r3 a, real s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline real dot3(r3 a, r3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline r3 cross3(r3 a, r3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* include/math/cuda_math.h:409-413: sqrtf(x*x + y*y + z*z) */
static inline real norm3(r3 v) { return (real)sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
/* D: normalise as v / sqrt(dot) (correctly rounded), not with rsqrtf. */
static inline r3 normalize3(r3 v) {
  real n = (real)sqrt(dot3(v, v));
  return mk3(v.x / n, v.y / n, v.z / n);
}
/* float4x4 * float3, include/math/utils_matrix.h:923-930: ((m11 x + m12 y) + m13 z) + m14 */
static inline r3 xform(const real *m, r3 v) {
  return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3],
             m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
             m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]);
}
static inline r3 rot3(const real *m, r3 v) {
  return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
             m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
static inline int isbad(real x) { return x != x || x == (real)INFINITY || x == -(real)INFINITY; }

/* CUDA float->int conversion semantics: truncate, saturate, NaN -> 0. */
static inline int32_t to_i32(real x) {
  if (x != x) return 0;
  if (x >= (real)2147483648.0) return INT32_MAX;
  if (x <= (real)-2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
static inline int32_t to_i32_d(double x) {
  if (x != x) return 0;
  if (x >= 2147483648.0) return INT32_MAX;
  if (x <= -2147483648.0) return INT32_MIN;
  return (int32_t)x;
}
static inline real rround(real x) { return sizeof(real) == 4 ? (real)roundf((float)x) : (real)round((double)x); }
static inline real rfloor(real x) { return sizeof(real) == 4 ? (real)floorf((float)x) : (real)floor((double)x); }

/* rigid inverse of a row-major 4x4, computed in FP64 then rounded to real */
static void rigid_inverse(const real *T, real *Ti) {
  double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
  double t[3] = {T[3], T[7], T[11]};
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) Ti[4 * r + c] = (real)R[3 * c + r];
    Ti[4 * r + 3] = (real)(-(R[0 * 3 + r] * t[0] + R[1 * 3 + r] * t[1] + R[2 * 3 + r] * t[2]));
  }
  Ti[12] = 0; Ti[13] = 0; Ti[14] = 0; Ti[15] = 1;
}

/* ------------------------------------------------------------------ */
/* sensor                                                               */
/* ------------------------------------------------------------------ */
typedef struct { int rows, cols; real fx, fy, cx, cy; } or_sensor;

/* include/utils/utils.h:104-112 getPoint3d (scale 1) */
static inline r3 get_point3d(int i, real depth, const or_sensor *s) {
  int v = i / s->cols;
  int u = i - s->cols * v;
  r3 p;
  p.z = depth;
  p.x = (((real)u - s->cx) * p.z) / s->fx;
  p.y = (((real)v - s->cy) * p.z) / s->fy;
  return p;
}
/* include/utils/utils.h:114-119 Project: round((f x)/z + c) */
static inline void project(r3 p, const or_sensor *s, int *u, int *v) {
  *u = to_i32(rround((s->fx * p.x) / p.z + s->cx));
  *v = to_i32(rround((s->fy * p.y) / p.z + s->cy));
}

/* ------------------------------------------------------------------ */
/* voxel-block hash table                                              */
/* ------------------------------------------------------------------ */
#define OR_KEY_RANGE (1 << 20)
typedef struct {
  real voxel_size, trunc, max_depth, min_depth;
  int num_buckets, bucket_size, num_blocks, max_weight;
  uint64_t *keys;      /* [num_buckets * bucket_size], 0 = free */
  int32_t *ptrs;       /* pool index per slot */
  int32_t *count;      /* occupied slots per bucket (always a prefix) */
  int32_t *block_key;  /* [num_blocks * 3] */
  real *sdf;           /* [num_blocks * 512] */
  uint8_t *rgb;        /* [num_blocks * 512 * 3] */
  uint8_t *w;          /* [num_blocks * 512] */
  uint8_t *live;       /* [num_blocks] 1 = pool index in use */
  int32_t *free_list;  /* [num_blocks] freed pool indices (a stack) */
  int num_alloc;       /* pool high-water mark */
  int num_free;
  int64_t dropped_bucket, dropped_pool, dropped_range;
} or_volume;

/* include/dataStructure/hashTable/hashTable.cu:152-163.  The int32 products
 * wrap (two's complement), the C remainder truncates toward zero. */
OR_EXPORT int or_hash(int x, int y, int z, int num_buckets, int bucket_size) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
  int32_t r = (int32_t)h % num_buckets;
  if (r < 0) r += num_buckets;
  return r * bucket_size;
}

OR_EXPORT uint64_t or_pack_key(int x, int y, int z) {
  const uint64_t M = (1u << 21) - 1u;
  return (1ULL << 63) | (((uint64_t)(uint32_t)x & M) << 42) | (((uint64_t)(uint32_t)y & M) << 21) |
         ((uint64_t)(uint32_t)z & M);
}
static inline int in_key_range(int x, int y, int z) {
  return x >= -OR_KEY_RANGE && x < OR_KEY_RANGE && y >= -OR_KEY_RANGE && y < OR_KEY_RANGE &&
         z >= -OR_KEY_RANGE && z < OR_KEY_RANGE;
}

OR_EXPORT or_volume *or_volume_create(double voxel_size, int num_buckets, int bucket_size, int num_blocks,
                                      int max_weight, double trunc, double max_depth, double min_depth) {
  or_volume *v = (or_volume *)calloc(1, sizeof(or_volume));
  v->voxel_size = (real)voxel_size; v->trunc = (real)trunc;
  v->max_depth = (real)max_depth; v->min_depth = (real)min_depth;
  v->num_buckets = num_buckets; v->bucket_size = bucket_size;
  v->num_blocks = num_blocks; v->max_weight = max_weight;
  size_t E = (size_t)num_buckets * bucket_size;
  v->keys = (uint64_t *)calloc(E, sizeof(uint64_t));
  v->ptrs = (int32_t *)malloc(E * sizeof(int32_t));
  for (size_t i = 0; i < E; ++i) v->ptrs[i] = -1;
  v->count = (int32_t *)calloc((size_t)num_buckets, sizeof(int32_t));
  v->block_key = (int32_t *)calloc((size_t)num_blocks * 3, sizeof(int32_t));
  v->sdf = (real *)calloc((size_t)num_blocks * 512, sizeof(real));
  v->rgb = (uint8_t *)calloc((size_t)num_blocks * 512 * 3, 1);
  v->w = (uint8_t *)calloc((size_t)num_blocks * 512, 1);
  v->live = (uint8_t *)calloc((size_t)num_blocks, 1);
  v->free_list = (int32_t *)malloc((size_t)num_blocks * sizeof(int32_t));
  return v;
}

OR_EXPORT void or_volume_destroy(or_volume *v) {
  if (!v) return;
  free(v->keys); free(v->ptrs); free(v->count); free(v->block_key);
  free(v->sdf); free(v->rgb); free(v->w); free(v->live); free(v->free_list); free(v);
}

/* live blocks */
OR_EXPORT int or_volume_num_alloc(const or_volume *v) { return v->num_alloc - v->num_free; }
OR_EXPORT void or_volume_dropped(const or_volume *v, int64_t out[3]) {
  out[0] = v->dropped_bucket; out[1] = v->dropped_pool; out[2] = v->dropped_range;
}

/* hashTable.cu:136-150 FindHashEntry; returns pool index or -1 */
static int find_block(const or_volume *v, int x, int y, int z) {
  if (!in_key_range(x, y, z)) return -1;
  uint64_t k = or_pack_key(x, y, z);
  int base = or_hash(x, y, z, v->num_buckets, v->bucket_size);
  int b = base / v->bucket_size;
  for (int s = 0; s < v->count[b]; ++s)
    if (v->keys[base + s] == k) return v->ptrs[base + s];
  return -1;
}

/* ------------------------------------------------------------------ */
/* world <-> voxel maps, include/geometry/tsdf/tsdf.cu:189-239          */
/* ------------------------------------------------------------------ */
static inline real signf_(real x) { return (real)((x > 0) - (x < 0)); }
static inline int world_to_voxel(real x, real vs) { return to_i32(x / vs + signf_(x) * (real)0.5); }
static inline int floordiv8(int g) { return g < 0 ? (g - 8 + 1) / 8 : g / 8; }
static inline int mod8(int g) { int m = g % 8; return m < 0 ? m + 8 : m; }

typedef struct { real sdf; uint8_t r, g, b, w; } or_voxel;

static or_voxel get_voxel(const or_volume *v, r3 p) {
  int gx = world_to_voxel(p.x, v->voxel_size), gy = world_to_voxel(p.y, v->voxel_size),
      gz = world_to_voxel(p.z, v->voxel_size);
  or_voxel out = {0, 0, 0, 0, 0};
  int ptr = find_block(v, floordiv8(gx), floordiv8(gy), floordiv8(gz));
  if (ptr < 0) return out;
  size_t idx = (size_t)ptr * 512 + mod8(gx) * 64 + mod8(gy) * 8 + mod8(gz);
  out.sdf = v->sdf[idx];
  out.r = v->rgb[3 * idx]; out.g = v->rgb[3 * idx + 1]; out.b = v->rgb[3 * idx + 2];
  out.w = v->w[idx];
  return out;
}

/* ------------------------------------------------------------------ */
/* block allocation, tsdf.cu:345-453 (ray DDA) + hashTable.cu:80-110    */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t *k; int32_t *xyz; size_t n, cap; } keyset;
static void ks_push(keyset *s, int x, int y, int z) {
  if (s->n == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 1024;
    s->xyz = (int32_t *)realloc(s->xyz, s->cap * 3 * sizeof(int32_t));
  }
  s->xyz[3 * s->n] = x; s->xyz[3 * s->n + 1] = y; s->xyz[3 * s->n + 2] = z;
  s->n++;
}

/*
 * Visit the blocks of one pixel's ray exactly as AllocateFromDepthKernel
 * does (tsdf.cu:357-452).  K: the ray starts at the camera centre, the
 * boundary uses (b + step) * bs for both directions, and the "neg_ray" extra
 * block is visited.  D: the traversal loop is bounded by the L1 distance from
 * the (shifted) start block to the end block, the number of steps a
 * traversal that only moves towards the end needs; the reference loop has no
 * bound and would not terminate if a step overshot.
 * Returns the number of blocks visited (written to out_xyz when non-NULL).
 */
static int ray_blocks(const real *T, int i, real d, const or_sensor *s, real trunc, real block_m,
                      keyset *out) {
  r3 start = mk3(T[3], T[7], T[11]);
  r3 point = xform(T, get_point3d(i, d, s));
  if (point.x == 0 && point.y == 0 && point.z == 0) return 0;
  r3 dir = normalize3(sub3(point, start));
  real sd = norm3(sub3(start, point));
  r3 end = add3(start, scl3(dir, sd + trunc));
  int bx = to_i32(rfloor(start.x / block_m)), by = to_i32(rfloor(start.y / block_m)),
      bz = to_i32(rfloor(start.z / block_m));
  int ex = to_i32(rfloor(end.x / block_m)), ey = to_i32(rfloor(end.y / block_m)),
      ez = to_i32(rfloor(end.z / block_m));
  int stx = (dir.x > 0) - (dir.x < 0), sty = (dir.y > 0) - (dir.y < 0), stz = (dir.z > 0) - (dir.z < 0);
  real dtx = dir.x != 0 ? (real)fabs((double)(block_m / dir.x)) : (real)FLT_MAX;
  real dty = dir.y != 0 ? (real)fabs((double)(block_m / dir.y)) : (real)FLT_MAX;
  real dtz = dir.z != 0 ? (real)fabs((double)(block_m / dir.z)) : (real)FLT_MAX;
  real bdx = ((real)bx + (real)stx) * block_m, bdy = ((real)by + (real)sty) * block_m,
       bdz = ((real)bz + (real)stz) * block_m;
  real mtx = dir.x != 0 ? (bdx - start.x) / dir.x : (real)FLT_MAX;
  real mty = dir.y != 0 ? (bdy - start.y) / dir.y : (real)FLT_MAX;
  real mtz = dir.z != 0 ? (bdz - start.z) / dir.z : (real)FLT_MAX;
  int n = 0;
  int dx = 0, dy = 0, dz = 0, neg = 0;
  if (bx != ex && dir.x < 0) { dx--; neg = 1; }
  if (by != ey && dir.y < 0) { dy--; neg = 1; }
  if (bz != ez && dir.z < 0) { dz--; neg = 1; }
  if (out) ks_push(out, bx, by, bz);
  n++;
  if (neg) {
    bx += dx; by += dy; bz += dz;
    if (out) ks_push(out, bx, by, bz);
    n++;
  }
  int64_t bound = llabs((int64_t)ex - bx) + llabs((int64_t)ey - by) + llabs((int64_t)ez - bz);
  for (int64_t it = 0; it < bound && (bx != ex || by != ey || bz != ez); ++it) {
    if (mtx < mty) {
      if (mtx < mtz) { bx += stx; mtx += dtx; }
      else { bz += stz; mtz += dtz; }
    } else {
      if (mty < mtz) { by += sty; mty += dty; }
      else { bz += stz; mtz += dtz; }
    }
    if (out) ks_push(out, bx, by, bz);
    n++;
  }
  return n;
}

/* Blocks visited by one pixel's ray (test hook). */
OR_EXPORT int or_ray_blocks(const double Tin[16], int i, double d, int rows, int cols, double fx, double fy,
                            double cx, double cy, double trunc, double block_m, int32_t *out_xyz, int cap) {
  real T[16];
  for (int k = 0; k < 16; ++k) T[k] = (real)Tin[k];
  or_sensor s = {rows, cols, (real)fx, (real)fy, (real)cx, (real)cy};
  keyset ks = {0, 0, 0, 0};
  int n = ray_blocks(T, i, (real)d, &s, (real)trunc, (real)block_m, &ks);
  for (int k = 0; k < n && k < cap; ++k) {
    out_xyz[3 * k] = ks.xyz[3 * k]; out_xyz[3 * k + 1] = ks.xyz[3 * k + 1]; out_xyz[3 * k + 2] = ks.xyz[3 * k + 2];
  }
  free(ks.xyz);
  return n;
}

typedef struct { uint64_t key; int32_t bucket, x, y, z; } newkey;
static int cmp_newkey(const void *a, const void *b) {
  const newkey *p = (const newkey *)a, *q = (const newkey *)b;
  if (p->bucket != q->bucket) return p->bucket < q->bucket ? -1 : 1;
  if (p->key != q->key) return p->key < q->key ? -1 : 1;
  return 0;
}

/*
 * One frame of allocation with the canonical insertion rule.
 * K (hashTable.cu:80-110): a block whose bucket has no free slot is dropped;
 * there is no overflow chain.  D: the winners are canonical instead of
 * thread-order dependent: the new keys of a bucket are inserted in ascending
 * packed-key order into the free slots, the rest are dropped.  If the pool
 * cannot hold every insertion of the frame, no block is inserted in that
 * frame (counted in dropped_pool).  Pool indices are assigned in
 * (bucket, slot) order here; they are internal and not compared.
 */
OR_EXPORT void or_allocate(or_volume *v, const float *depth, int rows, int cols, double fx, double fy,
                           double cx, double cy, const double Tin[16]) {
  real T[16];
  for (int k = 0; k < 16; ++k) T[k] = (real)Tin[k];
  or_sensor s = {rows, cols, (real)fx, (real)fy, (real)cx, (real)cy};
  real block_m = (real)8 * v->voxel_size; /* tsdf.cu:353-354: blockSize * voxelSize */
  keyset all = {0, 0, 0, 0};
  for (int i = 0; i < rows * cols; ++i) {
    real d = (real)depth[i];
    if (d < v->min_depth || d > v->max_depth) continue;
    ray_blocks(T, i, d, &s, v->trunc, block_m, &all);
  }
  /* unique new keys */
  newkey *nk = (newkey *)malloc((all.n ? all.n : 1) * sizeof(newkey));
  size_t m = 0;
  for (size_t j = 0; j < all.n; ++j) {
    int x = all.xyz[3 * j], y = all.xyz[3 * j + 1], z = all.xyz[3 * j + 2];
    if (!in_key_range(x, y, z)) { v->dropped_range++; continue; }
    if (find_block(v, x, y, z) >= 0) continue;
    nk[m].key = or_pack_key(x, y, z);
    nk[m].bucket = or_hash(x, y, z, v->num_buckets, v->bucket_size) / v->bucket_size;
    nk[m].x = x; nk[m].y = y; nk[m].z = z;
    ++m;
  }
  free(all.xyz);
  qsort(nk, m, sizeof(newkey), cmp_newkey);
  size_t u = 0;
  for (size_t j = 0; j < m; ++j)
    if (u == 0 || nk[j].key != nk[u - 1].key) nk[u++] = nk[j];
  /* per-bucket decision */
  size_t n_ins = 0;
  for (size_t j = 0; j < u;) {
    size_t e = j;
    while (e < u && nk[e].bucket == nk[j].bucket) ++e;
    int freeb = v->bucket_size - v->count[nk[j].bucket];
    size_t take = (size_t)freeb < e - j ? (size_t)freeb : e - j;
    n_ins += take;
    j = e;
  }
  /* freed pool indices are reused first (Heap::Consume, heap.cu:14-17) */
  const size_t grow = n_ins > (size_t)v->num_free ? n_ins - (size_t)v->num_free : 0;
  if ((size_t)v->num_alloc + grow > (size_t)v->num_blocks) {
    v->dropped_pool += (int64_t)u;
    free(nk);
    return;
  }
  for (size_t j = 0; j < u;) {
    size_t e = j;
    int b = nk[j].bucket;
    while (e < u && nk[e].bucket == b) ++e;
    for (size_t q = j; q < e; ++q) {
      if (v->count[b] >= v->bucket_size) { v->dropped_bucket++; continue; }
      int slot = b * v->bucket_size + v->count[b];
      int ptr = v->num_free > 0 ? v->free_list[--v->num_free] : v->num_alloc++;
      v->live[ptr] = 1;
      v->keys[slot] = nk[q].key;
      v->ptrs[slot] = ptr;
      v->block_key[3 * ptr] = nk[q].x; v->block_key[3 * ptr + 1] = nk[q].y; v->block_key[3 * ptr + 2] = nk[q].z;
      v->count[b]++;
    }
    j = e;
  }
  free(nk);
}

/* ------------------------------------------------------------------ */
/* integration, tsdf.cu:267-344 + geom_Types.h:79-102                   */
/* ------------------------------------------------------------------ */
/*
 * K: per-block cull by the corner's camera z and the projection of
 * (corner_cam + half a block edge) (computed in double, as the reference's
 * 0.5 literal makes it); the per-voxel band rule; the merge order (sdf and
 * colour with the old weight, then the weight).  D: only allocated blocks are
 * visited (the reference also walks free entries whose position is (0,0,0)),
 * and each voxel is addressed directly instead of through the world->hash
 * round trip, which maps back to the same voxel.
 */
static void integrate_block(or_volume *v, int ptr, const float *depth, const uint8_t *rgb, const or_sensor *s,
                            const real *Tinv) {
  const int B = 8;
  real vs = v->voxel_size;
  int kx = v->block_key[3 * ptr], ky = v->block_key[3 * ptr + 1], kz = v->block_key[3 * ptr + 2];
  r3 pos = mk3((real)kx * vs * (real)B, (real)ky * vs * (real)B, (real)kz * vs * (real)B);
  r3 pc = xform(Tinv, pos);
  if (pc.z < 0) return;
  double half = 0.5 * (double)vs * (double)B;
  r3 bc = mk3((real)((double)pc.x + half), (real)((double)pc.y + half), (real)((double)pc.z + half));
  int iu, iv;
  project(bc, s, &iu, &iv);
  if (!(iu >= 0 && iv >= 0 && iu < s->cols && iv < s->rows)) return;
  for (int bx = 0; bx < B; ++bx)
    for (int by = 0; by < B; ++by)
      for (int bz = 0; bz < B; ++bz) {
        r3 vp = mk3(pos.x + (real)bx * vs, pos.y + (real)by * vs, pos.z + (real)bz * vs);
        r3 vc = xform(Tinv, vp);
        project(vc, s, &iu, &iv);
        if (!(iu >= 0 && iv >= 0 && iu < s->cols && iv < s->rows)) continue;
        int idx = iv * s->cols + iu;
        real d = (real)depth[idx];
        if (d < v->min_depth) continue;
        if (d > v->max_depth) continue;
        r3 p3 = get_point3d(idx, d, s);
        real surf = norm3(sub3(mk3(0, 0, 0), p3));
        real vox = norm3(sub3(mk3(0, 0, 0), vc));
        real nsdf;
        if (vox > surf - v->trunc && vox < surf + v->trunc && d < v->max_depth) nsdf = surf - vox;
        else if (vox < surf - v->trunc) nsdf = v->trunc;
        else continue;
        size_t vi = (size_t)ptr * 512 + bx * 64 + by * 8 + bz;
        uint8_t nr = rgb ? rgb[3 * idx] : 0, ng = rgb ? rgb[3 * idx + 1] : 0, nb = rgb ? rgb[3 * idx + 2] : 0;
        real w = (real)v->w[vi];
        v->sdf[vi] = (v->sdf[vi] * w + nsdf * (real)1) / (w + (real)1);
        real cr = ((real)v->rgb[3 * vi] * w + (real)nr * (real)1) / (w + (real)1);
        real cg = ((real)v->rgb[3 * vi + 1] * w + (real)ng * (real)1) / (w + (real)1);
        real cb = ((real)v->rgb[3 * vi + 2] * w + (real)nb * (real)1) / (w + (real)1);
        v->rgb[3 * vi] = (uint8_t)cr; v->rgb[3 * vi + 1] = (uint8_t)cg; v->rgb[3 * vi + 2] = (uint8_t)cb;
        int nw = v->w[vi] + 1;
        if (nw > v->max_weight) nw = v->max_weight;
        v->w[vi] = (uint8_t)nw;
      }
}

OR_EXPORT void or_integrate(or_volume *v, const float *depth, const uint8_t *rgb, int rows, int cols, double fx,
                            double fy, double cx, double cy, const double Tin[16]) {
  real T[16], Ti[16];
  for (int k = 0; k < 16; ++k) T[k] = (real)Tin[k];
  rigid_inverse(T, Ti);
  or_sensor s = {rows, cols, (real)fx, (real)fy, (real)cx, (real)cy};
#pragma omp parallel for schedule(dynamic, 64)
  for (int p = 0; p < v->num_alloc; ++p)
    if (v->live[p]) integrate_block(v, p, depth, rgb, &s, Ti);
}

/* IntegrateScan = allocation then integration, tsdf.cu:454-469 */
OR_EXPORT void or_integrate_scan(or_volume *v, const float *depth, const uint8_t *rgb, int rows, int cols,
                                 double fx, double fy, double cx, double cy, const double T[16]) {
  or_allocate(v, depth, rows, cols, fx, fy, cx, cy, T);
  or_integrate(v, depth, rgb, rows, cols, fx, fy, cx, cy, T);
}

/* Voxel of a block coordinate and local index (test hook). */
OR_EXPORT int or_get_block(const or_volume *v, int x, int y, int z, double *sdf, uint8_t *rgb, uint8_t *w) {
  int ptr = find_block(v, x, y, z);
  if (ptr < 0) return -1;
  for (int i = 0; i < 512; ++i) {
    size_t vi = (size_t)ptr * 512 + i;
    sdf[i] = (double)v->sdf[vi];
    rgb[3 * i] = v->rgb[3 * vi]; rgb[3 * i + 1] = v->rgb[3 * vi + 1]; rgb[3 * i + 2] = v->rgb[3 * vi + 2];
    w[i] = v->w[vi];
  }
  return ptr;
}

/* Hash-table contents in slot order: keys (0 = free). */
OR_EXPORT void or_get_table(const or_volume *v, uint64_t *keys, int32_t *counts) {
  memcpy(keys, v->keys, (size_t)v->num_buckets * v->bucket_size * sizeof(uint64_t));
  memcpy(counts, v->count, (size_t)v->num_buckets * sizeof(int32_t));
}

/* All allocated (live) blocks: coordinates, sdf, rgb, weight, in pool order. */
OR_EXPORT void or_dump_blocks(const or_volume *v, int32_t *xyz, double *sdf, uint8_t *rgb, uint8_t *w) {
  int o = 0;
  for (int p = 0; p < v->num_alloc; ++p) {
    if (!v->live[p]) continue;
    xyz[3 * o] = v->block_key[3 * p]; xyz[3 * o + 1] = v->block_key[3 * p + 1]; xyz[3 * o + 2] = v->block_key[3 * p + 2];
    for (int i = 0; i < 512; ++i) {
      size_t vi = (size_t)o * 512 + i, pi = (size_t)p * 512 + i;
      if (sdf) sdf[vi] = (double)v->sdf[pi];
      if (rgb) { rgb[3 * vi] = v->rgb[3 * pi]; rgb[3 * vi + 1] = v->rgb[3 * pi + 1]; rgb[3 * vi + 2] = v->rgb[3 * pi + 2]; }
      if (w) w[vi] = v->w[pi];
    }
    ++o;
  }
}

/*
 * HashTable::DeleteBlock (hashTable.cu:112-134) for n blocks (absent or
 * repeated ones skipped): the block's voxels are reset and its pool index
 * pushed on the free stack (Heap::Append, heap.cu:19-22).  D: the remaining
 * entries of the bucket close up in their order, so the occupied slots stay
 * a prefix.  Returns the number deleted.
 */
OR_EXPORT int64_t or_delete_blocks(or_volume *v, const int32_t *xyz, int64_t n) {
  int64_t deleted = 0;
  for (int64_t j = 0; j < n; ++j) {
    const int x = xyz[3 * j], y = xyz[3 * j + 1], z = xyz[3 * j + 2];
    if (!in_key_range(x, y, z)) continue;
    const uint64_t k = or_pack_key(x, y, z);
    const int base = or_hash(x, y, z, v->num_buckets, v->bucket_size);
    const int b = base / v->bucket_size;
    for (int s = 0; s < v->count[b]; ++s) {
      if (v->keys[base + s] != k) continue;
      const int ptr = v->ptrs[base + s];
      for (int t = s; t + 1 < v->count[b]; ++t) {
        v->keys[base + t] = v->keys[base + t + 1];
        v->ptrs[base + t] = v->ptrs[base + t + 1];
      }
      v->count[b]--;
      v->keys[base + v->count[b]] = 0;
      v->ptrs[base + v->count[b]] = -1;
      memset(v->sdf + (size_t)ptr * 512, 0, 512 * sizeof(real));
      memset(v->rgb + (size_t)ptr * 512 * 3, 0, 512 * 3);
      memset(v->w + (size_t)ptr * 512, 0, 512);
      v->live[ptr] = 0;
      v->free_list[v->num_free++] = ptr;
      ++deleted;
      break;
    }
  }
  return deleted;
}

/* Pool recycling: delete every live block whose centre ((x + 1/2) * 8 *
 * voxel_size per axis, in double) lies farther than radius from center. */
OR_EXPORT int64_t or_recycle(or_volume *v, const double center[3], double radius) {
  int32_t *sel = (int32_t *)malloc((size_t)(v->num_alloc > 0 ? v->num_alloc : 1) * 3 * sizeof(int32_t));
  int64_t m = 0;
  const double bm = 8.0 * (double)v->voxel_size, r2 = radius * radius;
  for (int p = 0; p < v->num_alloc; ++p) {
    if (!v->live[p]) continue;
    const int32_t *k = v->block_key + 3 * p;
    const double dx = ((double)k[0] + 0.5) * bm - center[0], dy = ((double)k[1] + 0.5) * bm - center[1],
                 dz = ((double)k[2] + 0.5) * bm - center[2];
    if (dx * dx + dy * dy + dz * dz > r2) { memcpy(sel + 3 * m, k, 3 * sizeof(int32_t)); ++m; }
  }
  const int64_t d = or_delete_blocks(v, sel, m);
  free(sel);
  return d;
}

/* ------------------------------------------------------------------ */
/* trilinear sampling, tsdf.cu:47-175 GetInterpolatedVoxel             */
/* ------------------------------------------------------------------ */
/* K: a corner with weight 0 contributes the centre voxel's values; an
 * unobserved centre voxel makes the whole sample unobserved. */
static int interp_voxel(const or_volume *v, r3 p, real *sdf_out, real col[3]) {
  or_voxel v0 = get_voxel(v, p);
  if (v0.w == 0) return 0;
  real vs = v->voxel_size;
  r3 dual = sub3(p, mk3(vs / (real)2, vs / (real)2, vs / (real)2));
  r3 vp = mk3(p.x / vs, p.y / vs, p.z / vs);
  r3 wt = mk3(vp.x - rfloor(vp.x), vp.y - rfloor(vp.y), vp.z - rfloor(vp.z));
  static const int cx_[8] = {0, 1, 0, 0, 1, 0, 1, 1};
  static const int cy_[8] = {0, 0, 1, 0, 1, 1, 0, 1};
  static const int cz_[8] = {0, 0, 0, 1, 0, 1, 1, 1};
  real dist = 0, c0 = 0, c1 = 0, c2 = 0;
  for (int c = 0; c < 8; ++c) {
    r3 q = add3(dual, mk3(cx_[c] ? vs : (real)0, cy_[c] ? vs : (real)0, cz_[c] ? vs : (real)0));
    or_voxel vv = get_voxel(v, q);
    real a = cx_[c] ? wt.x : (real)1 - wt.x;
    real b = cy_[c] ? wt.y : (real)1 - wt.y;
    real g = cz_[c] ? wt.z : (real)1 - wt.z;
    real coef = a * b * g;
    const or_voxel *src = vv.w == 0 ? &v0 : &vv;
    dist += coef * src->sdf;
    c0 += coef * (real)src->r;
    c1 += coef * (real)src->g;
    c2 += coef * (real)src->b;
  }
  *sdf_out = dist;
  if (col) { col[0] = c0; col[1] = c1; col[2] = c2; }
  return 1;
}

/* Test hook: interp_voxel at n world points (x, y, z triples); ok[k] = 0
 * where the centre voxel is unobserved (sdf and colour then left at 0). */
OR_EXPORT void or_interp(const or_volume *v, int n, const double *pts, double *sdf, double *col, int *ok) {
  for (int k = 0; k < n; ++k) {
    real s = 0, c[3] = {0, 0, 0};
    ok[k] = interp_voxel(v, mk3((real)pts[3 * k], (real)pts[3 * k + 1], (real)pts[3 * k + 2]), &s, c);
    sdf[k] = (double)s;
    for (int a = 0; a < 3; ++a) col[3 * k + a] = (double)c[a];
  }
}

static inline uint8_t to_u8(real x) {
  if (!(x > 0)) return 0;
  if (x >= (real)255) return 255;
  return (uint8_t)x;
}

/* ------------------------------------------------------------------ */
/* raycast, tsdf.cu:470-526 GenerateDepth/GenerateRgb                  */
/* ------------------------------------------------------------------ */
/*
 * Per pixel, march the z-depth parameter from min_depth towards max_depth.
 * K: samples are GetInterpolatedVoxel at T * getPoint3d(i, z).
 * D (raycast of these workloads): the step is 0.5*trunc through unobserved space and
 * min(sdf, 0.5*trunc) (at least half a voxel) through observed space; a hit
 * is the first +/- zero crossing, located by linear interpolation between the
 * bracketing samples (the reference stops at sdf < voxel_size, up to one
 * voxel short of the surface).  A negative first observed sample (entering
 * from behind a surface) ends the ray as a miss.  Normals are the central
 * difference gradient of the interpolated SDF (step one voxel, world axes),
 * rotated into the camera frame; they are NaN if any of the six samples is
 * unobserved.
 */
OR_EXPORT void or_raycast(const or_volume *v, int rows, int cols, double fx, double fy, double cx, double cy,
                          const double Tin[16], float *depth_out, float *vert_out, float *norm_out,
                          uint8_t *rgb_out) {
  real T[16], Ti[16];
  for (int k = 0; k < 16; ++k) T[k] = (real)Tin[k];
  rigid_inverse(T, Ti);
  or_sensor s = {rows, cols, (real)fx, (real)fy, (real)cx, (real)cy};
  const real vs = v->voxel_size, tau = v->trunc;
  const real big = (real)0.5 * tau, small = (real)0.5 * vs;
  const int max_steps = (int)ceil((double)((v->max_depth - v->min_depth) / small)) + 2;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < rows * cols; ++i) {
    real z = v->min_depth, pz = 0, ps = 0;
    int pobs = 0, hit = 0;
    real zh = 0;
    for (int it = 0; it < max_steps && z < v->max_depth; ++it) {
      real sdf;
      r3 p = xform(T, get_point3d(i, z, &s));
      if (!interp_voxel(v, p, &sdf, NULL)) { pobs = 0; z = z + big; continue; }
      if (sdf <= 0) {
        if (pobs && ps > 0) { zh = pz + (z - pz) * (ps / (ps - sdf)); hit = 1; }
        break;
      }
      pobs = 1; ps = sdf; pz = z;
      real st = sdf < big ? sdf : big;
      if (st < small) st = small;
      z = z + st;
    }
    float *vo = vert_out + 3 * (size_t)i, *no = norm_out + 3 * (size_t)i;
    if (!hit) {
      depth_out[i] = 0.0f;
      vo[0] = vo[1] = vo[2] = 0.0f;
      no[0] = no[1] = no[2] = NAN;
      if (rgb_out) rgb_out[3 * i] = rgb_out[3 * i + 1] = rgb_out[3 * i + 2] = 0;
      continue;
    }
    r3 pc = get_point3d(i, zh, &s);
    r3 pw = xform(T, pc);
    depth_out[i] = (float)zh;
    vo[0] = (float)pc.x; vo[1] = (float)pc.y; vo[2] = (float)pc.z;
    real sp[3], sm[3], col[3];
    int ok = 1;
    for (int a = 0; a < 3; ++a) {
      r3 e = mk3(a == 0 ? vs : (real)0, a == 1 ? vs : (real)0, a == 2 ? vs : (real)0);
      ok &= interp_voxel(v, add3(pw, e), &sp[a], NULL);
      ok &= interp_voxel(v, sub3(pw, e), &sm[a], NULL);
    }
    if (ok) {
      r3 g = mk3(sp[0] - sm[0], sp[1] - sm[1], sp[2] - sm[2]);
      r3 n = rot3(Ti, normalize3(g));
      no[0] = (float)n.x; no[1] = (float)n.y; no[2] = (float)n.z;
    } else {
      no[0] = no[1] = no[2] = NAN;
    }
    if (rgb_out) {
      real sdf;
      if (interp_voxel(v, pw, &sdf, col)) {
        rgb_out[3 * i] = to_u8(col[0]); rgb_out[3 * i + 1] = to_u8(col[1]); rgb_out[3 * i + 2] = to_u8(col[2]);
      } else {
        rgb_out[3 * i] = rgb_out[3 * i + 1] = rgb_out[3 * i + 2] = 0;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* vertex / normal maps and pyramid                                     */
/* ------------------------------------------------------------------ */
/* include/image/sensor/rgbdImage.cu:22-38 (D: depth <= 0 gives (0,0,0)).
 * D: evaluated in FP64 and rounded to float (both builds), so the live-frame
 * maps are identical to the FP64 oracle's. */
OR_EXPORT void or_vertex_map(const float *depth, int rows, int cols, double fx, double fy, double cx, double cy,
                             float *vert) {
  for (int i = 0; i < rows * cols; ++i) {
    float *o = vert + 3 * (size_t)i;
    if (!(depth[i] > 0.0f)) { o[0] = o[1] = o[2] = 0.0f; continue; }
    int v = i / cols, u = i - cols * v;
    double d = (double)depth[i];
    o[0] = (float)((((double)u - cx) * d) / fx);
    o[1] = (float)((((double)v - cy) * d) / fy);
    o[2] = depth[i];
  }
}

/*
 * rgbdImage.cu:126-145 vertextoNormal: cross(V[i+2] - V[i-2], V[i+W] - V[i-W])
 * with out-of-range neighbours replaced by the pixel itself (K, including the
 * 1-D index that wraps across rows and the "<= 0" test for the upper
 * neighbour).  D: a neighbour without depth (vertex z == 0) makes the normal
 * NaN instead of a garbage direction; evaluated in FP64, rounded to float.
 */
OR_EXPORT void or_normal_map(const float *vert, int rows, int cols, float *norm) {
  int size = rows * cols;
  for (int i = 0; i < size; ++i) {
    int l = i - 2 < 0 ? i : i - 2;
    int r = i + 2 >= size ? i : i + 2;
    int u = i - cols <= 0 ? i : i - cols;
    int d = i + cols >= size ? i : i + cols;
    const float *L = vert + 3 * (size_t)l, *R = vert + 3 * (size_t)r, *U = vert + 3 * (size_t)u,
                *D = vert + 3 * (size_t)d;
    float *o = norm + 3 * (size_t)i;
    if (L[2] == 0.0f || R[2] == 0.0f || U[2] == 0.0f || D[2] == 0.0f || vert[3 * (size_t)i + 2] == 0.0f) {
      o[0] = o[1] = o[2] = NAN;
      continue;
    }
    double ax = (double)R[0] - (double)L[0], ay = (double)R[1] - (double)L[1], az = (double)R[2] - (double)L[2];
    double bx = (double)D[0] - (double)U[0], by = (double)D[1] - (double)U[1], bz = (double)D[2] - (double)U[2];
    double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    double n = sqrt(cx * cx + cy * cy + cz * cz);
    o[0] = (float)(cx / n); o[1] = (float)(cy / n); o[2] = (float)(cz / n);
  }
}

/* D: true half-resolution level by nearest subsampling, out(u,v) = in(2u,2v)
 * (the reference writes 1/s images into a centred window of the full-res
 * buffer with offsets hard-coded for 640x480, pyramid.cu:48-82). */
OR_EXPORT void or_subsample(const float *in, int rows, int cols, int channels, float *out) {
  int r2 = rows / 2, c2 = cols / 2;
  for (int v = 0; v < r2; ++v)
    for (int u = 0; u < c2; ++u)
      for (int c = 0; c < channels; ++c)
        out[((size_t)v * c2 + u) * channels + c] = in[((size_t)(2 * v) * cols + 2 * u) * channels + c];
}

/* ------------------------------------------------------------------ */
/* ICP normal equations, gnSolver.cu:25-196                             */
/* ------------------------------------------------------------------ */
/*
 * One Gauss-Newton linearisation of frame-to-model projective ICP at one
 * pyramid level.  Source maps are in the current camera, target maps in the
 * model camera; Tinc (row-major) maps source into target camera.
 * K: projective association by Project(); the normal-angle gate
 * |cos| > cos(pi/10); point-to-plane residual r = (p - q).n with
 * J = [p x n; n] for the perturbation exp(xi) * Tinc, xi = (w, t).
 * D: the source normal is rotated by Tinc before the angle test; the
 * projected pixel is bounds-checked; a point distance gate |p - q| <= dist_max
 * is applied; the depth gate uses the volume's [min, max] (the reference uses
 * the hard-coded 0.15..2.0 m); the transform, projection and gates are
 * evaluated in FP64 (the 10/5/4 damped schedule turns single flipped
 * decisions into ~1e-6 pose differences).  Photometric term (D, enabled): residual
 * I_t(pix) - I_s with the central-difference target gradient, chained through
 * the pinhole projection, weighted by w_photo.
 * Output: A (6x6 row-major) = sum J J^T, b (6) = sum J r, sums[0] = #geometric
 * residuals, sums[1] = sum r^2, sums[2] = #photometric residuals.
 * Accumulation is in double in both builds (the FP32 build still evaluates
 * each term in float).
 */
OR_EXPORT void or_icp_system(const float *sv, const float *sn, const float *sint, int rows, int cols,
                             double fx, double fy, double cx, double cy, const float *tv, const float *tn,
                             const float *tint, const double Tinc_in[16], double dmin, double dmax,
                             double dist_max, double w_photo, double *A, double *b, double *sums) {
  or_sensor s = {rows, cols, (real)fx, (real)fy, (real)cx, (real)cy};
  double Aa[36] = {0}, ba[6] = {0}, n_geo = 0, rr = 0, n_ph = 0;
  const double cos_gate = cos(3.14159265358979323846 / 10);
  for (int i = 0; i < rows * cols; ++i) {
    const float *V = sv + 3 * (size_t)i, *N = sn + 3 * (size_t)i;
    real sz = (real)V[2];
    if (!(sz >= (real)dmin && sz <= (real)dmax)) continue;
    r3 ns = mk3((real)N[0], (real)N[1], (real)N[2]);
    if (isbad(ns.x) || isbad(ns.y) || isbad(ns.z) || (ns.x == 0 && ns.y == 0 && ns.z == 0)) continue;
    /* D: association and gates in FP64 in both builds, so that the pixel
       choice and the inclusion decisions are those of the FP64 geometry; the
       residual and the Jacobian below are evaluated in `real` */
    const double *Td = Tinc_in;
    double px = Td[0] * V[0] + Td[1] * V[1] + Td[2] * V[2] + Td[3];
    double py = Td[4] * V[0] + Td[5] * V[1] + Td[6] * V[2] + Td[7];
    double pz = Td[8] * V[0] + Td[9] * V[1] + Td[10] * V[2] + Td[11];
    int pu = to_i32_d(round(((double)s.fx * px) / pz + (double)s.cx));
    int pv = to_i32_d(round(((double)s.fy * py) / pz + (double)s.cy));
    if (pu < 0 || pv < 0 || pu >= cols || pv >= rows) continue;
    int j = pv * cols + pu;
    const float *Q = tv + 3 * (size_t)j, *NT = tn + 3 * (size_t)j;
    real tz = (real)Q[2];
    if (!(tz >= (real)dmin && tz <= (real)dmax)) continue;
    r3 n = mk3((real)NT[0], (real)NT[1], (real)NT[2]);
    if (isbad(n.x) || isbad(n.y) || isbad(n.z) || (n.x == 0 && n.y == 0 && n.z == 0)) continue;
    double nrx = Td[0] * N[0] + Td[1] * N[1] + Td[2] * N[2];
    double nry = Td[4] * N[0] + Td[5] * N[1] + Td[6] * N[2];
    double nrz = Td[8] * N[0] + Td[9] * N[1] + Td[10] * N[2];
    double ct = (double)NT[0] * nrx + (double)NT[1] * nry + (double)NT[2] * nrz;
    if (!(fabs(ct) > cos_gate)) continue;
    double dqx = px - (double)Q[0], dqy = py - (double)Q[1], dqz = pz - (double)Q[2];
    if (!(sqrt(dqx * dqx + dqy * dqy + dqz * dqz) <= (double)(float)dist_max)) continue;
    r3 p = mk3((real)px, (real)py, (real)pz);
    r3 pq = mk3((real)dqx, (real)dqy, (real)dqz);
    real r = dot3(pq, n);
    r3 c = cross3(p, n);
    real J[6] = {c.x, c.y, c.z, n.x, n.y, n.z};
    for (int a = 0; a < 6; ++a) {
      for (int bb = 0; bb < 6; ++bb) Aa[6 * a + bb] += (double)(J[a] * J[bb]);
      ba[a] += (double)(J[a] * r);
    }
    n_geo += 1;
    rr += (double)(r * r);
    if (w_photo > 0 && sint && tint && pu >= 1 && pv >= 1 && pu < cols - 1 && pv < rows - 1) {
      real gx = ((real)tint[j + 1] - (real)tint[j - 1]) * (real)0.5;
      real gy = ((real)tint[j + cols] - (real)tint[j - cols]) * (real)0.5;
      real rp = (real)tint[j] - (real)sint[i];
      real iz = (real)1 / p.z;
      r3 G = mk3(gx * s.fx * iz, gy * s.fy * iz, -(gx * s.fx * p.x + gy * s.fy * p.y) * iz * iz);
      r3 cg = cross3(p, G);
      real Jp[6] = {cg.x, cg.y, cg.z, G.x, G.y, G.z};
      real wp = (real)w_photo;
      for (int a = 0; a < 6; ++a) {
        for (int bb = 0; bb < 6; ++bb) Aa[6 * a + bb] += (double)(wp * Jp[a] * Jp[bb]);
        ba[a] += (double)(wp * Jp[a] * rp);
      }
      n_ph += 1;
    }
  }
  memcpy(A, Aa, sizeof Aa);
  memcpy(b, ba, sizeof ba);
  sums[0] = n_geo; sums[1] = rr; sums[2] = n_ph;
}

/* ------------------------------------------------------------------ */
/* FP64 6x6 LDLT and SE(3) exp (gnSolver.cu:389-455 Eigen replacements)  */
/* ------------------------------------------------------------------ */
/* LDL^T without pivoting; A must be symmetric positive definite. */
OR_EXPORT int or_ldlt6(const double *Ain, const double *b, double *x) {
  double L[36] = {0}, D[6];
  for (int j = 0; j < 6; ++j) {
    double d = Ain[6 * j + j];
    for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k] * D[k];
    D[j] = d;
    if (!(d > 0)) return -1;
    L[6 * j + j] = 1;
    for (int i = j + 1; i < 6; ++i) {
      double a = Ain[6 * i + j];
      for (int k = 0; k < j; ++k) a -= L[6 * i + k] * L[6 * j + k] * D[k];
      L[6 * i + j] = a / d;
    }
  }
  double y[6];
  for (int i = 0; i < 6; ++i) {
    double a = b[i];
    for (int k = 0; k < i; ++k) a -= L[6 * i + k] * y[k];
    y[i] = a;
  }
  for (int i = 0; i < 6; ++i) y[i] /= D[i];
  for (int i = 5; i >= 0; --i) {
    double a = y[i];
    for (int k = i + 1; k < 6; ++k) a -= L[6 * k + i] * x[k];
    x[i] = a;
  }
  return 0;
}

/* exp of the 4x4 twist v2t(xi) (gnSolver.cu:446-455): xi = (w, t).
 * Closed form: R = I + a[w] + b[w]^2, V = I + b[w] + c[w]^2 with
 * a = sin/th, b = (1-cos)/th^2, c = (th - sin)/th^3 (series near 0). */
OR_EXPORT void or_se3_exp(const double xi[6], double T[16]) {
  double w0 = xi[0], w1 = xi[1], w2 = xi[2];
  double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
  double a, bb, c;
  if (th < 1e-5) {
    a = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    bb = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    c = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
  } else {
    a = sin(th) / th;
    bb = (1.0 - cos(th)) / th2;
    c = (th - sin(th)) / (th2 * th);
  }
  double W[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
  double W2[9];
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc) W2[3 * r + cc] = W[3 * r] * W[cc] + W[3 * r + 1] * W[3 + cc] + W[3 * r + 2] * W[6 + cc];
  double R[9], V[9];
  for (int k = 0; k < 9; ++k) {
    double I = (k % 4 == 0) ? 1.0 : 0.0;
    R[k] = I + a * W[k] + bb * W2[k];
    V[k] = I + bb * W[k] + c * W2[k];
  }
  for (int r = 0; r < 3; ++r) {
    T[4 * r] = R[3 * r]; T[4 * r + 1] = R[3 * r + 1]; T[4 * r + 2] = R[3 * r + 2];
    T[4 * r + 3] = V[3 * r] * xi[3] + V[3 * r + 1] * xi[4] + V[3 * r + 2] * xi[5];
  }
  T[12] = 0; T[13] = 0; T[14] = 0; T[15] = 1;
}

static void mat4_mul(const double *A, const double *B, double *C) {
  double t[16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      t[4 * r + c] = A[4 * r] * B[c] + A[4 * r + 1] * B[4 + c] + A[4 * r + 2] * B[8 + c] + A[4 * r + 3] * B[12 + c];
  memcpy(C, t, sizeof t);
}

/*
 * gnSolver.cu:389-416 step rule: scale A, b by 1/max(A); add
 * (lambda * max(A_scaled) / it) I; xi = LDLT^-1 (-b).
 */
OR_EXPORT int or_icp_step(const double *A_in, const double *b_in, int it, double lambda, double xi[6]) {
  double A[36], b[6];
  double mx = A_in[0];
  for (int k = 1; k < 36; ++k) if (A_in[k] > mx) mx = A_in[k];
  if (!(mx > 0)) { for (int k = 0; k < 6; ++k) xi[k] = 0; return -1; }
  double sc = 1.0 / mx;
  for (int k = 0; k < 36; ++k) A[k] = A_in[k] * sc;
  for (int k = 0; k < 6; ++k) b[k] = -b_in[k] * sc;
  double mx2 = A[0];
  for (int k = 1; k < 36; ++k) if (A[k] > mx2) mx2 = A[k];
  double alpha = lambda * mx2 / (double)it;
  for (int k = 0; k < 6; ++k) A[7 * k] += alpha;
  if (or_ldlt6(A, b, xi) != 0) { for (int k = 0; k < 6; ++k) xi[k] = 0; return -1; }
  return 0;
}

/*
 * Full multi-level frame-to-model ICP (gnSolver.cu:317-438 restated).
 * Levels run coarse to fine with at most iters[l] Gauss-Newton steps each;
 * Tinc <- exp(xi) Tinc.  K: a level ends once the summed twist changed by at
 * most min_inc since the last step that did not end a level
 * (gnSolver.cu:413-415); the final guard resets the increment when the
 * summed twist is NaN/Inf or its component sum is >= max_inc.
 * src/tgt_*: per level l = 0..2 maps of (rows >> l) x (cols >> l).
 * Output Tinc (row-major, source camera -> target camera).
 */
OR_EXPORT void or_icp_track(const float *const *sv, const float *const *sn, const float *const *sint,
                            const float *const *tv, const float *const *tn, const float *const *tint, int rows,
                            int cols, double fx, double fy, double cx, double cy, const int iters[3],
                            double dmin, double dmax, double dist_max, double w_photo, double lambda,
                            double max_inc, double min_inc, double Tinc[16]) {
  double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  double acc[6] = {0}, prev[6] = {0};
  for (int l = 2; l >= 0; --l) {
    double sc = (double)(1 << l);
    int r = rows >> l, c = cols >> l;
    for (int it = 1; it <= iters[l]; ++it) {
      double A[36], b[6], sums[3], xi[6], E[16];
      or_icp_system(sv[l], sn[l], sint ? sint[l] : NULL, r, c, fx / sc, fy / sc, cx / sc, cy / sc, tv[l], tn[l],
                    tint ? tint[l] : NULL, T, dmin, dmax, dist_max, w_photo, A, b, sums);
      or_icp_step(A, b, it, lambda, xi);
      or_se3_exp(xi, E);
      mat4_mul(E, T, T);
      for (int k = 0; k < 6; ++k) acc[k] += xi[k];
      double c2 = 0.0;
      for (int k = 0; k < 6; ++k) {
        const double c = acc[k] - prev[k];
        c2 += c * c;
      }
      if (sqrt(c2) <= min_inc) break;
      memcpy(prev, acc, sizeof acc);
    }
  }
  double s = acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5];
  if (s != s || isinf(s) || s >= max_inc) {
    double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    memcpy(T, I, sizeof I);
  }
  memcpy(Tinc, T, sizeof T);
}
#include "dm_oracle_mc.inc"
