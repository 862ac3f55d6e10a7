// Voxel-block allocation, TSDF/colour integration and raycasting.
//
// Reference: include/geometry/tsdf/tsdf.cu (AllocateFromDepthKernel :345-453,
// IntegrateScanKernel :267-344, GenerateDepthKernel/GenerateRgbKernel
// :470-526), include/dataStructure/hashTable/hashTable.cu (:80-163).
//
// HBM layout per session:
//   buckets  num_buckets x 128 B (for bucket_size <= 10): u64 key[S] (bit 63 =
//            occupied, 21 bits per coordinate), i32 pool index[S], i32 count.
//            One cache line per probe; the occupied slots are a prefix.
//   vox      num_blocks x 512 x 8 B voxel words, x-major local index
//            (x*64 + y*8 + z, include/geometry/geom_Types.h:116-119).
//   block_key pool index -> block coordinate.
#include "dm_device.h"

namespace dm {

__device__ __forceinline__ Intr intr_of(const DmView &v) {
  return Intr{v.rows, v.cols, v.fx, v.fy, v.cx, v.cy};
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}

// Bit of block (x, y, z) in the dense window bitmaps (kWin^3 blocks centred
// on block (0, 0, 0), x fastest); false outside the window.
__device__ __forceinline__ bool win_bit(int x, int y, int z, uint32_t &idx) {
  const unsigned X = (unsigned)(x + kWin / 2), Y = (unsigned)(y + kWin / 2), Z = (unsigned)(z + kWin / 2);
  if (X >= (unsigned)kWin || Y >= (unsigned)kWin || Z >= (unsigned)kWin) return false;
  idx = (Z * (unsigned)kWin + Y) * (unsigned)kWin + X;
  return true;
}
__device__ __forceinline__ bool test_bit(const uint32_t *map, uint32_t idx) { return (ldg(map + (idx >> 5)) >> (idx & 31)) & 1u; }

// Record a block that is missing from the table in the per-frame key set.
// A plain read first: once a key is in the set, later visitors return
// without an atomic (a missing free-space block near the camera is visited
// by thousands of rays).
__device__ __forceinline__ void note_new_key(const DmView &v, uint64_t key) {
  uint32_t h = (uint32_t)mix64(key) & (uint32_t)v.dset_mask;
  for (int probe = 0; probe <= v.dset_mask; ++probe) {
    uint64_t cur = __atomic_load_n((unsigned long long *)&v.dset[h], __ATOMIC_RELAXED);
    if (cur == key) return;
    if (cur == 0ull) {
      const uint64_t prev = atomicCAS((unsigned long long *)&v.dset[h], 0ull, (unsigned long long)key);
      if (prev == 0ull) {
        int idx = atomicAdd(&v.ctr[C_N_NEW], 1);
        if (idx < v.newcap) {
          v.newkeys[idx] = key;
          v.newdslot[idx] = (int32_t)h;
        } else {
          // no room in the frame's key list: release the slot so that a
          // later frame can insert the key
          atomicExch((unsigned long long *)&v.dset[h], 0ull);
          atomicAdd(&v.ctr[C_DSET_OVF], 1);
        }
        return;
      }
      if (prev == key) return;
    }
    h = (h + 1u) & (uint32_t)v.dset_mask;
  }
  atomicAdd(&v.ctr[C_DSET_OVF], 1);
}

// Per-workgroup set of the blocks its rays have already handled (looked up,
// and noted when missing): neighbouring rays visit mostly the same blocks,
// so each block is looked up in the global table about once per workgroup.
constexpr int kSeenSlots = 1024;
constexpr int kSeenProbes = 8;

// kCount (profiling only): also mark the pool index of every block found in
// the table in v.visited, so that the distinct blocks of a traversal can be
// counted.
// Inside the window the allocated-block bitmap answers "is it in the
// table" with one (mostly cached) word read, and only blocks missing from it
// reach the workgroup set and the per-frame key set; outside the window the
// hash table is probed.
template <bool kCount>
__device__ __forceinline__ void visit_block(const DmView &v, int x, int y, int z, unsigned long long *seen) {
  if (!in_key_range(x, y, z)) { atomicAdd(&v.ctr[C_DROP_RANGE], 1); return; }
  uint32_t widx;
  const bool in_win = !kCount && win_bit(x, y, z, widx);
  if (in_win && test_bit(v.allocmap, widx)) return;
  const uint64_t key = pack_key(x, y, z);
  uint32_t h = (uint32_t)mix64(key) & (kSeenSlots - 1);
  for (int probe = 0; probe < kSeenProbes; ++probe) {
    const unsigned long long cur = seen[h];
    if (cur == key) return;
    if (cur == 0ull) {
      const unsigned long long prev = atomicCAS(&seen[h], 0ull, (unsigned long long)key);
      if (prev == key) return;
      if (prev == 0ull) break;  // this thread handles the block for the workgroup
    }
    h = (h + 1u) & (kSeenSlots - 1);
  }
  if (!in_win) {
    const int ptr = find_block(v, x, y, z);
    if (ptr >= 0) {
      if (kCount) atomicOr(&v.visited[ptr >> 5], 1u << (ptr & 31));
      return;
    }
  }
  note_new_key(v, key);
}

// AllocateFromDepthKernel (tsdf.cu:357-452), one thread per pixel, 16x16
// pixel tiles so that neighbouring rays (which share most blocks) run in the
// same workgroup.  K: ray from the camera centre to surface + trunc, the
// (b + step) * bs boundary for both directions and the extra "neg_ray" block.
// D: the traversal is bounded by the L1 distance to the end block.
template <bool kCount>
__device__ __forceinline__ int alloc_ray(const DmView &v, int u, int r, unsigned long long *seen) {
  if (u >= v.cols || r >= v.rows) return 0;
  const int i = r * v.cols + u;
  const float d = v.depth[i];
  if (d < v.min_depth || d > v.max_depth) {
    v.surf[i] = make_uint2(__float_as_uint(__builtin_nanf("")), 0u);
    return 0;
  }
  const Intr s = intr_of(v);
  const f3 p3 = get_point3d_uv(u, r, d, s);
  // the integration's surface distance and colour of this pixel
  // (tsdf.cu:319-322, :329-331)
  const float surf = norm(mk(0.f, 0.f, 0.f) - p3);
  const uint32_t rgb = v.rgb ? (uint32_t)v.rgb[3 * i] | ((uint32_t)v.rgb[3 * i + 1] << 8) |
                                   ((uint32_t)v.rgb[3 * i + 2] << 16)
                             : 0u;
  v.surf[i] = make_uint2(__float_as_uint(d < v.max_depth ? surf : -surf), rgb);
  const M4 T = load_m4(v.posef + PF_T);
  const float block_m = 8.0f * v.voxel_size;
  const f3 start = mk(T.m[3], T.m[7], T.m[11]);
  const f3 point = xform(T, p3);
  if (point.x == 0 && point.y == 0 && point.z == 0) return 0;
  const f3 dir = normalize(point - start);
  const float sd = norm(start - point);
  const f3 end = start + dir * (sd + v.trunc);
  int bx = to_i32(floorf(start.x / block_m)), by = to_i32(floorf(start.y / block_m)),
      bz = to_i32(floorf(start.z / block_m));
  const int ex = to_i32(floorf(end.x / block_m)), ey = to_i32(floorf(end.y / block_m)),
            ez = to_i32(floorf(end.z / block_m));
  const int stx = (dir.x > 0) - (dir.x < 0), sty = (dir.y > 0) - (dir.y < 0), stz = (dir.z > 0) - (dir.z < 0);
  const float FMAX = 3.40282347e+38f;
  const float dtx = dir.x != 0 ? fabsf(block_m / dir.x) : FMAX;
  const float dty = dir.y != 0 ? fabsf(block_m / dir.y) : FMAX;
  const float dtz = dir.z != 0 ? fabsf(block_m / dir.z) : FMAX;
  const float bdx = ((float)bx + (float)stx) * block_m, bdy = ((float)by + (float)sty) * block_m,
              bdz = ((float)bz + (float)stz) * block_m;
  float mtx = dir.x != 0 ? (bdx - start.x) / dir.x : FMAX;
  float mty = dir.y != 0 ? (bdy - start.y) / dir.y : FMAX;
  float mtz = dir.z != 0 ? (bdz - start.z) / dir.z : FMAX;
  int dx = 0, dy = 0, dz = 0;
  bool neg = false;
  if (bx != ex && dir.x < 0) { dx--; neg = true; }
  if (by != ey && dir.y < 0) { dy--; neg = true; }
  if (bz != ez && dir.z < 0) { dz--; neg = true; }
  visit_block<kCount>(v, bx, by, bz, seen);
  int nvisit = 1;
  if (neg) {
    bx += dx; by += dy; bz += dz;
    visit_block<kCount>(v, bx, by, bz, seen);
    ++nvisit;
  }
  const long long bound64 = llabs((long long)ex - bx) + llabs((long long)ey - by) + llabs((long long)ez - bz);
  const int bound = (int)min(bound64, (long long)0x7fffffff);
  // Every step moves one axis by one block, so the traversal stays within
  // `bound` blocks of (bx, by, bz), and the L1 distance to the end block,
  // `bound` at the start, falls by at most one per step: the end block can
  // only be reached at step `bound`, so the loop takes exactly `bound` steps
  // and needs no end-block test.  When that cube lies inside the window (and
  // so inside the key range), only the window index of the current block is
  // stepped; a block whose allocated bit is clear (rare once the scene is
  // mapped) has its coordinates decoded from that index for visit_block
  // (which starts with the same bit test).
  uint32_t widx;
  if (!kCount && bound < kWin / 2 && win_bit(bx - bound, by - bound, bz - bound, widx) &&
      win_bit(bx + bound, by + bound, bz + bound, widx)) {
    win_bit(bx, by, bz, widx);
    const uint32_t ix = (uint32_t)stx, iy = (uint32_t)(sty * kWin), iz = (uint32_t)(stz * kWin * kWin);
    for (int it = 0; it < bound; ++it) {
      const bool sx = mtx < mty && mtx < mtz;
      const bool sy = !(mtx < mty) && mty < mtz;
      const bool sz = !sx && !sy;
      mtx += sx ? dtx : 0.0f;
      mty += sy ? dty : 0.0f;
      mtz += sz ? dtz : 0.0f;
      widx += sx ? ix : sy ? iy : iz;
      if (!test_bit(v.allocmap, widx))
        visit_block<kCount>(v, (int)(widx & (kWin - 1)) - kWin / 2, (int)((widx / kWin) & (kWin - 1)) - kWin / 2,
                            (int)(widx / (kWin * kWin)) - kWin / 2, seen);
    }
    return nvisit + bound;
  }
  for (int it = 0; it < bound; ++it) {
    // the reference's nested comparisons, as selects: x when mtx is below
    // both, else y when mty < mtz, else z
    const bool sx = mtx < mty && mtx < mtz;
    const bool sy = !(mtx < mty) && mty < mtz;
    const bool sz = !sx && !sy;
    bx += sx ? stx : 0;
    mtx += sx ? dtx : 0.0f;
    by += sy ? sty : 0;
    mty += sy ? dty : 0.0f;
    bz += sz ? stz : 0;
    mtz += sz ? dtz : 0.0f;
    visit_block<kCount>(v, bx, by, bz, seen);
  }
  return nvisit + bound;
}

template <bool kCount>
__global__ __launch_bounds__(256) void k_alloc_rays(const DmView *__restrict__ views,
                                                     const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int tiles_x = (v.cols + 15) >> 4;
  const int tile = blockIdx.x;
  if (tile >= tiles_x * ((v.rows + 15) >> 4)) return;
  __shared__ int visits;
  __shared__ unsigned long long seen[kSeenSlots];
  for (int k = threadIdx.x; k < kSeenSlots; k += 256) seen[k] = 0ull;
  if (threadIdx.x == 0) visits = 0;
  __syncthreads();
  const int u = (tile % tiles_x) * 16 + (threadIdx.x & 15);
  const int r = (tile / tiles_x) * 16 + (threadIdx.x >> 4);
  const int n = alloc_ray<kCount>(v, u, r, seen);
  if (n) atomicAdd(&visits, n);
  __syncthreads();
  if (threadIdx.x == 0 && visits) atomicAdd(&v.ctr[C_NVISIT], visits);
}


__device__ __forceinline__ int key_bucket(const DmView &v, uint64_t key) {
  int x, y, z;
  unpack_key(key, x, y, z);
  return hash_bucket(x, y, z, v.num_buckets);
}

// Chain the new keys of each bucket into a list (order irrelevant).
__global__ __launch_bounds__(256) void k_alloc_link(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  int n = min(v.ctr[C_N_NEW], v.newcap);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    int b = key_bucket(v, v.newkeys[i]);
    v.newnext[i] = atomicExch(&v.bucket_head[b], i);
  }
}

// Canonical rank of each new key within its bucket (ascending packed key);
// the first `free` keys of a bucket win (K: no overflow chain, the rest drop).
__global__ __launch_bounds__(256) void k_alloc_rank(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  int n = min(v.ctr[C_N_NEW], v.newcap);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint64_t key = v.newkeys[i];
    const int b = key_bucket(v, key);
    int rank = 0, len = 0;
    for (int j = v.bucket_head[b]; j >= 0; j = v.newnext[j]) {
      rank += v.newkeys[j] < key;
      ++len;
    }
    const int freeb = v.bucket_size - bucket_count(v, bucket_ptr(v, b));
    if (rank < freeb) {
      v.newrank[i] = rank;
      atomicAdd(&v.ctr[C_N_INS], 1);
    } else {
      v.newrank[i] = -1;
    }
    v.newlen[i] = len;
  }
}

// Insert the winners.  D: if the pool cannot take every insertion of the
// frame, nothing is inserted and all new keys count as dropped_pool.
__global__ __launch_bounds__(256) void k_alloc_commit(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int n = min(v.ctr[C_N_NEW], v.newcap);
  const int nalloc = v.ctr[C_NUM_ALLOC];
  const int nins = v.ctr[C_N_INS];
  const int nfree = v.ctr[C_NUM_FREE];
  if (nalloc + max(0, nins - nfree) > v.num_blocks) return;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int rank = v.newrank[i];
    if (rank < 0) continue;
    const uint64_t key = v.newkeys[i];
    int x, y, z;
    unpack_key(key, x, y, z);
    const int b = hash_bucket(x, y, z, v.num_buckets);
    uint8_t *bk = v.buckets + (size_t)b * (size_t)v.bstride;
    const int slot = *(const int32_t *)(bk + 12 * v.bucket_size) + rank;
    // freed pool indices first, top of the stack down (Heap::Consume,
    // heap.cu:14-17), then fresh ones
    const int q = atomicAdd(&v.ctr[C_CURSOR], 1);
    const int ptr = q < nfree ? v.free_list[nfree - 1 - q] : nalloc + (q - nfree);
    ((uint64_t *)bk)[slot] = key;
    ((int32_t *)(bk + 8 * v.bucket_size))[slot] = ptr;
    v.block_key[ptr] = make_int4(x, y, z, 1);
    uint32_t widx;
    if (win_bit(x, y, z, widx)) atomicOr(&v.allocmap[widx >> 5], 1u << (widx & 31));
  }
}

// Publish bucket counts, clear the per-frame structures.
__global__ __launch_bounds__(256) void k_alloc_finish(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int n = min(v.ctr[C_N_NEW], v.newcap);
  const bool pool_ok = v.ctr[C_NUM_ALLOC] + max(0, v.ctr[C_N_INS] - v.ctr[C_NUM_FREE]) <= v.num_blocks;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const uint64_t key = v.newkeys[i];
    const int b = key_bucket(v, key);
    v.dset[v.newdslot[i]] = 0ull;
    v.bucket_head[b] = -1;
    if (pool_ok && v.newrank[i] == 0) {
      int32_t *cnt = (int32_t *)(v.buckets + (size_t)b * (size_t)v.bstride + 12 * v.bucket_size);
      const int freeb = v.bucket_size - *cnt;
      *cnt += min(v.newlen[i], freeb);
    }
  }
}

// Counter epilogue (one thread per session): runs after k_alloc_finish.
__global__ void k_alloc_close(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.x];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  if (threadIdx.x != 0) return;
  const int n_all = v.ctr[C_N_NEW];
  const int n = min(n_all, v.newcap);
  const int nins = v.ctr[C_N_INS];
  const int nfree = v.ctr[C_NUM_FREE];
  if (v.ctr[C_NUM_ALLOC] + max(0, nins - nfree) <= v.num_blocks) {
    v.ctr[C_NUM_ALLOC] += max(0, nins - nfree);
    v.ctr[C_NUM_FREE] = max(0, nfree - nins);
    v.ctr[C_DROP_BUCKET] += n - nins;
    v.ctr[C_NNEW] = nins;
  } else {
    v.ctr[C_DROP_POOL] += n;
    v.ctr[C_NNEW] = 0;
  }
  v.ctr[C_N_NEW] = 0;
  v.ctr[C_N_INS] = 0;
  v.ctr[C_CURSOR] = 0;
}


// ---------------------------------------------------------------------------
// Integration (IntegrateScanKernel, tsdf.cu:267-344; merge geom_Types.h:79-102)
// One workgroup per voxel block (grid-stride over the allocated blocks), two
// voxels per thread, coalesced 8-byte voxel words; unchanged voxels are not
// written back.
// One colour channel of Voxel::mergeVoxel (geom_Types.h:85-90):
// uchar((c*w + c_new) / (w + 1)) with the float quotient truncated.  Here n =
// c*w + c_new and m = w + 1 are integers (c, c_new, w <= 255, so n < 2^16 and
// n/m <= 255), and the correctly rounded float quotient truncates to
// floor(n/m): a non-integer n/m lies at least 1/m >= 2^-8 from the integers,
// far beyond the quotient's rounding error (2^-24 * 255).  floor(n/m) is
// taken as trunc((n + 1/2) * rm) with rm = rcp(m) (v_rcp_f32, 1 ulp):
// (n + 1/2)/m lies at least 1/(2m) >= 2^-9 from every integer, while the
// reciprocal's and the product's rounding move it by at most
// 255.5 * (2^-23 + 2^-24) < 2^-14.
__device__ __forceinline__ uint32_t merge_channel(uint32_t c, uint32_t w, uint32_t c_new, float rm) {
  return (uint32_t)(((float)(c * w + c_new) + 0.5f) * rm);
}

// One voxel of IntegrateScanKernel (tsdf.cu:301-341): returns the new voxel
// word (the old one when the voxel is not updated).
__device__ __forceinline__ uint64_t integrate_voxel(const DmView &v, const Intr &s, const M4 &Ti, f3 pos, int t,
                                                    uint64_t old) {
  const float vs = v.voxel_size, tau = v.trunc;
  const int bx = t >> 6, by = (t >> 3) & 7, bz = t & 7;
  const f3 vp = mk(pos.x + (float)bx * vs, pos.y + (float)by * vs, pos.z + (float)bz * vs);
  const f3 vc = xform(Ti, vp);
  int pu, pv;
  project(vc, s, pu, pv);
  if (!(pu >= 0 && pv >= 0 && pu < s.cols && pv < s.rows)) return old;
  const int idx = pv * s.cols + pu;
  // |getPoint3d(pu, pv, d)| from the allocation pass, sign bit set when
  // d == max_depth (no band update); NaN when d is out of [min_depth,
  // max_depth] or NaN, where the reference's comparisons all fail and the
  // voxel keeps its value
  const uint64_t sp = ldg((const uint64_t *)(v.surf + idx));
  const float sv = __uint_as_float((uint32_t)sp);
  const uint32_t prgb = (uint32_t)(sp >> 32);
  if (sv != sv) return old;
  const float surf = fabsf(sv);
  const float vox = norm(mk(0.f, 0.f, 0.f) - vc);
  float nsdf;
  if (vox > surf - tau && vox < surf + tau && !signbit(sv)) nsdf = surf - vox;
  else if (vox < surf - tau) nsdf = tau;
  else return old;
  const uint32_t nr = prgb & 0xff, ng = (prgb >> 8) & 0xff, nb = (prgb >> 16) & 0xff;
  const uint32_t rgbw = vox_rgbw(old);
  const uint32_t wi = rgbw >> 24;
  const float w = (float)wi;
  const float sdf = (vox_sdf(old) * w + nsdf * 1.0f) / (w + 1.0f);
  const float rm = __builtin_amdgcn_rcpf((float)(wi + 1u));
  const uint32_t cr = merge_channel(rgbw & 0xff, wi, nr, rm);
  const uint32_t cg = merge_channel((rgbw >> 8) & 0xff, wi, ng, rm);
  const uint32_t cb = merge_channel((rgbw >> 16) & 0xff, wi, nb, rm);
  int nw = (int)wi + 1;
  if (nw > v.max_weight) nw = v.max_weight;
  const uint32_t nrgbw = cr | (cg << 8) | (cb << 16) | ((uint32_t)nw << 24);
  return ((uint64_t)nrgbw << 32) | (uint64_t)__float_as_uint(sdf);
}

// Far threshold of the raycast's empty-space test: a sample whose voxels all
// have sdf at or above it (or are unobserved) interpolates to at least
// 0.5*trunc, the march's full step (margin for the rounding of the
// trilinear weights).
__device__ __forceinline__ float far_threshold(float trunc) {
  const float big = 0.5f * trunc;
  return big + big * (1.0f / 256.0f);
}

// Cell of 2^S blocks per axis of a block coordinate (S = 3: the coarse
// cells, S = 2: the 4^3 cells), and its bit in a map of (kWin >> S)^3 cells
// over the window; false outside it.
template <int S> __device__ __forceinline__ int cell_of(int b) { return b >> S; }
template <int S> __device__ __forceinline__ bool cell_bit(int cx, int cy, int cz, int &word, int &bit) {
  constexpr int n = kWin >> S;
  const int x = cx + n / 2, y = cy + n / 2, z = cz + n / 2;
  if ((unsigned)x >= (unsigned)n || (unsigned)y >= (unsigned)n || (unsigned)z >= (unsigned)n) return false;
  const int idx = (z * n + y) * n + x;
  word = idx >> 5;
  bit = idx & 31;
  return true;
}
__device__ __forceinline__ int coarse_of(int b) { return cell_of<3>(b); }
__device__ __forceinline__ bool coarse_bit(int cx, int cy, int cz, int &word, int &bit) {
  return cell_bit<3>(cx, cy, cz, word, bit);
}
// Sets the bits of the cells (2^S blocks per axis) that hold a block of the
// 3^3 neighbourhood of block (x, y, z): at most 2 per axis.
template <int S> __device__ __forceinline__ void mark_cells(uint32_t *map, int x, int y, int z) {
  for (int cz = cell_of<S>(z - 1); cz <= cell_of<S>(z + 1); ++cz)
    for (int cy = cell_of<S>(y - 1); cy <= cell_of<S>(y + 1); ++cy)
      for (int cx = cell_of<S>(x - 1); cx <= cell_of<S>(x + 1); ++cx) {
        int w, b;
        if (!cell_bit<S>(cx, cy, cz, w, b)) continue;
        if (!(map[w] & (1u << b))) atomicOr(&map[w], 1u << b);
      }
}

// Bit of block (x, y, z) in the block window, false outside it.

// Row (dy, dz) = (row % 3 - 1, row / 3 - 1) of the 3^3 neighbourhood of block
// (x, y, z) in the window: bits x-1..x+1 (clipped to the window) lie in words
// w0 and w1 (equal or adjacent) under mask0 and mask1.  False when the row is
// outside the window.
__device__ __forceinline__ bool nbr_row(int x, int y, int z, int row, uint32_t &w0, uint32_t &w1, uint32_t &mask0,
                                        uint32_t &mask1) {
  const int yy = y + row % 3 - 1, zz = z + row / 3 - 1;
  const int xl = max(x - 1, -kWin / 2), xh = min(x + 1, kWin / 2 - 1);
  uint32_t i0;
  if (xl > xh || !win_bit(xl, yy, zz, i0)) return false;
  const uint32_t i1 = i0 + (uint32_t)(xh - xl);
  w0 = i0 >> 5;
  w1 = i1 >> 5;
  mask0 = 0u;
  mask1 = 0u;
  for (uint32_t i = i0; i <= i1; ++i) {
    if ((i >> 5) == w0) mask0 |= 1u << (i & 31);
    else mask1 |= 1u << (i & 31);
  }
  return true;
}

__global__ __launch_bounds__(256) void k_integrate(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int nalloc = v.ctr[C_NUM_ALLOC];
  const Intr s = intr_of(v);
  const M4 Ti = load_m4(v.posef + PF_TINV);
  const float vs = v.voxel_size, tau = v.trunc;
  const double half = 0.5 * (double)vs * 8.0;
  const float far_thr = far_threshold(tau);
  // the raycast's bitmaps are rebuilt by k_dilate_near: the cell maps are
  // cleared whole, the block-level ones at the words that the near blocks of
  // the previous frame set (whole words: every bit in them belongs to a block
  // near one of those, and k_dilate_near sets it again if it still is)
  for (int w = blockIdx.x * 256 + threadIdx.x; w < kCoarseWords; w += gridDim.x * 256) v.coarse[w] = 0u;
  for (int w = blockIdx.x * 256 + threadIdx.x; w < kCell4Words; w += gridDim.x * 256) v.cell4[w] = 0u;
  int nvis = 0;
  for (int blk = blockIdx.x; blk < nalloc; blk += gridDim.x) {
    const int4 key = v.block_key[blk];
    if (key.w == 0) continue;  // freed pool index (uniform per workgroup)
    if (threadIdx.x < 10 && v.blk_near[blk]) {
      if (threadIdx.x < 9) {
        uint32_t w0, w1, m0, m1;
        if (nbr_row(key.x, key.y, key.z, threadIdx.x, w0, w1, m0, m1)) {
          v.dilmap[w0] = 0u;
          v.dilmap[w1] = 0u;
        }
      } else {
        uint32_t idx;
        if (win_bit(key.x, key.y, key.z, idx)) v.nearmap[idx >> 5] = 0u;
      }
    }
    const f3 pos = mk((float)key.x * vs * 8.0f, (float)key.y * vs * 8.0f, (float)key.z * vs * 8.0f);
    const f3 pc = xform(Ti, pos);
    if (pc.z < 0) continue;
    const f3 bc = mk((float)((double)pc.x + half), (float)((double)pc.y + half), (float)((double)pc.z + half));
    int iu, iv;
    project(bc, s, iu, iv);
    if (!(iu >= 0 && iv >= 0 && iu < s.cols && iv < s.rows)) continue;
    ++nvis;
    uint64_t *bv = v.vox + (size_t)blk * kBlockVox;
    bool near = false;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = threadIdx.x + 256 * h;
      const uint64_t old = bv[t];
      const uint64_t nword = integrate_voxel(v, s, Ti, pos, t, old);
      if (nword != old) bv[t] = nword;
      near |= (nword >> 56) != 0 && vox_sdf(nword) < far_thr;
    }
    // block summary used by the raycast's empty-space test
    near = __syncthreads_or(near);
    if (threadIdx.x == 0) v.blk_near[blk] = near ? 1 : 0;
  }
  if (threadIdx.x == 0 && nvis) atomicAdd(&v.ctr[C_NVIS], nvis);
}

// Rebuild the raycast's bitmaps from the near blocks (after the integration
// has cleared the words of the previous frame's near blocks): the window bit
// of every near block, the dilated bits of its 3^3 neighbourhood, and the
// 4^3-block and coarse cells that hold a block of that neighbourhood (at most
// 8 each).
__global__ __launch_bounds__(256) void k_dilate_near(const DmView *__restrict__ views, const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int nalloc = v.ctr[C_NUM_ALLOC];
  for (int blk = blockIdx.x * 256 + threadIdx.x; blk < nalloc; blk += gridDim.x * 256) {
    if (!v.blk_near[blk]) continue;
    const int4 key = v.block_key[blk];
    uint32_t idx;
    if (win_bit(key.x, key.y, key.z, idx)) atomicOr(&v.nearmap[idx >> 5], 1u << (idx & 31));
    for (int row = 0; row < 9; ++row) {
      uint32_t w0, w1, m0, m1;
      if (!nbr_row(key.x, key.y, key.z, row, w0, w1, m0, m1)) continue;
      if ((v.dilmap[w0] & m0) != m0) atomicOr(&v.dilmap[w0], m0);
      if (m1 && (v.dilmap[w1] & m1) != m1) atomicOr(&v.dilmap[w1], m1);
    }
    mark_cells<2>(v.cell4, key.x, key.y, key.z);
    mark_cells<3>(v.coarse, key.x, key.y, key.z);
  }
}

// Window bit of every allocated block in the allocation bitmap (rebuild).
__global__ __launch_bounds__(256) void k_mark_allocated(const DmView *__restrict__ views,
                                                       const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const int nalloc = v.ctr[C_NUM_ALLOC];
  for (int blk = blockIdx.x * 256 + threadIdx.x; blk < nalloc; blk += gridDim.x * 256) {
    const int4 key = v.block_key[blk];
    uint32_t idx;
    if (key.w != 0 && win_bit(key.x, key.y, key.z, idx)) atomicOr(&v.allocmap[idx >> 5], 1u << (idx & 31));
  }
}

// ---------------------------------------------------------------------------
// Block deletion (HashTable::DeleteBlock, hashTable.cu:112-134, with
// Heap::Append, heap.cu:19-22).  D: a bucket keeps its occupied slots as a
// prefix, so the survivors of a bucket close up in their order (a stable
// compaction), which keeps the table canonical; freed pool indices go on a
// stack that later insertions take first.  Three kernels per request list:
// mark (one thread per requested block), compact (one thread per touched
// bucket), clear (one workgroup per freed block).

// One requested block: mark its slot for removal.  dmask (one word per
// bucket, zero outside a request) collects the slots; the first mark of a
// bucket lists it in touched.  Freed pool indices are pushed on free_list.
__global__ __launch_bounds__(256) void k_del_mark(const DmView *__restrict__ views, const int4 *__restrict__ req,
                                                 int n, unsigned long long *__restrict__ dmask,
                                                 int *__restrict__ touched, int *__restrict__ counts) {
  const DmView &v = views[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int4 q = req[i];
  if (!in_key_range(q.x, q.y, q.z)) return;
  const uint64_t key = pack_key(q.x, q.y, q.z);
  const int b = hash_bucket(q.x, q.y, q.z, v.num_buckets);
  const uint8_t *bk = bucket_ptr(v, b);
  const int cnt = bucket_count(v, bk);
  for (int s = 0; s < cnt; ++s) {
    if (((const uint64_t *)bk)[s] != key) continue;
    const unsigned long long bit = 1ull << s;
    const unsigned long long old = atomicOr(&dmask[b], bit);
    if (old & bit) return;  // the same block listed twice
    if (old == 0ull) touched[atomicAdd(&counts[0], 1)] = b;
    const int ptr = ((const int32_t *)(bk + 8 * v.bucket_size))[s];
    v.block_key[ptr].w = 0;
    v.free_list[v.ctr[C_NUM_FREE] + atomicAdd(&counts[1], 1)] = ptr;
    return;
  }
}

// Stable compaction of each touched bucket's occupied prefix.
__global__ __launch_bounds__(256) void k_del_compact(const DmView *__restrict__ views,
                                                    unsigned long long *__restrict__ dmask,
                                                    const int *__restrict__ touched, const int *__restrict__ counts) {
  const DmView &v = views[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= counts[0]) return;
  const int b = touched[i];
  uint8_t *bk = v.buckets + (size_t)b * (size_t)v.bstride;
  uint64_t *keys = (uint64_t *)bk;
  int32_t *ptrs = (int32_t *)(bk + 8 * v.bucket_size);
  int32_t *cnt = (int32_t *)(bk + 12 * v.bucket_size);
  const unsigned long long m = dmask[b];
  const int c = *cnt;
  int o = 0;
  for (int s = 0; s < c; ++s) {
    if ((m >> s) & 1ull) continue;
    keys[o] = keys[s];
    ptrs[o] = ptrs[s];
    ++o;
  }
  for (int s = o; s < c; ++s) { keys[s] = 0ull; ptrs[s] = 0; }
  *cnt = o;
  dmask[b] = 0ull;
}

// Zero the freed blocks' voxels and flags (DeleteBlock's voxel reset) and
// clear their allocation bits; then the counters.
__global__ __launch_bounds__(256) void k_del_clear(const DmView *__restrict__ views, const int *__restrict__ counts) {
  const DmView &v = views[0];
  const int nfreed = counts[1];
  const int base = v.ctr[C_NUM_FREE];
  for (int k = blockIdx.x; k < nfreed; k += gridDim.x) {
    const int ptr = v.free_list[base + k];
    uint64_t *bv = v.vox + (size_t)ptr * kBlockVox;
    bv[threadIdx.x] = 0ull;
    bv[threadIdx.x + 256] = 0ull;
    if (threadIdx.x == 0) {
      v.blk_near[ptr] = 0;
      const int4 key = v.block_key[ptr];
      uint32_t idx;
      if (win_bit(key.x, key.y, key.z, idx)) atomicAnd(&v.allocmap[idx >> 5], ~(1u << (idx & 31)));
    }
  }
}

__global__ void k_del_close(const DmView *__restrict__ views, int *__restrict__ counts) {
  const DmView &v = views[0];
  if (threadIdx.x != 0) return;
  v.ctr[C_NUM_FREE] += counts[1];
}

// Recycling selection: every live block whose centre ((key + 1/2) * 8 * vs)
// lies farther than radius from center, listed in req.
__global__ __launch_bounds__(256) void k_select_far(const DmView *__restrict__ views, double cx, double cy, double cz,
                                                   double r2, int4 *__restrict__ req, int *__restrict__ nreq) {
  const DmView &v = views[0];
  const int nalloc = v.ctr[C_NUM_ALLOC];
  const double bm = 8.0 * (double)v.voxel_size;
  for (int blk = blockIdx.x * 256 + threadIdx.x; blk < nalloc; blk += gridDim.x * 256) {
    const int4 key = v.block_key[blk];
    if (key.w == 0) continue;
    const double dx = ((double)key.x + 0.5) * bm - cx, dy = ((double)key.y + 0.5) * bm - cy,
                 dz = ((double)key.z + 0.5) * bm - cz;
    if (dx * dx + dy * dy + dz * dz > r2) req[atomicAdd(nreq, 1)] = key;
  }
}

// Clears every bitmap word (grid-stride) before a full rebuild.
__global__ __launch_bounds__(256) void k_clear_far_maps(const DmView *__restrict__ views,
                                                        const int *__restrict__ active) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < kWinWords; w += stride) {
    v.nearmap[w] = 0u;
    v.dilmap[w] = 0u;
    v.allocmap[w] = 0u;
  }
  for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < (size_t)kCoarseWords; w += stride) v.coarse[w] = 0u;
  for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < (size_t)kCell4Words; w += stride) v.cell4[w] = 0u;
}

// ---------------------------------------------------------------------------
// Raycast (GenerateDepthKernel / GenerateRgbKernel, tsdf.cu:470-526).
// K: samples are GetInterpolatedVoxel at T * getPoint3d(i, z).
// D: step 0.5*trunc through unobserved space, min(sdf, 0.5*trunc) (at least
// half a voxel) through observed space; the hit is the first +/- zero
// crossing, linearly interpolated; normals are the central-difference
// gradient of the interpolated SDF, in camera coordinates.  The model maps of
// all pyramid levels are written by subsampling (pixel (2^l u, 2^l v)).
__device__ __forceinline__ float intensity_of(uint32_t r, uint32_t g, uint32_t b) {
  return (0.2126f * (float)r + 0.7152f * (float)g + 0.0722f * (float)b) * (1.0f / 255.0f);
}

// GetInterpolatedVoxel (tsdf.cu:47-175) for the raycast.  Returns false when
// the centre voxel (p/vs rounded half away from zero) is unobserved; the 8
// corners are the voxels of p - vs/2 and p - vs/2 + vs per axis, weighted by
// p/vs - floor(p/vs), and an unobserved corner (integration weight 0)
// contributes the centre voxel instead.
// The block pointers of the samples are kept in a per-thread window of 2x2x2
// blocks.  The window's pool indices sit in LDS (thread-strided), a register
// bit per slot says which are looked up, and a slot is filled from the hash
// table on first use, an unallocated block with the index of the all-zero
// block past the pool; the nine voxel words are then loaded together.  A
// sample whose voxels span more than two blocks on an axis (possible only
// through rounding at huge coordinates) looks its voxels up one by one.
struct BlockWin {
  int wx, wy, wz;      // window base block
  unsigned resolved;   // bit k < 8: LDS slot k holds the pool index; bits 8, 9, 10: the ray
                       // runs towards -x, -y, -z (the window extends that way)
  int *slot;           // this thread's slots: slot[k * 256], k = x + 2y + 4z
};

// voxel word of global voxel (gx, gy, gz) through the hash table
__device__ __forceinline__ uint64_t voxel_lookup(const DmView &v, int gx, int gy, int gz) {
  const int ptr = find_block(v, floordiv8(gx), floordiv8(gy), floordiv8(gz));
  if (ptr < 0) return 0ull;
  return ldg(v.vox + ((uint32_t)ptr * (uint32_t)kBlockVox + (uint32_t)(mod8(gx) * 64 + mod8(gy) * 8 + mod8(gz))));
}

// The per-axis part of GetInterpolatedVoxel at coordinate q (the three axes
// are independent): centre voxel (q/vs rounded half away from zero), the two
// corner voxels (those of q - vs/2 and q - vs/2 + vs) and the trilinear
// weight q/vs - floor(q/vs).
struct AxisSample {
  int c, i0, i1;
  float w;
};
// The corner voxels world_to_voxel(q - vs/2) and world_to_voxel(q - vs/2 +
// vs) are floor(q/vs) and floor(q/vs) + 1 in exact arithmetic (the +-1/2 of
// the dual offset and of the rounding cancel, on either sign); with the
// float roundings they still are unless q/vs lies within a few ulp of an
// integer.  vq = q * (1/vs) is within 1.5 ulp of q / vs, far inside the
// margin |vq| * 1e-6 + 1e-6, so the corners are taken from floor(vq) when its
// fraction is further than that from 0 and 1 and evaluated exactly
// otherwise; the centre voxel is world_to_voxel_fast, exact by the same
// argument.  Only the weight carries vq's rounding (real-valued outputs).
__device__ __forceinline__ AxisSample axis_sample(float q, float vs, float inv_vs) {
  AxisSample a;
  const float vq = q * inv_vs;
  a.c = world_to_voxel_fast(q, vs, inv_vs);
  const float fl = floorf(vq);
  a.w = vq - fl;
  const float tol = fabsf(vq) * 1.0e-6f + 1.0e-6f;
  if (a.w > tol && a.w < 1.0f - tol && fabsf(vq) < 1.0e8f) {
    a.i0 = (int)fl;
    a.i1 = a.i0 + 1;
  } else {
    const float dual = q - vs / 2.0f;
    a.i0 = world_to_voxel_fast(dual + 0.0f, vs, inv_vs);
    a.i1 = world_to_voxel_fast(dual + vs, vs, inv_vs);
  }
  return a;
}

// getPoint3d of pixel (u, v) at depth z with the focal lengths' reciprocals
// (the raycast's sample and hit positions: a few ulp from the quotients,
// only real-valued outputs depend on them)
__device__ __forceinline__ f3 point3d_rc(int u, int v, float z, const Intr &s, float ifx, float ify) {
  f3 p;
  p.z = z;
  p.x = (((float)u - s.cx) * z) * ifx;
  p.y = (((float)v - s.cy) * z) * ify;
  return p;
}

__device__ __forceinline__ bool interp_axes(const DmView &v, const AxisSample &X, const AxisSample &Y,
                                            const AxisSample &Z, BlockWin &W, float &sdf_out, float *col) {
  const int cx = X.c, cy = Y.c, cz = Z.c;
  const int ix[2] = {X.i0, X.i1};
  const int iy[2] = {Y.i0, Y.i1};
  const int iz[2] = {Z.i0, Z.i1};
  // block range of the centre and corner voxels per axis
  const int lox = min(min(ix[0], ix[1]), cx) >> 3, hix = max(max(ix[0], ix[1]), cx) >> 3;
  const int loy = min(min(iy[0], iy[1]), cy) >> 3, hiy = max(max(iy[0], iy[1]), cy) >> 3;
  const int loz = min(min(iz[0], iz[1]), cz) >> 3, hiz = max(max(iz[0], iz[1]), cz) >> 3;
  const f3 wt = mk(X.w, Y.w, Z.w);
  if (hix - lox > 1 || hiy - loy > 1 || hiz - loz > 1) {
    // voxels over more than two blocks on an axis: one lookup per voxel (a
    // rolled loop, same order and arithmetic as below)
    const uint64_t w0 = voxel_lookup(v, cx, cy, cz);
    if ((w0 >> 56) == 0) return false;
    float dist = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
      const int ox = (0xD2 >> k) & 1, oy = (0xB4 >> k) & 1, oz = (0xE8 >> k) & 1;
      const uint64_t w = voxel_lookup(v, ox ? ix[1] : ix[0], oy ? iy[1] : iy[0], oz ? iz[1] : iz[0]);
      const float a = ox ? wt.x : 1.0f - wt.x;
      const float b = oy ? wt.y : 1.0f - wt.y;
      const float g = oz ? wt.z : 1.0f - wt.z;
      const float coef = a * b * g;
      const uint64_t sw = (w >> 56) == 0 ? w0 : w;
      dist += coef * vox_sdf(sw);
      if (col) {
        const uint32_t rgbw = vox_rgbw(sw);
        c0 += coef * (float)(rgbw & 0xff);
        c1 += coef * (float)((rgbw >> 8) & 0xff);
        c2 += coef * (float)((rgbw >> 16) & 0xff);
      }
    }
    sdf_out = dist;
    if (col) { col[0] = c0; col[1] = c1; col[2] = c2; }
    return true;
  }
  if (lox < W.wx || hix > W.wx + 1 || loy < W.wy || hiy > W.wy + 1 || loz < W.wz || hiz > W.wz + 1) {
    W.wx = lox - (hix == lox ? (int)((W.resolved >> 8) & 1u) : 0);
    W.wy = loy - (hiy == loy ? (int)((W.resolved >> 9) & 1u) : 0);
    W.wz = loz - (hiz == loz ? (int)((W.resolved >> 10) & 1u) : 0);
    W.resolved &= 0x700u;
  }
  const int kx[2] = {(ix[0] >> 3) - W.wx, (ix[1] >> 3) - W.wx};
  const int ky[2] = {2 * ((iy[0] >> 3) - W.wy), 2 * ((iy[1] >> 3) - W.wy)};
  const int kz[2] = {4 * ((iz[0] >> 3) - W.wz), 4 * ((iz[1] >> 3) - W.wz)};
  const int kc = ((cx >> 3) - W.wx) + 2 * ((cy >> 3) - W.wy) + 4 * ((cz >> 3) - W.wz);
  // slots of the 8 corners: the product of the per-axis slot sets (x bits
  // 0-1, y spread to bits 0/2, z to bits 0/4), and the centre's slot
  const unsigned mx = (1u << kx[0]) | (1u << kx[1]);
  const unsigned my = (1u << (ky[0] >> 1)) | (1u << (ky[1] >> 1));
  const unsigned mz = (1u << (kz[0] >> 2)) | (1u << (kz[1] >> 2));
  const unsigned need = (mx * ((my & 1u) | ((my & 2u) << 1)) * ((mz & 1u) | ((mz & 2u) << 3))) | (1u << kc);
  unsigned todo = need & ~W.resolved;
  while (todo) {
    const int k = __builtin_ctz(todo);
    todo &= todo - 1u;
    const int ptr = find_block(v, W.wx + (k & 1), W.wy + ((k >> 1) & 1), W.wz + (k >> 2));
    W.slot[k * 256] = ptr < 0 ? v.num_blocks : ptr;  // unallocated: the pool's all-zero block
    W.resolved |= 1u << k;
  }
  const uint64_t *vox = v.vox;
  // voxel word of (gx, gy, gz) in window slot k, 0 when the block is unallocated
  auto word = [&](int k, int gx, int gy, int gz) -> uint64_t {
    const uint32_t vi = (uint32_t)W.slot[k * 256] * (uint32_t)kBlockVox +
                        (uint32_t)(mod8(gx) * 64 + mod8(gy) * 8 + mod8(gz));
    return ldg(vox + vi);
  };
  const uint64_t w0 = word(kc, cx, cy, cz);
  uint64_t wc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = (0xD2 >> k) & 1, oy = (0xB4 >> k) & 1, oz = (0xE8 >> k) & 1;
    wc[k] = word(kx[ox] + ky[oy] + kz[oz], ix[ox], iy[oy], iz[oz]);
  }
  if ((w0 >> 56) == 0) return false;
  float dist = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = (0xD2 >> k) & 1, oy = (0xB4 >> k) & 1, oz = (0xE8 >> k) & 1;
    const float a = ox ? wt.x : 1.0f - wt.x;
    const float b = oy ? wt.y : 1.0f - wt.y;
    const float g = oz ? wt.z : 1.0f - wt.z;
    const float coef = a * b * g;
    const uint64_t sw = (wc[k] >> 56) == 0 ? w0 : wc[k];
    dist += coef * vox_sdf(sw);
    if (col) {
      const uint32_t rgbw = vox_rgbw(sw);
      c0 += coef * (float)(rgbw & 0xff);
      c1 += coef * (float)((rgbw >> 8) & 0xff);
      c2 += coef * (float)((rgbw >> 16) & 0xff);
    }
  }
  sdf_out = dist;
  if (col) { col[0] = c0; col[1] = c1; col[2] = c2; }
  return true;
}

__device__ __forceinline__ bool interp_voxel_win(const DmView &v, f3 p, float inv_vs, BlockWin &W, float &sdf_out,
                                                 float *col) {
  const float vs = v.voxel_size;
  return interp_axes(v, axis_sample(p.x, vs, inv_vs), axis_sample(p.y, vs, inv_vs), axis_sample(p.z, vs, inv_vs), W,
                     sdf_out, col);
}

// The hit's seven interpolations (the colour sample at pw and the central
// differences at pw +- vs e_a) from one stencil, when every axis is regular:
// the axis samples of q - vs, q and q + vs have corners (i0 - 1, i0),
// (i0, i0 + 1), (i0 + 1, i0 + 2) and centres c - 1, c, c + 1 with c in
// {i0, i0 + 1}.  The samples then read the 2^3 core voxels (i0 + {0,1}, ...)
// and per axis the 4 voxels before and the 4 after it: 32 voxels in 2^3
// blocks at most.  Each sample's centre is its corner (c - i0, ...), the
// same index for all seven.  Corner order, weights and sums are those of
// interp_axes, so the results are bit for bit the same.
__device__ __forceinline__ bool stencil_regular(const AxisSample &A, const AxisSample &P, const AxisSample &M) {
  return A.i1 == A.i0 + 1 && P.i0 == A.i1 && P.i1 == A.i0 + 2 && M.i0 == A.i0 - 1 && M.i1 == A.i0 &&
         (unsigned)(A.c - A.i0) <= 1u && P.c == A.c + 1 && M.c == A.c - 1;
}
// value of corner ci (0..7) of f, without indexing registers
__device__ __forceinline__ float pick8(const float (&f)[8], int ci) {
  const float a = (ci & 1) ? f[1] : f[0], b = (ci & 1) ? f[3] : f[2];
  const float c = (ci & 1) ? f[5] : f[4], d = (ci & 1) ? f[7] : f[6];
  const float lo = (ci & 2) ? b : a, hi = (ci & 2) ? d : c;
  return (ci & 4) ? hi : lo;
}
// interp_axes' sum for one sample: corners f (index ox + 2 oy + 4 oz), bit
// o of obs set when corner o is observed, centre corner ci
__device__ __forceinline__ bool blend8(const float (&f)[8], uint32_t obs, int ci, f3 wt, float &out) {
  if (!((obs >> ci) & 1u)) return false;
  const float fc = pick8(f, ci);
  float dist = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = (0xD2 >> k) & 1, oy = (0xB4 >> k) & 1, oz = (0xE8 >> k) & 1;
    const int o = ox + 2 * oy + 4 * oz;
    const float a = ox ? wt.x : 1.0f - wt.x;
    const float b = oy ? wt.y : 1.0f - wt.y;
    const float g = oz ? wt.z : 1.0f - wt.z;
    const float coef = a * b * g;
    dist += coef * (((obs >> o) & 1u) ? f[o] : fc);
  }
  out = dist;
  return true;
}
__device__ __forceinline__ void hit_stencil(const DmView &v, const AxisSample &X0, const AxisSample &Y0,
                                            const AxisSample &Z0, const float (&wp)[3], const float (&wm)[3],
                                            BlockWin &W, float (&grad)[3], bool &ok, uint32_t (&rgb)[3],
                                            bool &has_rgb) {
  const int i0 = X0.i0, j0 = Y0.i0, k0 = Z0.i0;
  // blocks of voxels i0 - 1 .. i0 + 2 per axis (two at most)
  const int lox = (i0 - 1) >> 3, hix = (i0 + 2) >> 3, loy = (j0 - 1) >> 3, hiy = (j0 + 2) >> 3,
            loz = (k0 - 1) >> 3, hiz = (k0 + 2) >> 3;
  if (lox < W.wx || hix > W.wx + 1 || loy < W.wy || hiy > W.wy + 1 || loz < W.wz || hiz > W.wz + 1) {
    W.wx = lox - (hix == lox ? (int)((W.resolved >> 8) & 1u) : 0);
    W.wy = loy - (hiy == loy ? (int)((W.resolved >> 9) & 1u) : 0);
    W.wz = loz - (hiz == loz ? (int)((W.resolved >> 10) & 1u) : 0);
    W.resolved &= 0x700u;
  }
  // slots of the block ranges: x bits 0-1, y spread to bits 0/2, z to 0/4
  const unsigned mx = (1u << (lox - W.wx)) | (1u << (hix - W.wx));
  const unsigned my = (1u << (loy - W.wy)) | (1u << (hiy - W.wy));
  const unsigned mz = (1u << (loz - W.wz)) | (1u << (hiz - W.wz));
  const unsigned need = mx * ((my & 1u) | ((my & 2u) << 1)) * ((mz & 1u) | ((mz & 2u) << 3));
  unsigned todo = need & ~W.resolved;
  while (todo) {
    const int k = __builtin_ctz(todo);
    todo &= todo - 1u;
    const int ptr = find_block(v, W.wx + (k & 1), W.wy + ((k >> 1) & 1), W.wz + (k >> 2));
    W.slot[k * 256] = ptr < 0 ? v.num_blocks : ptr;
    W.resolved |= 1u << k;
  }
  const uint64_t *vox = v.vox;
  auto word = [&](int gx, int gy, int gz) -> uint64_t {
    const int k = ((gx >> 3) - W.wx) + 2 * ((gy >> 3) - W.wy) + 4 * ((gz >> 3) - W.wz);
    const uint32_t vi = (uint32_t)W.slot[k * 256] * (uint32_t)kBlockVox +
                        (uint32_t)(mod8(gx) * 64 + mod8(gy) * 8 + mod8(gz));
    return ldg(vox + vi);
  };
  const int ci = (X0.c - i0) + 2 * (Y0.c - j0) + 4 * (Z0.c - k0);
  float f[8];
  uint32_t obs = 0u;
  {
    uint64_t wc[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) wc[o] = word(i0 + (o & 1), j0 + ((o >> 1) & 1), k0 + (o >> 2));
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      f[o] = vox_sdf(wc[o]);
      obs |= ((wc[o] >> 56) != 0 ? 1u : 0u) << o;
    }
    // colour sample: interp_axes' colour sums with the centre standing in
    // for unobserved corners
    has_rgb = (obs >> ci) & 1u;
    if (has_rgb) {
      const uint32_t cc = vox_rgbw(ci & 4 ? ((ci & 2) ? ((ci & 1) ? wc[7] : wc[6]) : ((ci & 1) ? wc[5] : wc[4]))
                                          : ((ci & 2) ? ((ci & 1) ? wc[3] : wc[2]) : ((ci & 1) ? wc[1] : wc[0])));
      float c0 = 0.f, c1 = 0.f, c2 = 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ox = (0xD2 >> k) & 1, oy = (0xB4 >> k) & 1, oz = (0xE8 >> k) & 1;
        const int o = ox + 2 * oy + 4 * oz;
        const float a = ox ? X0.w : 1.0f - X0.w;
        const float b = oy ? Y0.w : 1.0f - Y0.w;
        const float g = oz ? Z0.w : 1.0f - Z0.w;
        const float coef = a * b * g;
        const uint32_t rgbw = ((obs >> o) & 1u) ? vox_rgbw(wc[o]) : cc;
        c0 += coef * (float)(rgbw & 0xff);
        c1 += coef * (float)((rgbw >> 8) & 0xff);
        c2 += coef * (float)((rgbw >> 16) & 0xff);
      }
      rgb[0] = to_u8(c0); rgb[1] = to_u8(c1); rgb[2] = to_u8(c2);
    }
  }
  ok = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    float sides[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      // one side's 4 loads at a time (keeps the other words out of registers)
      __builtin_amdgcn_sched_barrier(0);
      // side 0: the - sample, whose corners along the axis are the 4 voxels
      // before the core (index i - 1) and then the core's lower layer; side 1:
      // the + sample, the core's upper layer and then the 4 voxels after it
      // (index i + 2).  t = a1 + 2 a2 numbers the other two axes' bits in x,
      // y, z order.
      float fe[4];
      uint32_t oe = 0u;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int a1 = t & 1, a2 = t >> 1;
        const int d = side ? 2 : -1;
        const int gx = ax == 0 ? i0 + d : i0 + a1, gy = ax == 1 ? j0 + d : j0 + (ax == 0 ? a1 : a2),
                  gz = ax == 2 ? k0 + d : k0 + a2;
        const uint64_t w = word(gx, gy, gz);
        fe[t] = vox_sdf(w);
        oe |= ((w >> 56) != 0 ? 1u : 0u) << t;
      }
      float F[8];
      uint32_t O = 0u;
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int oa = (o >> ax) & 1;
        const int t = ax == 0 ? ((o >> 1) & 1) + 2 * (o >> 2)
                      : ax == 1 ? (o & 1) + 2 * (o >> 2) : (o & 1) + 2 * ((o >> 1) & 1);
        const bool core = side ? !oa : oa;  // the core voxel on this corner, else the extra one
        const int oc = side ? (o | (1 << ax)) : (o & ~(1 << ax));
        F[o] = core ? f[oc] : fe[t];
        O |= (core ? ((obs >> oc) & 1u) : ((oe >> t) & 1u)) << o;
      }
      const float wa = side ? wp[ax] : wm[ax];
      const f3 wt = mk(ax == 0 ? wa : X0.w, ax == 1 ? wa : Y0.w, ax == 2 ? wa : Z0.w);
      float sv = 0.f;
      ok &= blend8(F, O, ci, wt, sv);
      sides[side] = sv;
    }
    grad[ax] = sides[1] - sides[0];
  }
}

// Empty-space test of the raycast.  A sample is "far" when every block that
// can hold one of the voxels GetInterpolatedVoxel reads at it is unallocated
// or has no observed voxel below the far threshold: the sample is then
// unobserved or interpolates to at least 0.5*trunc, and the march takes its
// full step without interpolating.  The test uses an approximate sample
// position (ray origin + z * direction, no divisions) and a voxel range
// widened by a margin far larger than its error, so it is conservative: a
// sample it calls far is far for the exact evaluation too.
// Far test of one block for the precise path: the window's near bit, or
// the hash table outside the window.
__device__ __forceinline__ bool block_far_any(const DmView &v, int bx, int by, int bz) {
  uint32_t idx;
  if (win_bit(bx, by, bz, idx)) return !test_bit(v.nearmap, idx);
  const int ptr = find_block(v, bx, by, bz);
  return ptr < 0 || ldg(v.blk_near + ptr) == 0;
}

// Conservative empty-space run.  pv = (origin + z * direction) / vs is the
// approximate sample position in voxel units (error about 1e-3 voxel); the
// voxels GetInterpolatedVoxel reads at the exact position lie in
// [floor(pv - e) - 1, floor(pv + 1 + e)] per axis (the exact indices round
// p/vs -+ 0.5 half away from zero).  B is the block holding the approximate
// centre voxel; it lies in that range.
//  - B's dilated bit is clear (no near block among its 3^3 neighbourhood):
//    the run extends over the largest clear region holding B: its coarse
//    cell (8^3 blocks) with a one-block margin, else its 4^3-block cell with
//    that margin, else the neighbourhood.  A cell's bit is clear only when no
//    block of it has its dilated bit set, so the cell and its margin hold no
//    near block.  A set dilated bit implies set cell bits, so the cells are
//    only tested here.
//  - otherwise B itself near ends the test (B is in the voxel range), and
//    else every block of the voxel range must be far (near bit clear, or
//    unallocated / not near through the hash table outside the window), and
//    the run extends over those blocks.
// Runs end e + 1 voxel inside the region, so the exact positions up to zmax
// keep their voxel range inside it.  Which region a run takes changes how
// many tests the march makes, never its samples.
__device__ __forceinline__ bool far_run(const DmView &v, f3 po, f3 pd, float z, float &zmax) {
  const float e = 0.05f;
  const float px = po.x + pd.x * z, py = po.y + pd.y * z, pz = po.z + pd.z * z;
  if (!(fabsf(px) < 1.0e6f && fabsf(py) < 1.0e6f && fabsf(pz) < 1.0e6f)) return false;
  const int bx = floordiv8((int)floorf(px + 0.5f)), by = floordiv8((int)floorf(py + 0.5f)),
            bz = floordiv8((int)floorf(pz + 0.5f));
  int x0, x1, y0, y1, z0, z1;
  uint32_t idx;
  const bool bwin = win_bit(bx, by, bz, idx);
  if (bwin && !test_bit(v.dilmap, idx)) {
    int cw, cb;
    if (coarse_bit(coarse_of(bx), coarse_of(by), coarse_of(bz), cw, cb) && !(ldg(v.coarse + cw) & (1u << cb))) {
      const int cx = coarse_of(bx), cy = coarse_of(by), cz = coarse_of(bz);
      x0 = 8 * cx - 1; x1 = 8 * cx + 8; y0 = 8 * cy - 1; y1 = 8 * cy + 8; z0 = 8 * cz - 1; z1 = 8 * cz + 8;
    } else if (cell_bit<2>(cell_of<2>(bx), cell_of<2>(by), cell_of<2>(bz), cw, cb) &&
               !(ldg(v.cell4 + cw) & (1u << cb))) {
      const int cx = cell_of<2>(bx), cy = cell_of<2>(by), cz = cell_of<2>(bz);
      x0 = 4 * cx - 1; x1 = 4 * cx + 4; y0 = 4 * cy - 1; y1 = 4 * cy + 4; z0 = 4 * cz - 1; z1 = 4 * cz + 4;
    } else {
      x0 = bx - 1; x1 = bx + 1; y0 = by - 1; y1 = by + 1; z0 = bz - 1; z1 = bz + 1;
    }
  } else {
    if (bwin && test_bit(v.nearmap, idx)) return false;
    x0 = floordiv8((int)floorf(px - e) - 1); x1 = floordiv8((int)floorf(px + 1.0f + e));
    y0 = floordiv8((int)floorf(py - e) - 1); y1 = floordiv8((int)floorf(py + 1.0f + e));
    z0 = floordiv8((int)floorf(pz - e) - 1); z1 = floordiv8((int)floorf(pz + 1.0f + e));
    // x-adjacent blocks are adjacent bits of the window: one word (two at a
    // word boundary) covers the x range of a (y, z) row; blocks outside the
    // window go through the hash table
    const bool xin = x0 >= -kWin / 2 && x1 < kWin / 2 && x1 - x0 <= 1;
    for (int qz = z0; qz <= z1; ++qz)
      for (int qy = y0; qy <= y1; ++qy) {
        uint32_t i0;
        if (xin && win_bit(x0, qy, qz, i0)) {
          const uint32_t i1 = i0 + (uint32_t)(x1 - x0);
          const bool split = (i1 >> 5) != (i0 >> 5);
          uint32_t hit = (ldg(v.nearmap + (i0 >> 5)) >> (i0 & 31)) & (split ? 1u : (2u << (i1 - i0)) - 1u);
          if (split) hit |= ldg(v.nearmap + (i1 >> 5)) & 1u;
          if (hit) return false;
        } else {
          for (int qx = x0; qx <= x1; ++qx)
            if (!block_far_any(v, qx, qy, qz)) return false;
        }
      }
  }
  float zm = 3.0e38f;
  // the hardware reciprocal (1 ulp): zmax only needs to be accurate to well
  // within e, and the direction is recomputed here rather than kept live
  const f3 ipd = mk(__builtin_amdgcn_rcpf(pd.x), __builtin_amdgcn_rcpf(pd.y), __builtin_amdgcn_rcpf(pd.z));
  const float m0 = 1.0f + 2.0f * e, m1 = 7.0f - 2.0f * e;
  // reciprocal direction: zmax only needs to be accurate to well within e
  if (pd.x > 0.f) zm = fminf(zm, ((float)(8 * x1) + m1 - po.x) * ipd.x);
  else if (pd.x < 0.f) zm = fminf(zm, ((float)(8 * x0) + m0 - po.x) * ipd.x);
  if (pd.y > 0.f) zm = fminf(zm, ((float)(8 * y1) + m1 - po.y) * ipd.y);
  else if (pd.y < 0.f) zm = fminf(zm, ((float)(8 * y0) + m0 - po.y) * ipd.y);
  if (pd.z > 0.f) zm = fminf(zm, ((float)(8 * z1) + m1 - po.z) * ipd.z);
  else if (pd.z < 0.f) zm = fminf(zm, ((float)(8 * z0) + m0 - po.z) * ipd.z);
  zmax = zm;
  return true;
}

__global__ __launch_bounds__(256, 6) void k_raycast(const DmView *__restrict__ views, const int *__restrict__ active,
                                                        int tw_log2) {
  const int slot_ = active[blockIdx.y];
  if (slot_ < 0) return;
  const DmView &v = views[slot_];
  // tile of tw x th pixels (tw * th = 256), shape chosen by the autotuner
  const int tw = 1 << tw_log2, th = 256 >> tw_log2;
  const int tiles_x = (v.cols + tw - 1) >> tw_log2;
  const int tile = blockIdx.x;
  if (tile >= tiles_x * ((v.rows + th - 1) / th)) return;
  // lanes of a wavefront: one 8x8 quadrant of a 16x16 tile (rays closer
  // together take more alike marches than a 16x4 strip's), else tw x 64/tw
  int lx = threadIdx.x & (tw - 1), ly = threadIdx.x >> tw_log2;
  if (tw_log2 == 4) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    lx = (l & 7) + 8 * (w & 1);
    ly = (l >> 3) + 8 * (w >> 1);
  }
  const int u = (tile % tiles_x) * tw + lx;
  const int r = (tile / tiles_x) * th + ly;
  if (u >= v.cols || r >= v.rows) return;
  const int i = r * v.cols + u;
  const Intr s = intr_of(v);
  const M4 T = load_m4(v.posef + PF_T);
  const float vs = v.voxel_size, tau = v.trunc;
  const float big = 0.5f * tau, small = 0.5f * vs;
  const int max_steps = (int)ceil((double)((v.max_depth - v.min_depth) / small)) + 2;
  const int it_last = max_steps - 1;
  const float zpre_max = nextafterf(v.max_depth, -INFINITY);  // largest float below max_depth
  const float inv_vs = 1.0f / vs;
  const float ifx = 1.0f / s.fx, ify = 1.0f / s.fy;
  // ray in voxel units for the approximate empty-space test
  const f3 po = mk(T.m[3], T.m[7], T.m[11]) * inv_vs;
  const f3 pd = rot(T, mk(((float)u - s.cx) * ifx, ((float)r - s.cy) * ify, 1.0f)) * inv_vs;
  __shared__ int win_lds[8 * 256];
  BlockWin c;
  c.wx = 0x3fffffff; c.wy = 0; c.wz = 0;
  c.resolved = (pd.x < 0.f ? 0x100u : 0u) | (pd.y < 0.f ? 0x200u : 0u) | (pd.z < 0.f ? 0x400u : 0u);
  c.slot = win_lds + threadIdx.x;
  float z = v.min_depth, pz = 0.f, ps = 0.f, zh = 0.f, zprev = 0.f;
  bool pobs = false, hit = false, prev_fast = false;
  for (int it = 0; it < max_steps && z < v.max_depth; ++it) {
    float sdf, zmax;
    // far samples: the step is 0.5*trunc whether they are observed or not;
    // the run of far samples up to zmax is stepped over without lookups (the
    // last one's interpolation is recomputed below if the next sample needs it)
    if (far_run(v, po, pd, z, zmax)) {
      prev_fast = true;
      zprev = z;
      z = z + big;
      const float zlim = fminf(zmax, zpre_max);  // z < max_depth <=> z <= zpre_max
      // while (z <= zlim && it < it_last) { ++it; zprev = z; z = z + big; }, four
      // steps at a time: the sums increase, so the fourth step's start within
      // zlim admits all four
      while (it + 4 <= it_last) {
        const float z1 = z + big, z2 = z1 + big, z3 = z2 + big;
        if (!(z3 <= zlim)) break;
        it += 4;
        zprev = z3;
        z = z3 + big;
      }
      while (z <= zlim && it < it_last) {
        ++it;
        zprev = z;
        z = z + big;
      }
      continue;
    }
    const f3 p = xform(T, point3d_rc(u, r, z, s, ifx, ify));
    if (!interp_voxel_win(v, p, inv_vs, c, sdf, nullptr)) { pobs = false; prev_fast = false; zprev = z; z = z + big; continue; }
    if (sdf <= 0.f) {
      if (prev_fast) {
        float s2;
        pobs = interp_voxel_win(v, xform(T, point3d_rc(u, r, zprev, s, ifx, ify)), inv_vs, c, s2, nullptr);
        ps = s2;
        pz = zprev;
      }
      if (pobs && ps > 0.f) { zh = pz + (z - pz) * (ps * __builtin_amdgcn_rcpf(ps - sdf)); hit = true; }
      break;
    }
    pobs = true; ps = sdf; pz = z; prev_fast = false;
    float st = sdf < big ? sdf : big;
    if (st < small) st = small;
    zprev = z;
    z = z + st;
  }
  float4 vert = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 nrm = make_float4(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000), __int_as_float(0x7fc00000), 0.f);
  uint32_t cr = 0, cg = 0, cb = 0;
  float dout = 0.f;
  if (hit) {
    const f3 pc = point3d_rc(u, r, zh, s, ifx, ify);
    const f3 pw = xform(T, pc);
    dout = zh;
    // The six central-difference samples pw +- vs e_a and the colour sample
    // at pw.  The axes are independent in the interpolation, and an axis
    // without offset (pw.y + 0.0f for the x samples) gives the same axis
    // sample as pw itself, so 9 axis samples serve the 7 interpolations.
    // One (not unrolled) copy of the gather keeps the kernel's code small.
    const AxisSample X0 = axis_sample(pw.x, vs, inv_vs), Y0 = axis_sample(pw.y, vs, inv_vs),
                     Z0 = axis_sample(pw.z, vs, inv_vs);
    float wp[3], wm[3];
    bool reg = true;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const float q = ax == 0 ? pw.x : ax == 1 ? pw.y : pw.z;
      const AxisSample P = axis_sample(q + vs, vs, inv_vs), M = axis_sample(q - vs, vs, inv_vs);
      reg &= stencil_regular(ax == 0 ? X0 : ax == 1 ? Y0 : Z0, P, M);
      wp[ax] = P.w;
      wm[ax] = M.w;
    }
    float gx = 0.f, gy = 0.f, gz = 0.f;
    bool ok = true;
    if (reg) {
      float g[3];
      uint32_t rgb[3];
      bool has_rgb;
      hit_stencil(v, X0, Y0, Z0, wp, wm, c, g, ok, rgb, has_rgb);
      gx = g[0]; gy = g[1]; gz = g[2];
      if (has_rgb) { cr = rgb[0]; cg = rgb[1]; cb = rgb[2]; }
    } else {
#pragma unroll 1
      for (int ax = 0; ax < 3; ++ax) {
        const float q = ax == 0 ? pw.x : ax == 1 ? pw.y : pw.z;
        const AxisSample P = axis_sample(q + vs, vs, inv_vs), M = axis_sample(q - vs, vs, inv_vs);
        float sp, sm;
        ok &= interp_axes(v, ax == 0 ? P : X0, ax == 1 ? P : Y0, ax == 2 ? P : Z0, c, sp, nullptr);
        ok &= interp_axes(v, ax == 0 ? M : X0, ax == 1 ? M : Y0, ax == 2 ? M : Z0, c, sm, nullptr);
        const float g = sp - sm;
        if (ax == 0) gx = g; else if (ax == 1) gy = g; else gz = g;
      }
      float sv, col[3];
      if (interp_axes(v, X0, Y0, Z0, c, sv, col)) { cr = to_u8(col[0]); cg = to_u8(col[1]); cb = to_u8(col[2]); }
    }
    if (ok) {
      const M4 Ti = load_m4(v.posef + PF_TINV);
      // unit gradient through the hardware reciprocal square root (a zero
      // gradient gives NaN, as the exact quotient does)
      const float rn = __builtin_amdgcn_rsqf(gx * gx + gy * gy + gz * gz);
      const f3 n = rot(Ti, mk(gx * rn, gy * rn, gz * rn));
      nrm = make_float4(n.x, n.y, n.z, 0.f);
    }
    vert = make_float4(pc.x, pc.y, pc.z, intensity_of(cr, cg, cb));
  }
  v.tvert[0][i] = vert;
  v.tnorm[0][i] = nrm;
  if (v.ray_depth) v.ray_depth[i] = dout;
  if (v.ray_rgb) {
    uint8_t *o = v.ray_rgb + 3 * (size_t)i;
    o[0] = (uint8_t)cr; o[1] = (uint8_t)cg; o[2] = (uint8_t)cb;
  }
  // coarser model levels by nearest subsampling
  if (((u | r) & 1) == 0 && (u >> 1) < (v.cols >> 1) && (r >> 1) < (v.rows >> 1)) {
    const int j = (r >> 1) * (v.cols >> 1) + (u >> 1);
    v.tvert[1][j] = vert;
    v.tnorm[1][j] = nrm;
  }
  if (((u | r) & 3) == 0 && (u >> 2) < (v.cols >> 2) && (r >> 2) < (v.rows >> 2)) {
    const int j = (r >> 2) * (v.cols >> 2) + (u >> 2);
    v.tvert[2][j] = vert;
    v.tnorm[2][j] = nrm;
  }
  if (i == 0) v.ctr[C_HAS_MODEL] = 1;
}

// ---------------------------------------------------------------------------
// launchers
static constexpr unsigned kListWG = 64;  // workgroups per session for the new-key passes

void launch_alloc(const Launch &L) {
  if (L.nact <= 0) return;
  hipLaunchKernelGGL(k_alloc_rays<false>, dim3((unsigned)L.max_tiles, (unsigned)L.nact), dim3(256), 0, L.stream,
                     L.views, L.active);
  hipLaunchKernelGGL(k_alloc_link, dim3(kListWG, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_alloc_rank, dim3(kListWG, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_alloc_commit, dim3(kListWG, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_alloc_finish, dim3(kListWG, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_alloc_close, dim3((unsigned)L.nact), dim3(64), 0, L.stream, L.views, L.active);
}

void launch_alloc_rays(const Launch &L, bool count_unique) {
  if (L.nact <= 0) return;
  if (count_unique)
    hipLaunchKernelGGL(k_alloc_rays<true>, dim3((unsigned)L.max_tiles, (unsigned)L.nact), dim3(256), 0, L.stream,
                       L.views, L.active);
  else
    hipLaunchKernelGGL(k_alloc_rays<false>, dim3((unsigned)L.max_tiles, (unsigned)L.nact), dim3(256), 0, L.stream,
                       L.views, L.active);
}

void launch_integrate(const Launch &L) {
  if (L.nact <= 0) return;
  hipLaunchKernelGGL(k_integrate, dim3((unsigned)L.integ_wg, (unsigned)L.nact), dim3(256), 0, L.stream, L.views,
                     L.active);
  hipLaunchKernelGGL(k_dilate_near, dim3(256, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
}

void rebuild_far_maps(const Launch &L) {
  if (L.nact <= 0) return;
  hipLaunchKernelGGL(k_clear_far_maps, dim3(1024, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_dilate_near, dim3(256, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
  hipLaunchKernelGGL(k_mark_allocated, dim3(256, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active);
}

void launch_delete(const DmView *view, const int4 *req, int n, unsigned long long *dmask, int *touched, int *counts,
                   hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_del_mark, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, view, req, n, dmask, touched,
                     counts);
  hipLaunchKernelGGL(k_del_compact, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, view, dmask, touched,
                     counts);
  hipLaunchKernelGGL(k_del_clear, dim3(1024), dim3(256), 0, st, view, counts);
  hipLaunchKernelGGL(k_del_close, dim3(1), dim3(64), 0, st, view, counts);
}

void launch_select_far(const DmView *view, const double c[3], double radius, int4 *req, int *nreq, hipStream_t st) {
  hipLaunchKernelGGL(k_select_far, dim3(1024), dim3(256), 0, st, view, c[0], c[1], c[2], radius * radius, req, nreq);
}

void launch_raycast(const Launch &L) {
  if (L.nact <= 0) return;
  const int tw = 1 << L.ray_tw_log2, th = 256 >> L.ray_tw_log2;
  const unsigned tiles = (unsigned)(((L.max_cols + tw - 1) / tw) * ((L.max_rows + th - 1) / th));
  hipLaunchKernelGGL(k_raycast, dim3(tiles, (unsigned)L.nact), dim3(256), 0, L.stream, L.views, L.active,
                     L.ray_tw_log2);
}

}  // namespace dm
