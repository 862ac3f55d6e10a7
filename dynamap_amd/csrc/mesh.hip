// Marching cubes over the allocated blocks of a volume
// (include/geometry/mesh/meshExtractor/meshExtractor.cu:22-242, 364-402).
//
// Output order is canonical: blocks in hash (bucket, slot) order, the 512
// cells of a block in x-major local order, the triangles of a cell in table
// order -- the order of the CPU restatement, independent of scheduling.
//
// Passes (all on the volume's stream):
//   1. gather the per-bucket occupancy counts, exclusive-scan them, and write
//      the pool index of every occupied slot at its canonical position;
//   2. k_mc_count: one workgroup per block, one thread per cell.  The block's
//      voxels plus a one-voxel halo (10^3 sdf values and observed flags) are
//      staged in LDS; each cell takes its 8 strict trilinear corner samples
//      from there, stores its cube index (0 = no triangles) and the
//      workgroup sums the triangle count of the block;
//   3. exclusive scan of the per-block counts;
//   4. k_mc_emit: blocks with triangles rescan their cells' counts in LDS,
//      recompute the corner samples of active cells, and write positions,
//      colours and edge ids at their canonical offset.
// Cell rules (K): cell centre g * vs, corners at +-vs/2, strict trilinear
// samples (any corner voxel of weight 0 gives no sample), the reference's
// corner-bit convention, Bourke's tables, the 1e-5 snaps of
// VertexInterpolation, the centre voxel's colour.  The float operations are
// those of the restatement in the same order, so positions are bit-identical.
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include "dm_device.h"
#include "dm_session.h"
#include "mc_table.h"
#include "dm_scan.h"

namespace dm {

__constant__ uint64_t c_mc_tri[256] = {DM_MC_TRI_TABLE};
static const uint64_t h_mc_tri[256] = {DM_MC_TRI_TABLE};

// corner index ci = 4a + 2b + c for the offset (a, b, c) in {-, +}^3; edge
// endpoints and the cube-index bit of each corner (meshExtractor.cu:181-189)
__constant__ int8_t c_edge_a[12] = {2, 6, 4, 0, 3, 7, 5, 1, 2, 6, 4, 0};
__constant__ int8_t c_edge_b[12] = {6, 4, 0, 2, 7, 5, 1, 3, 3, 7, 5, 1};
__constant__ int8_t c_edge_axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
__constant__ uint8_t c_corner_bit[8] = {8, 128, 1, 16, 4, 64, 2, 32};

constexpr int kHalo = 10;         // block edge + one voxel each side
constexpr int kCellThreads = 512;

__global__ __launch_bounds__(256) void k_gather_counts(const DmView *__restrict__ vp, int *__restrict__ counts) {
  const DmView &v = *vp;
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b < v.num_buckets) counts[b] = bucket_count(v, bucket_ptr(v, b));
}

// pool index of every occupied slot at its (bucket, slot) rank
__global__ __launch_bounds__(256) void k_block_list(const DmView *__restrict__ vp, const long long *__restrict__ offs,
                                                   int *__restrict__ list) {
  const DmView &v = *vp;
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= v.num_buckets) return;
  const uint8_t *bk = bucket_ptr(v, b);
  const int cnt = bucket_count(v, bk);
  const int32_t *ptrs = (const int32_t *)(bk + 8 * v.bucket_size);
  const long long o = offs[b];
  for (int s = 0; s < cnt; ++s) list[o + s] = ptrs[s];
}

// ---------------------------------------------------------------- cells
struct Halo {
  float sdf[kHalo * kHalo * kHalo];
  uint8_t obs[kHalo * kHalo * kHalo];
  int nb[27];
};

// Stage the block's voxels and a one-voxel halo.  All threads call it.
__device__ void load_halo(const DmView &v, int kx, int ky, int kz, Halo &H) {
  const int t = threadIdx.x;
  if (t < 27) H.nb[t] = find_block(v, kx + t % 3 - 1, ky + (t / 3) % 3 - 1, kz + t / 9 - 1);
  __syncthreads();
  for (int h = t; h < kHalo * kHalo * kHalo; h += blockDim.x) {
    const int lx = h / 100 - 1, ly = (h / 10) % 10 - 1, lz = h % 10 - 1;
    const int ox = lx < 0 ? 0 : (lx > 7 ? 2 : 1), oy = ly < 0 ? 0 : (ly > 7 ? 2 : 1), oz = lz < 0 ? 0 : (lz > 7 ? 2 : 1);
    const int ptr = H.nb[ox + 3 * oy + 9 * oz];
    uint64_t w = 0;
    if (ptr >= 0) w = v.vox[(size_t)ptr * kBlockVox + ((lx + 8) & 7) * 64 + ((ly + 8) & 7) * 8 + ((lz + 8) & 7)];
    H.sdf[h] = vox_sdf(w);
    H.obs[h] = (uint8_t)((w >> 56) != 0);
  }
  __syncthreads();
}

// Voxel at global index g: from the halo when inside it, else from the table.
__device__ __forceinline__ bool halo_voxel(const DmView &v, const Halo &H, int kx, int ky, int kz, int gx, int gy,
                                           int gz, float &sdf) {
  const int hx = gx - 8 * kx + 1, hy = gy - 8 * ky + 1, hz = gz - 8 * kz + 1;
  if ((unsigned)hx < kHalo && (unsigned)hy < kHalo && (unsigned)hz < kHalo) {
    const int h = hx * 100 + hy * 10 + hz;
    sdf = H.sdf[h];
    return H.obs[h] != 0;
  }
  BlockCache bc;
  bc.init();
  const uint64_t w = voxel_at(v, gx, gy, gz, bc);
  sdf = vox_sdf(w);
  return (w >> 56) != 0;
}

struct BoxArg {
  int use;
  double lo[3], hi[3];
};

// Corner positions and samples of cell (gx, gy, gz); false when the cell is
// outside the box or a sample is missing.  Each sample is
// TrilinearInterpolation (meshExtractor.cu:22-93), strict: false when any of
// its 8 voxels is unobserved.  The corner samples sit at c +- vs/2 per axis,
// and every per-axis quantity of the interpolation (the voxel indices of
// dual = p - vs/2 and dual + vs, the weight p/vs - floor(p/vs)) depends only
// on the axis and the sign, so it is evaluated once per (axis, sign) and
// combined: the same arithmetic as interpolating each corner on its own.
struct AxisCorner {
  float p;     // corner coordinate c +- vs/2
  int i0, i1;  // voxel indices of dual + 0 and dual + vs
  float w;     // trilinear weight
};
__device__ __forceinline__ AxisCorner axis_corner(float c, float off, float vs) {
  AxisCorner a;
  a.p = c + off;
  const float dual = a.p - vs / 2.f;
  a.i0 = world_to_voxel(dual + 0.f, vs);
  a.i1 = world_to_voxel(dual + vs, vs);
  const float vp = a.p / vs;
  a.w = vp - floorf(vp);
  return a;
}

__device__ __forceinline__ bool halo_axis_ok(const AxisCorner *A, int k) {
  const int lo = 8 * k - 1, hi = 8 * k + 8;
  return A[0].i0 >= lo && A[0].i1 >= lo && A[1].i0 >= lo && A[1].i1 >= lo && A[0].i0 <= hi && A[0].i1 <= hi &&
         A[1].i0 <= hi && A[1].i1 <= hi;
}

// The 8 corner samples of a cell (kWant: positions and values, else only
// the cube index).  Each voxel comes from the block's LDS halo; a cell whose
// indices leave the halo (only through rounding at huge coordinates) reads
// them through the hash table in a rolled loop.
template <bool kWant>
__device__ __forceinline__ bool cell_eval(const DmView &v, const Halo &H, int kx, int ky, int kz, int gx, int gy,
                                          int gz, const BoxArg &box, f3 *pc, float *d, int &cube) {
  const float vs = v.voxel_size;
  const f3 c = mk((float)gx * vs + 0.f, (float)gy * vs + 0.f, (float)gz * vs + 0.f);
  if (box.use && ((double)c.x < box.lo[0] || (double)c.y < box.lo[1] || (double)c.z < box.lo[2] ||
                  (double)c.x > box.hi[0] || (double)c.y > box.hi[1] || (double)c.z > box.hi[2]))
    return false;
  const float P = vs / 2.f, M = -P;
  const AxisCorner X[2] = {axis_corner(c.x, M, vs), axis_corner(c.x, P, vs)};
  const AxisCorner Y[2] = {axis_corner(c.y, M, vs), axis_corner(c.y, P, vs)};
  const AxisCorner Z[2] = {axis_corner(c.z, M, vs), axis_corner(c.z, P, vs)};
  // corner c: offsets (cx, cy, cz) = (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (0,1,1) (1,0,1) (1,1,1)
  constexpr uint8_t cxm = 0xD2, cym = 0xB4, czm = 0xE8;
  cube = 0;
  if (halo_axis_ok(X, kx) && halo_axis_ok(Y, ky) && halo_axis_ok(Z, kz)) {
    const int bx = 1 - 8 * kx, by = 1 - 8 * ky, bz = 1 - 8 * kz;
#pragma unroll
    for (int ci = 0; ci < 8; ++ci) {
      const AxisCorner &ax = X[(ci >> 2) & 1], &ay = Y[(ci >> 1) & 1], &az = Z[ci & 1];
      float dd = 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const bool ox = (cxm >> k) & 1, oy = (cym >> k) & 1, oz = (czm >> k) & 1;
        const int h = ((ox ? ax.i1 : ax.i0) + bx) * 100 + ((oy ? ay.i1 : ay.i0) + by) * 10 + (oz ? az.i1 : az.i0) + bz;
        if (!H.obs[h]) return false;
        const float a = ox ? ax.w : 1.f - ax.w;
        const float b = oy ? ay.w : 1.f - ay.w;
        const float g = oz ? az.w : 1.f - az.w;
        dd += a * b * g * H.sdf[h];
      }
      if (kWant) { pc[ci] = mk(ax.p, ay.p, az.p); d[ci] = dd; }
      if (dd < 0.f) cube += c_corner_bit[ci];
    }
    return true;
  }
#pragma unroll 1
  for (int ci = 0; ci < 8; ++ci) {
    const AxisCorner &ax = X[(ci >> 2) & 1], &ay = Y[(ci >> 1) & 1], &az = Z[ci & 1];
    float dd = 0.f;
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
      const bool ox = (cxm >> k) & 1, oy = (cym >> k) & 1, oz = (czm >> k) & 1;
      float sv;
      if (!halo_voxel(v, H, kx, ky, kz, ox ? ax.i1 : ax.i0, oy ? ay.i1 : ay.i0, oz ? az.i1 : az.i0, sv)) return false;
      const float a = ox ? ax.w : 1.f - ax.w;
      const float b = oy ? ay.w : 1.f - ay.w;
      const float g = oz ? az.w : 1.f - az.w;
      dd += a * b * g * sv;
    }
    if (kWant) { pc[ci] = mk(ax.p, ay.p, az.p); d[ci] = dd; }
    if (dd < 0.f) cube += c_corner_bit[ci];
  }
  return true;
}


__global__ __launch_bounds__(kCellThreads) void k_mc_count(const DmView *__restrict__ vp, const int *__restrict__ list,
                                                           BoxArg box, uint8_t *__restrict__ cubes,
                                                           int *__restrict__ block_tri) {
  __shared__ Halo H;
  __shared__ int red[kCellThreads / 64 + 1];
  const DmView &v = *vp;
  const int ptr = list[blockIdx.x];
  const int4 key = v.block_key[ptr];
  load_halo(v, key.x, key.y, key.z, H);
  const int li = threadIdx.x;
  const int gx = key.x * 8 + (li >> 6), gy = key.y * 8 + ((li >> 3) & 7), gz = key.z * 8 + (li & 7);
  int cube = 0, ntri = 0, cb;
  if (cell_eval<false>(v, H, key.x, key.y, key.z, gx, gy, gz, box, nullptr, nullptr, cb)) {
    cube = cb;
    ntri = mc_num_tri(c_mc_tri[cube]);
    if (ntri == 0) cube = 0;
  }
  cubes[(size_t)blockIdx.x * kBlockVox + li] = (uint8_t)cube;
  int tot;
  wg_exclusive_scan<kCellThreads>(ntri, red, tot);
  if (li == 0) block_tri[blockIdx.x] = tot;
}

// VertexInterpolation at isolevel 0 (meshExtractor.cu:95-123)
__device__ __forceinline__ f3 vertex_interp(f3 p1, f3 p2, float d1, float d2) {
  const float eps = 0.00001f;
  if (fabsf(0.f - d1) < eps) return p1;
  if (fabsf(0.f - d2) < eps) return p2;
  if (fabsf(d1 - d2) < eps) return p1;
  const float mu = (0.f - d1) / (d2 - d1);
  return mk(p1.x + mu * (p2.x - p1.x), p1.y + mu * (p2.y - p1.y), p1.z + mu * (p2.z - p1.z));
}

__device__ __forceinline__ uint64_t edge_id(int gx, int gy, int gz, int e) {
  const int a = c_edge_a[e], b = c_edge_b[e];
  const int lo = a < b ? a : b;
  const int cx = gx + ((lo >> 2) & 1), cy = gy + ((lo >> 1) & 1), cz = gz + (lo & 1);
  const uint64_t off = 1u << 19, m = (1u << 20) - 1;
  return ((uint64_t)c_edge_axis[e] << 60) | ((((uint64_t)(cx + (int)off)) & m) << 40) |
         ((((uint64_t)(cy + (int)off)) & m) << 20) | (((uint64_t)(cz + (int)off)) & m);
}

__global__ __launch_bounds__(kCellThreads) void k_mc_emit(const DmView *__restrict__ vp, const int *__restrict__ list,
                                                          BoxArg box, const uint8_t *__restrict__ cubes,
                                                          const int *__restrict__ block_tri,
                                                          const long long *__restrict__ block_off, long long cap,
                                                          float *__restrict__ pos, float *__restrict__ col,
                                                          uint64_t *__restrict__ eid) {
  if (block_tri[blockIdx.x] == 0 || block_off[blockIdx.x] >= cap) return;  // uniform per workgroup
  __shared__ Halo H;
  __shared__ int red[kCellThreads / 64 + 1];
  const DmView &v = *vp;
  const int ptr = list[blockIdx.x];
  const int4 key = v.block_key[ptr];
  load_halo(v, key.x, key.y, key.z, H);
  const int li = threadIdx.x;
  const int cube = cubes[(size_t)blockIdx.x * kBlockVox + li];
  const uint64_t word = c_mc_tri[cube];
  const int ntri = cube ? mc_num_tri(word) : 0;
  int tot;
  const long long t0 = block_off[blockIdx.x] + wg_exclusive_scan<kCellThreads>(ntri, red, tot);
  if (ntri == 0 || t0 >= cap) return;
  const int gx = key.x * 8 + (li >> 6), gy = key.y * 8 + ((li >> 3) & 7), gz = key.z * 8 + (li & 7);
  f3 pc[8];
  float d[8];
  int cube_e;
  cell_eval<true>(v, H, key.x, key.y, key.z, gx, gy, gz, box, pc, d, cube_e);  // true: the count pass kept it
  const uint32_t rgbw = vox_rgbw(v.vox[(size_t)ptr * kBlockVox + li]);
  const float cr = (float)(rgbw & 0xFF) / 255.f, cg = (float)((rgbw >> 8) & 0xFF) / 255.f,
              cb = (float)((rgbw >> 16) & 0xFF) / 255.f;
  for (int k = 0; k < ntri; ++k) {
    const long long t = t0 + k;
    if (t >= cap) break;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = mc_tri_edge(word, 3 * k + j);
      const int a = c_edge_a[e], b = c_edge_b[e];
      const f3 p = vertex_interp(pc[a], pc[b], d[a], d[b]);
      const size_t o = (size_t)t * 9 + 3 * j;
      if (pos) { pos[o] = p.x; pos[o + 1] = p.y; pos[o + 2] = p.z; }
      if (col) { col[o] = cr; col[o + 1] = cg; col[o + 2] = cb; }
      if (eid) eid[(size_t)t * 3 + j] = edge_id(gx, gy, gz, e);
    }
  }
}

// ---------------------------------------------------------------- host

// Scratch of dm_mesh_extract_noalloc: one fixed array per device, resident
// until the process ends.  Segments (bytes, each 256-aligned):
//   counts  int32 per bucket            (kNaMaxBuckets)
//   offs    int64 per bucket            (kNaMaxBuckets)
//   wsum1   scan workspace of the buckets
//   list    pool index per allocated block (kNaMaxBlocks)
//   cubes   cube index per cell of a chunk of kNaChunk blocks
//   btri    triangles per block of the chunk
//   boff    triangle offset per block of the chunk
//   wsum2   scan workspace of the chunk
// Every segment is written before it is read within one call, so nothing
// needs clearing between calls.
constexpr int kNaMaxBuckets = 1 << 20;
constexpr int kNaMaxBlocks = 1 << 21;
constexpr int kNaChunk = 1 << 14;
constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaOffCounts = 0;
constexpr size_t kNaOffOffs = kNaOffCounts + na_align((size_t)kNaMaxBuckets * 4);
constexpr size_t kNaOffWsum1 = kNaOffOffs + na_align((size_t)kNaMaxBuckets * 8);
constexpr size_t kNaOffList = kNaOffWsum1 + na_align((size_t)(kNaMaxBuckets / kScanItems + 2) * 8);
constexpr size_t kNaOffCubes = kNaOffList + na_align((size_t)kNaMaxBlocks * 4);
constexpr size_t kNaOffBtri = kNaOffCubes + na_align((size_t)kNaChunk * kBlockVox);
constexpr size_t kNaOffBoff = kNaOffBtri + na_align((size_t)kNaChunk * 4);
constexpr size_t kNaOffWsum2 = kNaOffBoff + na_align((size_t)kNaChunk * 8);
constexpr size_t kNaScratchBytes = kNaOffWsum2 + na_align((size_t)(kNaChunk / kScanItems + 2) * 8);

static __device__ __align__(256) uint8_t scratch[kNaScratchBytes];

// Device address of `scratch` per device, taken once with that device
// current.  The host-side name of the array is not a device address.
constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static uint8_t *g_na_ptr[kNaMaxDevices];

static int na_scratch(int device, uint8_t **out) {
  if (device < 0 || device >= kNaMaxDevices) {
    set_error("dm_mesh_extract_noalloc: device index out of range");
    return DM_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[device]) {
    void *p = nullptr;
    DM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[device] = (uint8_t *)p;
  }
  *out = g_na_ptr[device];
  return DM_OK;
}

// Makes `device` current for the scope and restores the previous one.
struct DeviceScope {
  int prev = -1;
  int enter(int device) {
    DM_HIP(hipGetDevice(&prev));
    if (prev != device) DM_HIP(hipSetDevice(device));
    return DM_OK;
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

}  // namespace dm

using namespace dm;
struct dm_volume : Sess {};

extern "C" {

int dm_mc_tri_table(uint64_t *out) {
  if (!out) { set_error("null argument"); return DM_E_INVALID; }
  memcpy(out, h_mc_tri, sizeof h_mc_tri);
  return DM_OK;
}

int dm_mesh_extract(dm_volume *vol, const double *box_lo, const double *box_hi, int64_t cap, float *pos, float *col,
                    uint64_t *edge_ids, int64_t *ntri_out) {
  if (!vol || !ntri_out || cap < 0 || (!box_lo) != (!box_hi)) {
    set_error("dm_mesh_extract: invalid argument");
    return DM_E_INVALID;
  }
  hipStream_t st = vol->stream;
  const int nb = vol->h.num_buckets;
  BoxArg box{};
  box.use = box_lo != nullptr;
  for (int k = 0; k < 3 && box.use; ++k) { box.lo[k] = box_lo[k]; box.hi[k] = box_hi[k]; }
  DevBuf counts, offs, wsum1, list, cubes, btri, boff, wsum2, dpos, dcol, deid;
  if (int rc = dev_alloc(counts, (size_t)nb * 4)) return rc;
  if (int rc = dev_alloc(offs, (size_t)nb * 8)) return rc;
  hipLaunchKernelGGL(k_gather_counts, dim3((nb + 255) / 256), dim3(256), 0, st, vol->d_view, counts.as<int>());
  long long *d_nblk = nullptr;
  if (int rc = scan_ints(counts.as<int>(), nb, offs.as<long long>(), wsum1, st, &d_nblk)) return rc;
  long long nblk = 0;
  DM_HIP(hipMemcpyAsync(&nblk, d_nblk, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  *ntri_out = 0;
  if (nblk == 0) return DM_OK;
  if (nblk > vol->h.num_blocks) {
    set_error("dm_mesh_extract: hash table holds more entries than the pool");
    return DM_E_HIP;
  }
  const int n = (int)nblk;
  if (int rc = dev_alloc(list, (size_t)n * 4)) return rc;
  if (int rc = dev_alloc(cubes, (size_t)n * kBlockVox)) return rc;
  if (int rc = dev_alloc(btri, (size_t)n * 4)) return rc;
  if (int rc = dev_alloc(boff, (size_t)n * 8)) return rc;
  hipLaunchKernelGGL(k_block_list, dim3((nb + 255) / 256), dim3(256), 0, st, vol->d_view, offs.as<long long>(),
                     list.as<int>());
  hipLaunchKernelGGL(k_mc_count, dim3(n), dim3(kCellThreads), 0, st, vol->d_view, list.as<int>(), box,
                     cubes.as<uint8_t>(), btri.as<int>());
  DM_HIP(hipGetLastError());
  long long *d_ntri = nullptr;
  if (int rc = scan_ints(btri.as<int>(), n, boff.as<long long>(), wsum2, st, &d_ntri)) return rc;
  long long T = 0;
  DM_HIP(hipMemcpyAsync(&T, d_ntri, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  *ntri_out = T;
  const long long m = std::min<long long>(T, cap);
  if (m == 0 || (!pos && !col && !edge_ids)) return DM_OK;
  if (pos) { if (int rc = dev_alloc(dpos, (size_t)m * 36)) return rc; }
  if (col) { if (int rc = dev_alloc(dcol, (size_t)m * 36)) return rc; }
  if (edge_ids) { if (int rc = dev_alloc(deid, (size_t)m * 24)) return rc; }
  hipLaunchKernelGGL(k_mc_emit, dim3(n), dim3(kCellThreads), 0, st, vol->d_view, list.as<int>(), box,
                     cubes.as<uint8_t>(), btri.as<int>(), boff.as<long long>(), m, dpos.as<float>(),
                     dcol.as<float>(), deid.as<uint64_t>());
  DM_HIP(hipGetLastError());
  if (pos) DM_HIP(hipMemcpyAsync(pos, dpos.p, (size_t)m * 36, hipMemcpyDeviceToHost, st));
  if (col) DM_HIP(hipMemcpyAsync(col, dcol.p, (size_t)m * 36, hipMemcpyDeviceToHost, st));
  if (edge_ids) DM_HIP(hipMemcpyAsync(edge_ids, deid.p, (size_t)m * 24, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  return DM_OK;
}

int dm_mesh_extract_noalloc(dm_volume *vol, const double *box_lo, const double *box_hi, int64_t cap, float *dev_pos,
                            float *dev_col, uint64_t *dev_edge_ids, int64_t *ntri_out) {
  if (!vol || !ntri_out || cap < 0 || (!box_lo) != (!box_hi)) {
    set_error("dm_mesh_extract_noalloc: invalid argument");
    return DM_E_INVALID;
  }
  const int nb = vol->h.num_buckets;
  if (nb > kNaMaxBuckets) {
    set_error("dm_mesh_extract_noalloc: the volume has more than 2^20 hash buckets (use dm_mesh_extract)");
    return DM_E_CAPACITY;
  }
  DeviceScope scope;
  DM_TRY(scope.enter(vol->device));
  uint8_t *base = nullptr;
  DM_TRY(na_scratch(vol->device, &base));
  int *counts = (int *)(base + kNaOffCounts);
  long long *offs = (long long *)(base + kNaOffOffs), *wsum1 = (long long *)(base + kNaOffWsum1);
  int *list = (int *)(base + kNaOffList);
  uint8_t *cubes = base + kNaOffCubes;
  int *btri = (int *)(base + kNaOffBtri);
  long long *boff = (long long *)(base + kNaOffBoff), *wsum2 = (long long *)(base + kNaOffWsum2);

  hipStream_t st = vol->stream;
  BoxArg box{};
  box.use = box_lo != nullptr;
  for (int k = 0; k < 3 && box.use; ++k) { box.lo[k] = box_lo[k]; box.hi[k] = box_hi[k]; }
  hipLaunchKernelGGL(k_gather_counts, dim3((nb + 255) / 256), dim3(256), 0, st, vol->d_view, counts);
  long long *d_total = nullptr;
  DM_TRY(scan_ints_ws(counts, nb, offs, wsum1, st, &d_total));
  long long nblk = 0;
  DM_HIP(hipMemcpyAsync(&nblk, d_total, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  *ntri_out = 0;
  if (nblk == 0) return DM_OK;
  if (nblk > vol->h.num_blocks) {
    set_error("dm_mesh_extract_noalloc: hash table holds more entries than the pool");
    return DM_E_HIP;
  }
  if (nblk > kNaMaxBlocks) {
    set_error("dm_mesh_extract_noalloc: more than 2^21 allocated blocks (use dm_mesh_extract)");
    return DM_E_CAPACITY;
  }
  const int n = (int)nblk;
  hipLaunchKernelGGL(k_block_list, dim3((nb + 255) / 256), dim3(256), 0, st, vol->d_view, offs, list);
  const bool want = dev_pos || dev_col || dev_edge_ids;
  long long T = 0;
  for (int c0 = 0; c0 < n; c0 += kNaChunk) {
    const int cn = std::min(kNaChunk, n - c0);
    hipLaunchKernelGGL(k_mc_count, dim3(cn), dim3(kCellThreads), 0, st, vol->d_view, list + c0, box, cubes, btri);
    DM_HIP(hipGetLastError());
    DM_TRY(scan_ints_ws(btri, cn, boff, wsum2, st, &d_total));
    long long Tc = 0;
    DM_HIP(hipMemcpyAsync(&Tc, d_total, 8, hipMemcpyDeviceToHost, st));
    DM_HIP(hipStreamSynchronize(st));
    const long long m = std::min<long long>(Tc, cap - std::min<long long>(T, cap));
    if (want && m > 0) {
      hipLaunchKernelGGL(k_mc_emit, dim3(cn), dim3(kCellThreads), 0, st, vol->d_view, list + c0, box, cubes, btri,
                         boff, m, dev_pos ? dev_pos + (size_t)T * 9 : nullptr,
                         dev_col ? dev_col + (size_t)T * 9 : nullptr,
                         dev_edge_ids ? dev_edge_ids + (size_t)T * 3 : nullptr);
      DM_HIP(hipGetLastError());
    }
    T += Tc;
  }
  DM_HIP(hipStreamSynchronize(st));
  *ntri_out = T;
  return DM_OK;
}

}  // extern "C"
