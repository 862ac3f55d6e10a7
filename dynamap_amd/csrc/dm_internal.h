// Internal definitions shared by the HIP translation units of libdynamap_hip.
//
// Every kernel works on a table of DmView records (one per reconstruction
// session) that lives in device memory, plus a list of the session slots that
// take part in a launch (blockIdx.y indexes that list).  A single-session call
// is the same launch with a one-entry list, so the per-frame path and the
// batched serving path run the same kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dm {

constexpr int kBlockEdge = 8;              // voxels per block edge
constexpr int kBlockVox = 512;             // voxels per block
constexpr int kKeyRange = 1 << 20;         // |block coordinate| limit of the packed key
constexpr int kLevels = 3;                 // ICP pyramid levels
constexpr int kIcpVals = 32;               // per-workgroup ICP partial record (30 used)
// Coarse empty-space bitmap of the raycast: cells of 8^3 blocks, a window of
// kCoarse^3 cells centred on block (0, 0, 0); cells outside it count as near.
constexpr int kCoarse = 64;
constexpr int kCoarseWords = kCoarse * kCoarse * kCoarse / 32;
// Block-level bitmaps over the same window (kWin^3 blocks centred on block
// (0, 0, 0), x fastest): "near" (an allocated block with an observed voxel
// below the far threshold) and "dilated" (a near block within one block).
// Blocks outside the window are looked up through the hash table.
constexpr int kWin = kCoarse * 8;
constexpr size_t kWinWords = (size_t)kWin * kWin * kWin / 32;
// Mid-level cells of 4^3 blocks over the same window (kCell4^3 cells).
constexpr int kCell4 = kWin / 4;
constexpr int kCell4Words = kCell4 * kCell4 * kCell4 / 32;

// per-volume counters, int32 each
enum Ctr : int {
  C_NUM_ALLOC = 0,   // pool high-water mark: indices 0..n-1 have been used (live, or freed
                     // and listed in free_list); live blocks = C_NUM_ALLOC - C_NUM_FREE
  C_N_NEW = 1,       // new keys collected this frame
  C_N_INS = 2,       // of which inserted
  C_CURSOR = 3,      // pool cursor within the frame
  C_DROP_BUCKET = 4, // keys dropped because their bucket was full (cumulative)
  C_DROP_POOL = 5,   // keys dropped because the pool was exhausted (cumulative)
  C_DROP_RANGE = 6,  // block visits outside the packed-key range (cumulative)
  C_DSET_OVF = 7,    // new keys lost to a full per-frame key set (cumulative)
  C_HAS_MODEL = 8,   // a raycast model exists (tracking can run)
  C_ICP_FAIL = 9,    // frames whose increment was reset by the guard (cumulative)
  C_NVIS = 10,       // blocks that passed the frustum test in the last integration
  C_NVISIT = 11,     // block visits of the last allocation's ray traversal
  C_NNEW = 12,       // keys inserted by the last allocation
  C_ICP_STOP = 13,   // level + 1 of the ICP level that ended early this frame, 0 = none
  C_NUM_FREE = 14,   // freed pool indices waiting in free_list (a stack; reused first)
  C_COUNT = 16
};

// pose block (double) offsets
enum PoseD : int {
  PD_T = 0,       // camera-to-world of the current frame (row-major 4x4)
  PD_TPREV = 16,  // camera-to-world of the model (last integrated) frame
  PD_TINC = 32,   // increment: current camera -> model camera
  PD_ACC = 48,    // summed twist of the frame (6)
  PD_TINIT = 64,  // pose used when no model exists yet
  PD_PERSIST = 80,  // the entries above persist between frames (snapshot)
  PD_PREVACC = 80,  // summed twist after the last step that did not end its level (6)
  PD_COUNT = 86
};
// pose block (float) offsets
enum PoseF : int { PF_T = 0, PF_TINV = 16, PF_TINC = 32, PF_TPREV = 48, PF_COUNT = 64 };

struct DmView {
  // sensor (level 0)
  int rows, cols;
  float fx, fy, cx, cy;
  // volume parameters
  float voxel_size, trunc, max_depth, min_depth;
  int num_buckets, bucket_size, bstride, num_blocks, max_weight;
  int dset_mask, newcap;
  // volume arrays
  uint8_t *buckets;     // num_buckets * bstride: u64 key[S] | i32 ptr[S] | i32 count
  uint64_t *vox;        // num_blocks * 512 voxel words (sdf bits | r | g<<8 | b<<16 | w<<24 <<32)
  int4 *block_key;      // num_blocks block coordinates, w = 1 live, 0 freed
  int32_t *free_list;   // num_blocks: pool indices returned by deletion (HashTable::DeleteBlock
                        // + Heap::Append), taken back by later insertions
  uint8_t *blk_near;    // num_blocks: 1 when an observed voxel lies below the raycast far threshold
  uint32_t *nearmap;    // kWinWords: window bit of every block with blk_near set
  uint32_t *dilmap;     // kWinWords: window bit of every block within one block of a near block
  uint32_t *coarse;     // kCoarse^3 bits: cell of 8^3 blocks holds a block with its dilmap bit set
  uint32_t *cell4;      // kCell4^3 bits: cell of 4^3 blocks holds a block with its dilmap bit set
  uint32_t *allocmap;   // kWinWords: window bit of every allocated block (the allocation's
                        // table test; blocks outside the window go through the hash table)
  uint64_t *dset;       // per-frame new-key set, dset_mask + 1 slots, 0 = empty
  uint64_t *newkeys;    // newcap
  int32_t *newdslot;    // newcap
  int32_t *newnext;     // newcap
  int32_t *newrank;     // newcap (rank, or -1 when dropped)
  int32_t *newlen;      // newcap (bucket list length, written by rank 0)
  int32_t *bucket_head; // num_buckets, -1 = empty
  int32_t *ctr;         // C_COUNT
  // current frame input (device)
  const float *depth;
  const float *depth_track;  // depth the tracking maps are built from (filtered), or null = depth
  const uint8_t *rgb;   // interleaved r,g,b (may be null)
  // per pixel of the current frame, written by the allocation pass for the
  // integration: x = bits of |getPoint3d(u, v, d)| (the ray distance to the
  // measured surface), sign bit set when d == max_depth, NaN when d is out of
  // range; y = the pixel's colour r | g << 8 | b << 16 (0 without colour)
  uint2 *surf;
  uint32_t *visited;     // num_blocks bits: blocks found by a counting traversal (profiling)
  // maps: vertex+intensity and normal per level, source (live frame) and
  // target (model raycast); level l is (rows >> l) x (cols >> l)
  float4 *svert[kLevels], *snorm[kLevels];
  float4 *tvert[kLevels], *tnorm[kLevels];
  // pose state
  double *posed;        // PD_COUNT
  float *posef;         // PF_COUNT
  float *partials;      // ICP partial sums: max_wg * kIcpVals
  int max_wg;
  // optional raycast outputs (level 0), may be null
  float *ray_depth;
  uint8_t *ray_rgb;
  // tracking parameters
  int iters[kLevels];
  float icp_dmin, icp_dmax, dist_max, w_photo, lambda, max_inc, min_inc;
  int pad_[2];
};

// ---- host-side launch helpers (defined in the .hip files) -----------------
struct Launch {
  const DmView *views;  // device table
  const int *active;    // device list of session slots
  int nact;             // entries in the list
  int max_pixels;       // largest rows*cols among active sessions
  int max_tiles;        // largest number of 16x16 pixel tiles among active sessions
  int max_rows, max_cols;  // largest frame among active sessions
  int ray_tw_log2 = 4;  // raycast tile: 2^ray_tw_log2 x (256 >> ray_tw_log2) pixels
  int integ_wg = 2048;  // integration workgroups per session
  int icp_mfma = 0;     // ICP linearisation on the matrix cores (k_icp_reduce_mfma) instead of the shuffle tree
  hipStream_t stream;
};

void launch_alloc(const Launch &L);
// traversal kernel only (profiling); count_unique also marks the pool index
// of every block found in v.visited
void launch_alloc_rays(const Launch &L, bool count_unique = false);
void launch_integrate(const Launch &L);
void launch_raycast(const Launch &L);
// HashTable::DeleteBlock over a list of n block coordinates (one session,
// view = its one-entry table): dmask (num_buckets 64-bit words, one bit per
// slot) must be zero and is left zero; touched holds n ints; counts =
// {touched buckets, freed blocks}, zero on entry.  Rebuild the raycast's
// bitmaps afterwards.
void launch_delete(const DmView *view, const int4 *req, int n, unsigned long long *dmask, int *touched, int *counts,
                   hipStream_t st);
// live blocks whose centre lies farther than radius from c, appended to req
// (capacity: the pool size) with *nreq counting them
void launch_select_far(const DmView *view, const double c[3], double radius, int4 *req, int *nreq, hipStream_t st);
void rebuild_far_maps(const Launch &L);  // clear and rebuild nearmap/dilmap/cell4/coarse/allocmap (not for graph capture)
void launch_source_maps(const Launch &L);
// ICP linearisation at one level (one FP32 record per workgroup); with solve,
// a second kernel sums the records and takes the Gauss-Newton step on the
// device.
void launch_icp_reduce(const Launch &L, int level, int it, bool solve);
void launch_pose_begin(const Launch &L);
void launch_pose_log(const Launch &L, const int *dest, double *log);
void launch_bilateral(const float *in, float *out, int rows, int cols, int r, float sigma_d, float sigma_s,
                      hipStream_t st);

constexpr int kIcpPix = 1024;   // pixels per ICP reduction workgroup (256 threads x 4)
__host__ __device__ inline int icp_wg_count(int pixels) { return (pixels + kIcpPix - 1) / kIcpPix; }
// The FP64 sum of the per-workgroup records in the order the device uses
// (thread t of 256 takes records t, t+256, ...; wave_sum's tree within each
// wavefront; the four wavefronts as (w0 + w1) + (w2 + w3)).
void icp_sum_records(const float *part, int nwg, double *sums);

}  // namespace dm
