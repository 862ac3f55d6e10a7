// C ABI: volume, session (per-frame pipeline) and tracking entry points.
#include <string.h>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>
#include "dm_scan.h"
#include "dm_session.h"
#include "dm_solve.h"

namespace dm {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? DM_E_CAPACITY : DM_E_HIP;
}

static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
static int pow2_at_least(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

int validate_props(const dm_tsdf_props &p) {
  if (!(p.voxel_size > 0.f) || p.num_buckets <= 0 || p.bucket_size <= 0 || p.bucket_size > 64 ||
      p.num_blocks <= 0 || p.num_blocks >= (1 << 23) || p.block_size != 8 || p.max_sdf_weight < 1 ||
      p.max_sdf_weight > 255 || !(p.truncation_distance > 0.f) || !(p.max_sensor_depth > p.min_sensor_depth)) {
    set_error("invalid tsdf properties (block_size must be 8, bucket_size 1..64, num_blocks 1..2^23-1, "
              "max weight 1..255)");
    return DM_E_INVALID;
  }
  return DM_OK;
}

static int bucket_stride(int S) { return (int)round_up((size_t)(12 * S + 4), 128); }
// Capacity of the per-frame new-key list.  A frame with more new keys than
// this cannot fit the pool either (it holds at least num_blocks), so it is
// rejected as a whole; only its drop counts are then lower bounds.
static int new_cap(int num_blocks) { return std::max(4096, std::min(num_blocks, 1 << 22)); }

size_t session_bytes(const dm_tsdf_props &p, uint32_t rows, uint32_t cols) {
  const size_t P = (size_t)rows * cols;
  const size_t nc = (size_t)new_cap(p.num_blocks);
  size_t b = (size_t)p.num_buckets * bucket_stride(p.bucket_size);
  b += ((size_t)p.num_blocks + 1) * kBlockVox * 8 + (size_t)p.num_blocks * (16 + 1 + 4);
  b += ((size_t)kCoarseWords + (size_t)kCell4Words + 3 * kWinWords) * 4;
  b += (size_t)pow2_at_least((int)(2 * nc)) * 8 + nc * (8 + 4 * 4) + (size_t)p.num_buckets * 4;
  size_t maps = 0;
  for (int l = 0; l < kLevels; ++l) maps += (size_t)(rows >> l) * (cols >> l) * 16 * 4;
  b += maps + P * (4 + 3) * 2 + P * 8 + ((P + 255) / 256 + 1) * kIcpVals * 4 + 4096;
  b += ((size_t)p.num_blocks + 31) / 32 * 4;
  return b;
}

template <class T> int Sess::dalloc(T **ptr, size_t count, bool zero) {
  void *p = nullptr;
  const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
  DM_HIP(hipMalloc(&p, bytes));
  allocs.push_back(p);
  hbm_bytes += bytes;
  if (zero) DM_HIP(hipMemsetAsync(p, 0, bytes, stream));
  *ptr = (T *)p;
  return DM_OK;
}


int Sess::init(const dm_tsdf_props &p, uint32_t rows, uint32_t cols) {
  DM_TRY(validate_props(p));
  if (rows == 0 || cols == 0 || rows > 8192 || cols > 8192) {
    set_error("invalid frame size");
    return DM_E_INVALID;
  }
  props = p;
  max_rows = rows;
  max_cols = cols;
  DM_HIP(hipGetDevice(&device));
  DM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  memset(&h, 0, sizeof(h));
  h.rows = (int)rows;
  h.cols = (int)cols;
  h.voxel_size = p.voxel_size;
  h.trunc = p.truncation_distance;
  h.max_depth = p.max_sensor_depth;
  h.min_depth = p.min_sensor_depth;
  h.num_buckets = p.num_buckets;
  h.bucket_size = p.bucket_size;
  h.bstride = bucket_stride(p.bucket_size);
  h.num_blocks = p.num_blocks;
  h.max_weight = p.max_sdf_weight;
  h.newcap = new_cap(p.num_blocks);
  const int dsz = pow2_at_least(2 * h.newcap);
  h.dset_mask = dsz - 1;
  const size_t P = (size_t)rows * cols;
  DM_TRY(dalloc(&h.buckets, (size_t)p.num_buckets * h.bstride));
  // one block more than the pool: block num_blocks stays all zero and stands
  // for unallocated blocks in the raycast's gathers
  DM_TRY(dalloc(&h.vox, ((size_t)p.num_blocks + 1) * kBlockVox));
  DM_TRY(dalloc(&h.block_key, (size_t)p.num_blocks));
  DM_TRY(dalloc(&h.free_list, (size_t)p.num_blocks, false));
  DM_TRY(dalloc(&h.blk_near, (size_t)p.num_blocks));
  DM_TRY(dalloc(&h.nearmap, kWinWords));
  DM_TRY(dalloc(&h.dilmap, kWinWords));
  DM_TRY(dalloc(&h.allocmap, kWinWords));
  DM_TRY(dalloc(&h.coarse, (size_t)kCoarseWords));
  DM_TRY(dalloc(&h.cell4, (size_t)kCell4Words));
  DM_TRY(dalloc(&h.dset, (size_t)dsz));
  DM_TRY(dalloc(&h.newkeys, (size_t)h.newcap, false));
  DM_TRY(dalloc(&h.newdslot, (size_t)h.newcap, false));
  DM_TRY(dalloc(&h.newnext, (size_t)h.newcap, false));
  DM_TRY(dalloc(&h.newrank, (size_t)h.newcap, false));
  DM_TRY(dalloc(&h.newlen, (size_t)h.newcap, false));
  DM_TRY(dalloc(&h.bucket_head, (size_t)p.num_buckets, false));
  DM_HIP(hipMemsetAsync(h.bucket_head, 0xff, (size_t)p.num_buckets * 4, stream));
  DM_TRY(dalloc(&h.ctr, (size_t)C_COUNT));
  DM_TRY(dalloc(&d_depth, P));
  DM_TRY(dalloc(&d_rgb, P * 3));
  DM_TRY(dalloc(&h.surf, P));
  DM_TRY(dalloc(&h.visited, ((size_t)p.num_blocks + 31) / 32));
  for (int l = 0; l < kLevels; ++l) {
    const size_t Pl = (size_t)(rows >> l) * (cols >> l);
    DM_TRY(dalloc(&h.svert[l], Pl));
    DM_TRY(dalloc(&h.snorm[l], Pl));
    DM_TRY(dalloc(&h.tvert[l], Pl));
    DM_TRY(dalloc(&h.tnorm[l], Pl));
  }
  DM_TRY(dalloc(&h.posed, (size_t)PD_COUNT));
  DM_TRY(dalloc(&h.posef, (size_t)PF_COUNT));
  h.max_wg = (int)((P + 255) / 256) + 1;
  DM_TRY(dalloc(&h.partials, (size_t)h.max_wg * kIcpVals));
  DM_TRY(dalloc(&d_ray_depth, P));
  DM_TRY(dalloc(&d_ray_rgb, P * 3));
  DM_TRY(dalloc(&d_view, 1));
  DM_TRY(dalloc(&d_active, 1));
  pinned_bytes = P * 7 + 64;
  DM_HIP(hipHostMalloc((void **)&h_pinned, pinned_bytes, hipHostMallocDefault));
  // identity poses
  double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  std::vector<double> pd(PD_COUNT, 0.0);
  for (int k = 0; k < 16; ++k) {
    pd[PD_T + k] = I[k]; pd[PD_TPREV + k] = I[k]; pd[PD_TINC + k] = I[k]; pd[PD_TINIT + k] = I[k];
  }
  DM_HIP(hipMemcpyAsync(h.posed, pd.data(), PD_COUNT * sizeof(double), hipMemcpyHostToDevice, stream));
  h.depth = d_depth;
  h.rgb = nullptr;
  h.iters[0] = 10; h.iters[1] = 5; h.iters[2] = 4;
  h.icp_dmin = p.min_sensor_depth; h.icp_dmax = p.max_sensor_depth;
  h.dist_max = 0.1f; h.w_photo = 0.f; h.lambda = 0.02f; h.max_inc = 0.3f;
  int zero = 0;
  DM_HIP(hipMemcpyAsync(d_active, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
  DM_TRY(push_view());
  DM_HIP(hipStreamSynchronize(stream));
  return DM_OK;
}

void Sess::release() {
  if (stream) hipStreamSynchronize(stream);
  for (void *p : allocs) hipFree(p);
  allocs.clear();
  if (h_pinned) hipHostFree(h_pinned);
  h_pinned = nullptr;
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

int Sess::push_view() {
  DM_HIP(hipMemcpyAsync(d_view, &h, sizeof(DmView), hipMemcpyHostToDevice, stream));
  // the host copy may change before the copy runs: wait for it
  DM_HIP(hipStreamSynchronize(stream));
  return DM_OK;
}

Launch Sess::launch() const {
  Launch L;
  L.views = d_view;
  L.active = d_active;
  L.nact = 1;
  L.max_pixels = h.rows * h.cols;
  L.max_tiles = ((h.cols + 15) / 16) * ((h.rows + 15) / 16);
  L.max_rows = h.rows;
  L.max_cols = h.cols;
  L.icp_mfma = icp_mfma;
  L.stream = stream;
  return L;
}

int Sess::set_sensor(const dm_sensor *s) {
  if (!s || s->rows == 0 || s->cols == 0 || s->rows > max_rows || s->cols > max_cols || !(s->fx > 0) ||
      !(s->fy > 0)) {
    set_error("invalid sensor (size must not exceed the object's frame size)");
    return DM_E_INVALID;
  }
  h.rows = (int)s->rows;
  h.cols = (int)s->cols;
  h.fx = s->fx; h.fy = s->fy; h.cx = s->cx; h.cy = s->cy;
  return DM_OK;
}

int Sess::set_pose_f(const float pose[16]) {
  float pf[32];
  double pd[16];
  for (int k = 0; k < 16; ++k) { pf[k] = pose[k]; pd[k] = (double)pose[k]; }
  rigid_inverse_ff(pose, pf + 16);
  DM_HIP(hipMemcpyAsync(h.posef + PF_T, pf, 32 * sizeof(float), hipMemcpyHostToDevice, stream));
  DM_HIP(hipMemcpyAsync(h.posed + PD_T, pd, 16 * sizeof(double), hipMemcpyHostToDevice, stream));
  DM_HIP(hipStreamSynchronize(stream));
  return DM_OK;
}

int Sess::stage_frame(const float *host_depth, const uint8_t *host_rgb) {
  const size_t P = (size_t)h.rows * h.cols;
  DM_HIP(hipStreamSynchronize(stream));
  memcpy(h_pinned, host_depth, P * 4);
  DM_HIP(hipMemcpyAsync(d_depth, h_pinned, P * 4, hipMemcpyHostToDevice, stream));
  h.depth = d_depth;
  if (host_rgb) {
    memcpy(h_pinned + P * 4, host_rgb, P * 3);
    DM_HIP(hipMemcpyAsync(d_rgb, h_pinned + P * 4, P * 3, hipMemcpyHostToDevice, stream));
    h.rgb = d_rgb;
  } else {
    h.rgb = nullptr;
  }
  return DM_OK;
}

// ---------------------------------------------------------------------------
// tracking orchestration shared by the session and the tracking test hooks

static int total_iters(const DmView &h) { return h.iters[0] + h.iters[1] + h.iters[2]; }

// Run the level loop.  Device solve: back-to-back kernels.  Host solve:
// read each reduction back and solve on the host (FP64), as the reference.
static int run_icp(Sess &S, bool host_solve) {
  const Launch L = S.launch();
  const int total = total_iters(S.h);
  if (!host_solve) {
    for (int l = kLevels - 1; l >= 0; --l)
      for (int it = 1; it <= S.h.iters[l]; ++it) {
        launch_icp_reduce(L, l, it, true);
      }
    DM_HIP(hipGetLastError());
    return DM_OK;
  }
  double pd[PD_COUNT];
  int has_model = 0;
  DM_HIP(hipMemcpyAsync(pd, S.h.posed, sizeof pd, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipMemcpyAsync(&has_model, S.h.ctr + C_HAS_MODEL, sizeof(int), hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  if (!has_model) return DM_OK;
  double *Tinc = pd + PD_TINC, *acc = pd + PD_ACC, *pacc = pd + PD_PREVACC;
  std::vector<float> part;
  for (int l = kLevels - 1; l >= 0; --l) {
    const int P = (S.h.rows >> l) * (S.h.cols >> l);
    const int nwg = icp_wg_count(P);
    part.resize((size_t)nwg * kIcpVals);
    for (int it = 1; it <= S.h.iters[l]; ++it) {
      launch_icp_reduce(L, l, it, false);
      DM_HIP(hipMemcpyAsync(part.data(), S.h.partials, part.size() * sizeof(float), hipMemcpyDeviceToHost,
                            S.stream));
      DM_HIP(hipStreamSynchronize(S.stream));
      double sums[30], A[36], b[6], xi[6], E[16];
      icp_sum_records(part.data(), nwg, sums);
      sym6_from_upper(sums, A);
      for (int k = 0; k < 6; ++k) b[k] = sums[21 + k];
      icp_step(A, b, it, (double)S.h.lambda, xi);
      se3_exp(xi, E);
      mat4_mul(E, Tinc, Tinc);
      for (int k = 0; k < 6; ++k) acc[k] += xi[k];
      const bool stop = level_converged(acc, pacc, (double)S.h.min_inc);
      float tf[16];
      for (int k = 0; k < 16; ++k) tf[k] = (float)Tinc[k];
      DM_HIP(hipMemcpyAsync(S.h.posef + PF_TINC, tf, sizeof tf, hipMemcpyHostToDevice, S.stream));
      DM_HIP(hipMemcpyAsync(S.h.posed + PD_TINC, Tinc, 16 * sizeof(double), hipMemcpyHostToDevice, S.stream));
      DM_HIP(hipStreamSynchronize(S.stream));
      if (stop) break;
    }
  }
  if (total > 0) {
    const double ssum = acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5];
    if (ssum != ssum || std::isinf(ssum) || ssum >= (double)S.h.max_inc)
      for (int k = 0; k < 16; ++k) Tinc[k] = (k % 5 == 0) ? 1.0 : 0.0;
    mat4_mul(pd + PD_TPREV, Tinc, pd + PD_T);
    float pf[48];
    for (int k = 0; k < 16; ++k) pf[k] = (float)pd[PD_T + k];
    rigid_inverse_ff(pf, pf + 16);
    for (int k = 0; k < 16; ++k) pf[32 + k] = (float)Tinc[k];
    DM_HIP(hipMemcpyAsync(S.h.posed, pd, sizeof pd, hipMemcpyHostToDevice, S.stream));
    DM_HIP(hipMemcpyAsync(S.h.posef, pf, sizeof pf, hipMemcpyHostToDevice, S.stream));
    DM_HIP(hipStreamSynchronize(S.stream));
  }
  return DM_OK;
}

static int process_frame(Sess &S, bool host_solve, double *pose_out) {
  const Launch L = S.launch();
  if (S.filt_r > 0)
    launch_bilateral(S.h.depth, S.d_depth_f, S.h.rows, S.h.cols, S.filt_r, S.filt_sd, S.filt_ss, S.stream);
  launch_pose_begin(L);
  launch_source_maps(L);
  DM_TRY(run_icp(S, host_solve));
  launch_alloc(L);
  launch_integrate(L);
  launch_raycast(L);
  DM_HIP(hipGetLastError());
  double T[16];
  DM_HIP(hipMemcpyAsync(T, S.h.posed + PD_T, sizeof T, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  if (pose_out) memcpy(pose_out, T, sizeof T);
  S.frames++;
  return DM_OK;
}

// ---------------------------------------------------------------- snapshot
struct SnapHeader {
  char magic[8];
  uint32_t version, hdr_bytes;
  dm_tsdf_props props;
  uint32_t max_rows, max_cols, rows, cols;
  float fx, fy, cx, cy;
  int32_t num_alloc, iters[3];
  float icp_f[7];  // icp dmin, dmax, dist_max, w_photo, lambda, max_inc, min_inc
  int32_t ctr[C_COUNT];
  uint64_t payload;
};
static const char kSnapMagic[8] = {'D', 'M', 'S', 'N', 'A', 'P', '0', '4'};
static const uint32_t kSnapVersion = 4;

// device sections of a volume / session state, in file order; the raycast's
// bitmaps are not stored (restore rebuilds them from blk_near and the keys)
static std::vector<std::pair<void *, size_t>> snap_sections(const Sess &S, int nalloc, int nfree) {
  const DmView &h = S.h;
  std::vector<std::pair<void *, size_t>> v = {
      {h.buckets, (size_t)h.num_buckets * h.bstride}, {h.block_key, (size_t)nalloc * sizeof(int4)},
      {h.vox, (size_t)nalloc * kBlockVox * 8},        {h.blk_near, (size_t)nalloc},
      {h.free_list, (size_t)nfree * 4},
      {h.posed, (size_t)PD_PERSIST * 8},             {h.posef, (size_t)PF_COUNT * 4}};
  for (int l = 0; l < kLevels; ++l) {
    const size_t Pl = (size_t)(h.rows >> l) * (h.cols >> l);
    v.push_back({h.tvert[l], Pl * 16});
    v.push_back({h.tnorm[l], Pl * 16});
  }
  return v;
}

}  // namespace dm

using namespace dm;

struct dm_volume : Sess {};
struct dm_session : Sess {};

extern "C" {

const char *dm_last_error(void) { return g_err.c_str(); }

int dm_device_count(int32_t *n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  if (n) *n = c;
  return DM_OK;
}

// ---------------------------------------------------------------- volume
int dm_volume_create(const dm_tsdf_props *props, uint32_t max_rows, uint32_t max_cols, dm_volume **out) {
  if (!props || !out) { set_error("null argument"); return DM_E_INVALID; }
  dm_volume *v = new dm_volume();
  int r = v->init(*props, max_rows, max_cols);
  if (r != DM_OK) { v->release(); delete v; return r; }
  *out = v;
  return DM_OK;
}

int dm_volume_destroy(dm_volume *vol) {
  if (!vol) return DM_OK;
  vol->release();
  delete vol;
  return DM_OK;
}

static int volume_frame(dm_volume *vol, const float *depth, const uint8_t *rgb, const dm_sensor *sensor,
                        const float pose[16], bool integrate) {
  if (!vol || !depth || !pose) { set_error("null argument"); return DM_E_INVALID; }
  DM_TRY(vol->set_sensor(sensor));
  DM_TRY(vol->stage_frame(depth, rgb));
  DM_TRY(vol->set_pose_f(pose));
  DM_TRY(vol->push_view());
  const Launch L = vol->launch();
  launch_alloc(L);
  if (integrate) launch_integrate(L);
  DM_HIP(hipGetLastError());
  DM_HIP(hipStreamSynchronize(vol->stream));
  return DM_OK;
}

int dm_volume_integrate(dm_volume *vol, const float *host_depth, const uint8_t *host_rgb, const dm_sensor *sensor,
                        const float pose[16]) {
  return volume_frame(vol, host_depth, host_rgb, sensor, pose, true);
}

int dm_volume_allocate(dm_volume *vol, const float *host_depth, const dm_sensor *sensor, const float pose[16]) {
  return volume_frame(vol, host_depth, nullptr, sensor, pose, false);
}

int dm_volume_raycast(dm_volume *vol, const float pose[16], const dm_sensor *sensor, float *host_depth,
                      float *host_vertex, float *host_normal, uint8_t *host_rgb) {
  if (!vol || !pose) { set_error("null argument"); return DM_E_INVALID; }
  DM_TRY(vol->set_sensor(sensor));
  DM_TRY(vol->set_pose_f(pose));
  vol->h.ray_depth = vol->d_ray_depth;
  vol->h.ray_rgb = vol->d_ray_rgb;
  DM_TRY(vol->push_view());
  launch_raycast(vol->launch());
  vol->h.ray_depth = nullptr;
  vol->h.ray_rgb = nullptr;
  DM_HIP(hipGetLastError());
  const size_t P = (size_t)vol->h.rows * vol->h.cols;
  std::vector<float4> vt(P), nt(P);
  DM_HIP(hipMemcpyAsync(vt.data(), vol->h.tvert[0], P * 16, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipMemcpyAsync(nt.data(), vol->h.tnorm[0], P * 16, hipMemcpyDeviceToHost, vol->stream));
  if (host_depth)
    DM_HIP(hipMemcpyAsync(host_depth, vol->d_ray_depth, P * 4, hipMemcpyDeviceToHost, vol->stream));
  if (host_rgb) DM_HIP(hipMemcpyAsync(host_rgb, vol->d_ray_rgb, P * 3, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  DM_TRY(vol->push_view());
  for (size_t i = 0; i < P; ++i) {
    if (host_vertex) { host_vertex[3 * i] = vt[i].x; host_vertex[3 * i + 1] = vt[i].y; host_vertex[3 * i + 2] = vt[i].z; }
    if (host_normal) { host_normal[3 * i] = nt[i].x; host_normal[3 * i + 1] = nt[i].y; host_normal[3 * i + 2] = nt[i].z; }
  }
  return DM_OK;
}

int dm_volume_num_allocated(dm_volume *vol, int32_t *n) {
  if (!vol || !n) { set_error("null argument"); return DM_E_INVALID; }
  int32_t c[C_COUNT];
  DM_HIP(hipMemcpyAsync(c, vol->h.ctr, sizeof c, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  *n = c[C_NUM_ALLOC] - c[C_NUM_FREE];
  return DM_OK;
}

// HashTable::DeleteBlock for a list of blocks (one device pass; see
// launch_delete), then the raycast's bitmaps are rebuilt.
static int delete_list(Sess &S, const int4 *d_req, int n, int64_t *deleted) {
  DevBuf dmask, touched, counts;
  DM_TRY(dev_alloc(dmask, (size_t)S.props.num_buckets * 8));
  DM_TRY(dev_alloc(touched, (size_t)std::max(n, 1) * 4));
  DM_TRY(dev_alloc(counts, 8));
  DM_HIP(hipMemsetAsync(dmask.p, 0, (size_t)S.props.num_buckets * 8, S.stream));
  DM_HIP(hipMemsetAsync(counts.p, 0, 8, S.stream));
  launch_delete(S.d_view, d_req, n, dmask.as<unsigned long long>(), touched.as<int>(), counts.as<int>(), S.stream);
  DM_HIP(hipGetLastError());
  int c[2] = {0, 0};
  DM_HIP(hipMemcpyAsync(c, counts.p, 8, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  if (c[1] > 0) {
    rebuild_far_maps(S.launch());
    DM_HIP(hipGetLastError());
    DM_HIP(hipStreamSynchronize(S.stream));
  }
  if (deleted) *deleted = c[1];
  return DM_OK;
}

int dm_volume_delete_blocks(dm_volume *vol, const int32_t *xyz, int64_t n, int64_t *deleted) {
  if (!vol || (n > 0 && !xyz) || n < 0 || n > (1LL << 28)) {
    set_error("dm_volume_delete_blocks: invalid argument");
    return DM_E_INVALID;
  }
  if (deleted) *deleted = 0;
  if (n == 0) return DM_OK;
  std::vector<int4> req((size_t)n);
  for (int64_t i = 0; i < n; ++i) req[i] = make_int4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0);
  DevBuf d_req;
  DM_TRY(dev_alloc(d_req, (size_t)n * sizeof(int4)));
  DM_HIP(hipMemcpyAsync(d_req.p, req.data(), (size_t)n * sizeof(int4), hipMemcpyHostToDevice, vol->stream));
  return delete_list(*vol, d_req.as<int4>(), (int)n, deleted);
}

int dm_volume_recycle(dm_volume *vol, const double center[3], double radius, int64_t *deleted) {
  if (!vol || !center || !(radius >= 0.0)) {
    set_error("dm_volume_recycle: invalid argument");
    return DM_E_INVALID;
  }
  Sess &S = *vol;
  DevBuf d_req, d_n;
  DM_TRY(dev_alloc(d_req, (size_t)S.props.num_blocks * sizeof(int4)));
  DM_TRY(dev_alloc(d_n, 4));
  DM_HIP(hipMemsetAsync(d_n.p, 0, 4, S.stream));
  launch_select_far(S.d_view, center, radius, d_req.as<int4>(), d_n.as<int>(), S.stream);
  DM_HIP(hipGetLastError());
  int n = 0;
  DM_HIP(hipMemcpyAsync(&n, d_n.p, 4, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  if (deleted) *deleted = 0;
  if (n == 0) return DM_OK;
  return delete_list(S, d_req.as<int4>(), n, deleted);
}

int dm_volume_num_entries(dm_volume *vol, int64_t *n) {
  if (!vol || !n) { set_error("null argument"); return DM_E_INVALID; }
  *n = (int64_t)vol->props.num_buckets * vol->props.bucket_size;
  return DM_OK;
}

int dm_volume_get_entries(dm_volume *vol, int32_t *out) {
  if (!vol || !out) { set_error("null argument"); return DM_E_INVALID; }
  const int B = vol->props.num_buckets, S = vol->props.bucket_size, st = vol->h.bstride;
  std::vector<uint8_t> raw((size_t)B * st);
  DM_HIP(hipMemcpyAsync(raw.data(), vol->h.buckets, raw.size(), hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  for (int b = 0; b < B; ++b) {
    const uint8_t *bk = raw.data() + (size_t)b * st;
    int cnt;
    memcpy(&cnt, bk + 12 * S, 4);
    for (int s = 0; s < S; ++s) {
      int32_t *o = out + 4 * ((size_t)b * S + s);
      if (s < cnt) {
        uint64_t k;
        int32_t p;
        memcpy(&k, bk + 8 * s, 8);
        memcpy(&p, bk + 8 * S + 4 * s, 4);
        auto sx = [](uint64_t v) { return (int)((int32_t)((uint32_t)(v & 0x1fffff) << 11) >> 11); };
        o[0] = sx(k >> 42); o[1] = sx(k >> 21); o[2] = sx(k); o[3] = p;
      } else {
        o[0] = 0; o[1] = 0; o[2] = 0; o[3] = -1;
      }
    }
  }
  return DM_OK;
}

int dm_volume_get_blocks(dm_volume *vol, int32_t cap, int32_t *xyz, float *sdf, uint8_t *rgb, uint8_t *weight,
                         int32_t *n_out) {
  if (!vol) { set_error("null argument"); return DM_E_INVALID; }
  int32_t c[C_COUNT];
  DM_HIP(hipMemcpyAsync(c, vol->h.ctr, sizeof c, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  const int hw = c[C_NUM_ALLOC], n = c[C_NUM_ALLOC] - c[C_NUM_FREE];
  if (n_out) *n_out = n;
  const int m = std::min(n, cap);
  if (m <= 0) return DM_OK;
  std::vector<int4> keys(hw);
  std::vector<uint64_t> vox((size_t)hw * kBlockVox);
  DM_HIP(hipMemcpyAsync(keys.data(), vol->h.block_key, (size_t)hw * 16, hipMemcpyDeviceToHost, vol->stream));
  if (sdf || rgb || weight)
    DM_HIP(hipMemcpyAsync(vox.data(), vol->h.vox, vox.size() * 8, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  // live blocks in pool order (freed pool indices skipped)
  for (int p = 0, b = 0; p < hw && b < m; ++p) {
    if (keys[p].w == 0) continue;
    if (xyz) { xyz[3 * b] = keys[p].x; xyz[3 * b + 1] = keys[p].y; xyz[3 * b + 2] = keys[p].z; }
    for (int i = 0; i < kBlockVox; ++i) {
      const size_t vi = (size_t)b * kBlockVox + i;
      const uint64_t w = vox[(size_t)p * kBlockVox + i];
      if (sdf) { uint32_t lo = (uint32_t)w; memcpy(&sdf[vi], &lo, 4); }
      if (rgb) { rgb[3 * vi] = (uint8_t)(w >> 32); rgb[3 * vi + 1] = (uint8_t)(w >> 40); rgb[3 * vi + 2] = (uint8_t)(w >> 48); }
      if (weight) weight[vi] = (uint8_t)(w >> 56);
    }
    ++b;
  }
  return DM_OK;
}

int dm_volume_counters(dm_volume *vol, int64_t out[5]) {
  if (!vol || !out) { set_error("null argument"); return DM_E_INVALID; }
  int32_t c[C_COUNT];
  DM_HIP(hipMemcpyAsync(c, vol->h.ctr, sizeof c, hipMemcpyDeviceToHost, vol->stream));
  DM_HIP(hipStreamSynchronize(vol->stream));
  out[0] = c[C_NUM_ALLOC] - c[C_NUM_FREE]; out[1] = c[C_DROP_BUCKET]; out[2] = c[C_DROP_POOL];
  out[3] = c[C_DROP_RANGE]; out[4] = c[C_DSET_OVF];
  return DM_OK;
}

// Snapshot of the whole reconstruction state: hash table, allocated blocks,
// block flags, pose state and the model maps.  buf NULL: *size only.
int dm_volume_snapshot(dm_volume *vol, void *buf, uint64_t cap, uint64_t *size) {
  if (!vol || !size) { set_error("null argument"); return DM_E_INVALID; }
  Sess &S = *vol;
  SnapHeader H{};
  memcpy(H.magic, kSnapMagic, 8);
  H.version = kSnapVersion;
  H.hdr_bytes = sizeof(SnapHeader);
  H.props = S.props;
  H.max_rows = S.max_rows; H.max_cols = S.max_cols; H.rows = S.h.rows; H.cols = S.h.cols;
  H.fx = S.h.fx; H.fy = S.h.fy; H.cx = S.h.cx; H.cy = S.h.cy;
  DM_HIP(hipMemcpyAsync(H.ctr, S.h.ctr, sizeof H.ctr, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  H.num_alloc = H.ctr[C_NUM_ALLOC];
  for (int l = 0; l < 3; ++l) H.iters[l] = S.h.iters[l];
  const float f7[7] = {S.h.icp_dmin, S.h.icp_dmax, S.h.dist_max, S.h.w_photo, S.h.lambda, S.h.max_inc,
                       S.h.min_inc};
  memcpy(H.icp_f, f7, sizeof f7);
  const auto secs = snap_sections(S, H.num_alloc, H.ctr[C_NUM_FREE]);
  uint64_t total = sizeof(SnapHeader);
  for (const auto &x : secs) total += x.second;
  H.payload = total - sizeof(SnapHeader);
  *size = total;
  if (!buf) return DM_OK;
  if (cap < total) { set_error("dm_volume_snapshot: buffer too small"); return DM_E_CAPACITY; }
  uint8_t *o = (uint8_t *)buf;
  memcpy(o, &H, sizeof H);
  o += sizeof H;
  for (const auto &x : secs) {
    DM_HIP(hipMemcpyAsync(o, x.first, x.second, hipMemcpyDeviceToHost, S.stream));
    o += x.second;
  }
  DM_HIP(hipStreamSynchronize(S.stream));
  return DM_OK;
}

// Restore a snapshot into a volume / session created with the same
// properties and frame size; the per-frame scratch is reset.
int dm_volume_restore(dm_volume *vol, const void *buf, uint64_t size) {
  if (!vol || !buf || size < sizeof(SnapHeader)) { set_error("invalid argument"); return DM_E_INVALID; }
  Sess &S = *vol;
  SnapHeader H;
  memcpy(&H, buf, sizeof H);
  if (memcmp(H.magic, kSnapMagic, 8) != 0 || H.version != kSnapVersion || H.hdr_bytes != sizeof(SnapHeader)) {
    set_error("dm_volume_restore: not a snapshot of this format");
    return DM_E_INVALID;
  }
  if (memcmp(&H.props, &S.props, sizeof(dm_tsdf_props)) != 0 || H.max_rows != S.max_rows ||
      H.max_cols != S.max_cols || H.num_alloc < 0 || H.num_alloc > S.props.num_blocks ||
      H.ctr[C_NUM_FREE] < 0 || H.ctr[C_NUM_FREE] > H.num_alloc || H.rows > S.max_rows || H.cols > S.max_cols) {
    set_error("dm_volume_restore: properties or frame size differ from the snapshot's");
    return DM_E_INVALID;
  }
  S.h.rows = (int)H.rows; S.h.cols = (int)H.cols;
  S.h.fx = H.fx; S.h.fy = H.fy; S.h.cx = H.cx; S.h.cy = H.cy;
  const auto secs = snap_sections(S, H.num_alloc, H.ctr[C_NUM_FREE]);
  uint64_t total = sizeof(SnapHeader);
  for (const auto &x : secs) total += x.second;
  if (size != total || H.payload != total - sizeof(SnapHeader)) {
    set_error("dm_volume_restore: size does not match the snapshot header");
    return DM_E_INVALID;
  }
  const uint8_t *p = (const uint8_t *)buf + sizeof H;
  for (const auto &x : secs) {
    DM_HIP(hipMemcpyAsync(x.first, p, x.second, hipMemcpyHostToDevice, S.stream));
    p += x.second;
  }
  DM_HIP(hipMemcpyAsync(S.h.ctr, H.ctr, sizeof H.ctr, hipMemcpyHostToDevice, S.stream));
  // blocks beyond the snapshot's are unused again
  const size_t rest = (size_t)(S.props.num_blocks - H.num_alloc);
  if (rest) {
    DM_HIP(hipMemsetAsync(S.h.vox + (size_t)H.num_alloc * kBlockVox, 0, rest * kBlockVox * 8, S.stream));
    DM_HIP(hipMemsetAsync(S.h.blk_near + H.num_alloc, 0, rest, S.stream));
  }
  DM_HIP(hipMemsetAsync(S.h.bucket_head, 0xff, (size_t)S.props.num_buckets * 4, S.stream));
  for (int l = 0; l < 3; ++l) S.h.iters[l] = H.iters[l];
  S.h.icp_dmin = H.icp_f[0]; S.h.icp_dmax = H.icp_f[1]; S.h.dist_max = H.icp_f[2];
  S.h.w_photo = H.icp_f[3]; S.h.lambda = H.icp_f[4]; S.h.max_inc = H.icp_f[5]; S.h.min_inc = H.icp_f[6];
  DM_TRY(S.push_view());
  rebuild_far_maps(S.launch());
  DM_HIP(hipGetLastError());
  DM_HIP(hipStreamSynchronize(S.stream));
  return DM_OK;
}

// ---------------------------------------------------------------- session
static int apply_icp(Sess &S, const dm_icp_props *icp) {
  if (!icp) return DM_OK;
  for (int l = 0; l < 3; ++l) {
    if (icp->lvl_iterations[l] < 0 || icp->lvl_iterations[l] > 1000) {
      set_error("invalid iteration count");
      return DM_E_INVALID;
    }
    S.h.iters[l] = icp->lvl_iterations[l];
  }
  S.icp = *icp;
  S.h.icp_dmin = icp->min_depth;
  S.h.icp_dmax = icp->max_depth;
  S.h.dist_max = icp->dist_max;
  S.h.w_photo = icp->photo_weight;
  S.h.lambda = icp->regularization;
  S.h.max_inc = icp->max_increment;
  S.h.min_inc = icp->min_increment;
  return DM_OK;
}

int dm_session_create(const dm_tsdf_props *props, const dm_icp_props *icp, const dm_sensor *sensor,
                      const double init_pose[16], dm_session **out) {
  if (!props || !sensor || !out) { set_error("null argument"); return DM_E_INVALID; }
  dm_session *s = new dm_session();
  int r = s->init(*props, sensor->rows, sensor->cols);
  if (r == DM_OK) r = s->set_sensor(sensor);
  if (r == DM_OK) r = apply_icp(*s, icp);
  if (r == DM_OK && init_pose) {
    hipError_t e = hipMemcpyAsync(s->h.posed + PD_TINIT, init_pose, 16 * sizeof(double), hipMemcpyHostToDevice,
                                  s->stream);
    if (e != hipSuccess) r = hip_fail(e, "hipMemcpyAsync(init pose)");
  }
  if (r == DM_OK) r = s->push_view();
  if (r != DM_OK) { s->release(); delete s; return r; }
  *out = s;
  return DM_OK;
}

int dm_session_destroy(dm_session *s) {
  if (!s) return DM_OK;
  s->release();
  delete s;
  return DM_OK;
}

dm_volume *dm_session_volume(dm_session *s) { return reinterpret_cast<dm_volume *>(s); }

int dm_session_process_frame(dm_session *s, const float *host_depth, const uint8_t *host_rgb, int32_t solve_on_host,
                             double pose_out[16]) {
  if (!s || !host_depth) { set_error("null argument"); return DM_E_INVALID; }
  DM_TRY(s->stage_frame(host_depth, host_rgb));
  DM_TRY(s->push_view());
  return process_frame(*s, solve_on_host != 0, pose_out);
}

int dm_session_process_frame_dev(dm_session *s, const float *dev_depth, const uint8_t *dev_rgb, int32_t solve_on_host,
                                 double pose_out[16]) {
  if (!s || !dev_depth) { set_error("null argument"); return DM_E_INVALID; }
  s->h.depth = dev_depth;
  s->h.rgb = dev_rgb;
  DM_TRY(s->push_view());
  return process_frame(*s, solve_on_host != 0, pose_out);
}

int dm_session_set_depth_filter(dm_session *s, int32_t radius, float sigma_depth, float sigma_space) {
  if (!s || radius < 0 || radius > 7 || (radius > 0 && (!(sigma_depth > 0.f) || !(sigma_space > 0.f)))) {
    set_error("dm_session_set_depth_filter: invalid argument (0 <= radius <= 7, sigmas > 0)");
    return DM_E_INVALID;
  }
  if (radius > 0 && !s->d_depth_f) DM_TRY(s->dalloc(&s->d_depth_f, (size_t)s->max_rows * s->max_cols));
  s->filt_r = radius;
  s->filt_sd = sigma_depth;
  s->filt_ss = sigma_space;
  s->h.depth_track = radius > 0 ? s->d_depth_f : nullptr;
  DM_TRY(s->push_view());
  DM_HIP(hipStreamSynchronize(s->stream));
  return DM_OK;
}

int dm_session_get_pose(dm_session *s, double pose[16]) {
  if (!s || !pose) { set_error("null argument"); return DM_E_INVALID; }
  DM_HIP(hipMemcpyAsync(pose, s->h.posed + PD_T, 16 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  return DM_OK;
}

int dm_session_source_maps(dm_session *s, const float *host_depth, const uint8_t *host_rgb, int32_t level,
                           float *host_vertex, float *host_normal, float *host_intensity) {
  if (!s || !host_depth || level < 0 || level >= kLevels) { set_error("invalid argument"); return DM_E_INVALID; }
  DM_TRY(s->stage_frame(host_depth, host_rgb));
  DM_TRY(s->push_view());
  launch_source_maps(s->launch());
  DM_HIP(hipGetLastError());
  const size_t P = (size_t)(s->h.rows >> level) * (s->h.cols >> level);
  std::vector<float4> vt(P), nt(P);
  DM_HIP(hipMemcpyAsync(vt.data(), s->h.svert[level], P * 16, hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipMemcpyAsync(nt.data(), s->h.snorm[level], P * 16, hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  for (size_t i = 0; i < P; ++i) {
    if (host_vertex) { host_vertex[3 * i] = vt[i].x; host_vertex[3 * i + 1] = vt[i].y; host_vertex[3 * i + 2] = vt[i].z; }
    if (host_normal) { host_normal[3 * i] = nt[i].x; host_normal[3 * i + 1] = nt[i].y; host_normal[3 * i + 2] = nt[i].z; }
    if (host_intensity) host_intensity[i] = vt[i].w;
  }
  return DM_OK;
}

int dm_session_set_model_maps(dm_session *s, int32_t level, const float *host_vertex, const float *host_normal,
                              const float *host_intensity) {
  if (!s || !host_vertex || !host_normal || level < 0 || level >= kLevels) {
    set_error("invalid argument");
    return DM_E_INVALID;
  }
  const size_t P = (size_t)(s->h.rows >> level) * (s->h.cols >> level);
  std::vector<float4> vt(P), nt(P);
  for (size_t i = 0; i < P; ++i) {
    vt[i] = make_float4(host_vertex[3 * i], host_vertex[3 * i + 1], host_vertex[3 * i + 2],
                        host_intensity ? host_intensity[i] : 0.f);
    nt[i] = make_float4(host_normal[3 * i], host_normal[3 * i + 1], host_normal[3 * i + 2], 0.f);
  }
  DM_HIP(hipMemcpyAsync(s->h.tvert[level], vt.data(), P * 16, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipMemcpyAsync(s->h.tnorm[level], nt.data(), P * 16, hipMemcpyHostToDevice, s->stream));
  int one = 1;
  DM_HIP(hipMemcpyAsync(s->h.ctr + C_HAS_MODEL, &one, 4, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  return DM_OK;
}

int dm_session_set_icp_reduce(dm_session *s, int32_t variant) {
  if (!s || (variant != 0 && variant != 1)) { set_error("ICP variant is 0 (shuffle tree) or 1 (matrix cores)"); return DM_E_INVALID; }
  s->icp_mfma = variant;
  return DM_OK;
}

int dm_session_icp_system(dm_session *s, int32_t level, const double Tinc[16], double A[36], double b[6],
                          double sums[3]) {
  if (!s || !Tinc || level < 0 || level >= kLevels) { set_error("invalid argument"); return DM_E_INVALID; }
  float tf[16];
  for (int k = 0; k < 16; ++k) tf[k] = (float)Tinc[k];
  DM_HIP(hipMemcpyAsync(s->h.posef + PF_TINC, tf, sizeof tf, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipMemcpyAsync(s->h.posed + PD_TINC, Tinc, 16 * sizeof(double), hipMemcpyHostToDevice, s->stream));
  int it1 = 1;
  if (s->h.iters[level] < 1) { set_error("level has no iterations"); return DM_E_STATE; }
  launch_icp_reduce(s->launch(), level, it1, false);
  DM_HIP(hipGetLastError());
  const int P = (s->h.rows >> level) * (s->h.cols >> level);
  const int nwg = icp_wg_count(P);
  std::vector<float> part((size_t)nwg * kIcpVals);
  DM_HIP(hipMemcpyAsync(part.data(), s->h.partials, part.size() * 4, hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  double sm[30];
  icp_sum_records(part.data(), nwg, sm);
  if (A) sym6_from_upper(sm, A);
  if (b) for (int k = 0; k < 6; ++k) b[k] = sm[21 + k];
  if (sums) { sums[0] = sm[27]; sums[1] = sm[28]; sums[2] = sm[29]; }
  return DM_OK;
}

int dm_session_icp_track(dm_session *s, int32_t solve_on_host, double Tinc_out[16]) {
  if (!s) { set_error("null argument"); return DM_E_INVALID; }
  double pd[PD_COUNT] = {0};
  for (int k = 0; k < 16; ++k) {
    const double I = (k % 5 == 0) ? 1.0 : 0.0;
    pd[PD_T + k] = I; pd[PD_TPREV + k] = I; pd[PD_TINC + k] = I; pd[PD_TINIT + k] = I;
  }
  float pf[PF_COUNT];
  for (int k = 0; k < PF_COUNT; ++k) pf[k] = (k % 16) % 5 == 0 ? 1.f : 0.f;
  DM_HIP(hipMemcpyAsync(s->h.posed, pd, sizeof pd, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipMemcpyAsync(s->h.posef, pf, sizeof pf, hipMemcpyHostToDevice, s->stream));
  const int one = 1, zero = 0;
  DM_HIP(hipMemcpyAsync(s->h.ctr + C_HAS_MODEL, &one, 4, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipMemcpyAsync(s->h.ctr + C_ICP_STOP, &zero, 4, hipMemcpyHostToDevice, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  DM_TRY(run_icp(*s, solve_on_host != 0));
  DM_HIP(hipMemcpyAsync(pd, s->h.posed, sizeof pd, hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  if (Tinc_out) memcpy(Tinc_out, pd + PD_TINC, 16 * sizeof(double));
  return DM_OK;
}

int dm_session_icp_increment(dm_session *s, double twist[6]) {
  if (!s || !twist) { set_error("null argument"); return DM_E_INVALID; }
  double acc[6];
  DM_HIP(hipMemcpyAsync(acc, s->h.posed + PD_ACC, sizeof acc, hipMemcpyDeviceToHost, s->stream));
  DM_HIP(hipStreamSynchronize(s->stream));
  // gnSolver.cu:426-428: a NaN, infinite or too large sum is replaced by zero
  const double ssum = acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5];
  const bool reset = ssum != ssum || std::isinf(ssum) || ssum >= (double)s->h.max_inc;
  for (int k = 0; k < 6; ++k) twist[k] = reset ? 0.0 : acc[k];
  return DM_OK;
}

}  // extern "C"
