// Serving loop: many concurrent reconstruction sessions on one GPU.
//
// No reference counterpart (the reference runs one synchronous frame at a
// time, src/dynaMap.cpp:218-301).  Each admitted session keeps its hash
// table, block pool, pose and model maps resident in HBM.  A batch ("tick")
// takes the next arrived frame of every session that has one and runs the
// whole frame pipeline for all of them in the same launches: blockIdx.y of
// every kernel indexes the batch's session list, unused list entries are -1.
// The launch sequence of a tick is captured once in a HIP graph and replayed;
// per tick only the session table and the list are copied (one small H2D).
// Sessions are admitted in request order while their state fits the HBM
// budget and a slot is free, and retired when their frames are done, which
// frees budget for waiting requests.
#include <string.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>
#include "dm_session.h"

namespace dm {

struct Trace {
  dm_tsdf_props props{};
  dm_icp_props icp{};
  dm_sensor sensor{};
  double init_pose[16];
  int nframes = 0;
  float *d_depth = nullptr;
  uint8_t *d_rgb = nullptr;
  size_t bytes = 0;
  Sess *sess = nullptr;
  int slot = -1;
  int next_frame = -1;  // trace frame the live session continues at; -1 when it cannot continue
};

int sess_reset(Sess &S, const double init_pose[16]);

}  // namespace dm

using namespace dm;

struct dm_server {
  dm_server_config cfg{};
  std::vector<Trace> traces;
  std::vector<Sess *> pool;          // retired sessions kept for reuse
  hipStream_t stream = nullptr;
  DmView *d_table = nullptr;
  int *d_list = nullptr;
  DmView *h_table[2] = {nullptr, nullptr};  // pinned, double buffered
  int *h_list[2] = {nullptr, nullptr};
  hipEvent_t ev_copy[2] = {nullptr, nullptr};
  std::vector<int> slot_trace;       // slot -> trace or -1
  uint64_t used_bytes = 0;
  hipGraphExec_t graph = nullptr;
  int graph_iters[3] = {0, 0, 0};
  int ray_tw_log2 = 4;               // launch shapes (dm_server_autotune, dm_server_set_shapes)
  int integ_wg = 2048;
  int icp_mfma = 0;                  // ICP linearisation: 0 shuffle tree, 1 matrix cores
};

namespace dm {

int sess_reset(Sess &S, const double init_pose[16]) {
  int n = 0;
  DM_HIP(hipMemcpyAsync(&n, S.h.ctr + C_NUM_ALLOC, 4, hipMemcpyDeviceToHost, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  n = std::min(std::max(n, 0), S.h.num_blocks);
  DM_HIP(hipMemsetAsync(S.h.buckets, 0, (size_t)S.h.num_buckets * S.h.bstride, S.stream));
  DM_HIP(hipMemsetAsync(S.h.vox, 0, (size_t)n * kBlockVox * 8, S.stream));
  DM_HIP(hipMemsetAsync(S.h.blk_near, 0, (size_t)n, S.stream));
  DM_HIP(hipMemsetAsync(S.h.nearmap, 0, kWinWords * 4, S.stream));
  DM_HIP(hipMemsetAsync(S.h.dilmap, 0, kWinWords * 4, S.stream));
  DM_HIP(hipMemsetAsync(S.h.allocmap, 0, kWinWords * 4, S.stream));
  DM_HIP(hipMemsetAsync(S.h.coarse, 0, (size_t)kCoarseWords * 4, S.stream));
  DM_HIP(hipMemsetAsync(S.h.cell4, 0, (size_t)kCell4Words * 4, S.stream));
  DM_HIP(hipMemsetAsync(S.h.ctr, 0, C_COUNT * 4, S.stream));
  double pd[PD_COUNT] = {0};
  for (int k = 0; k < 16; ++k) {
    const double I = (k % 5 == 0) ? 1.0 : 0.0;
    pd[PD_T + k] = I; pd[PD_TPREV + k] = I; pd[PD_TINC + k] = I;
    pd[PD_TINIT + k] = init_pose ? init_pose[k] : I;
  }
  DM_HIP(hipMemcpyAsync(S.h.posed, pd, sizeof pd, hipMemcpyHostToDevice, S.stream));
  DM_HIP(hipStreamSynchronize(S.stream));
  S.frames = 0;
  return DM_OK;
}

static bool same_shape(const Sess &S, const Trace &t) {
  return memcmp(&S.props, &t.props, sizeof(dm_tsdf_props)) == 0 && S.max_rows == t.sensor.rows &&
         S.max_cols == t.sensor.cols;
}

static int apply_trace(Sess &S, const Trace &t) {
  S.h.rows = (int)t.sensor.rows; S.h.cols = (int)t.sensor.cols;
  S.h.fx = t.sensor.fx; S.h.fy = t.sensor.fy; S.h.cx = t.sensor.cx; S.h.cy = t.sensor.cy;
  for (int l = 0; l < 3; ++l) S.h.iters[l] = t.icp.lvl_iterations[l];
  S.h.icp_dmin = t.icp.min_depth; S.h.icp_dmax = t.icp.max_depth; S.h.dist_max = t.icp.dist_max;
  S.h.w_photo = t.icp.photo_weight; S.h.lambda = t.icp.regularization; S.h.max_inc = t.icp.max_increment;
  S.h.min_inc = t.icp.min_increment;
  S.icp = t.icp;
  return DM_OK;
}

// Admit waiting traces in request order while slots and budget allow.
static int admit(dm_server *srv, const std::vector<int> &want) {
  for (int ti : want) {
    Trace &t = srv->traces[ti];
    if (t.sess) continue;
    int slot = -1;
    for (int s = 0; s < srv->cfg.max_sessions; ++s)
      if (srv->slot_trace[s] < 0) { slot = s; break; }
    if (slot < 0 || srv->used_bytes + t.bytes > srv->cfg.hbm_budget_bytes) break;  // FIFO: no overtaking
    Sess *S = nullptr;
    for (size_t k = 0; k < srv->pool.size(); ++k)
      if (same_shape(*srv->pool[k], t)) { S = srv->pool[k]; srv->pool.erase(srv->pool.begin() + k); break; }
    if (S) {
      int r = sess_reset(*S, t.init_pose);
      if (r != DM_OK) return r;
    } else {
      // free pooled sessions of other shapes if their memory is needed
      while (!srv->pool.empty()) {
        Sess *o = srv->pool.back();
        srv->pool.pop_back();
        o->release();
        delete o;
      }
      S = new Sess();
      int r = S->init(t.props, t.sensor.rows, t.sensor.cols);
      if (r == DM_OK) r = sess_reset(*S, t.init_pose);
      if (r != DM_OK) { S->release(); delete S; return r; }
    }
    apply_trace(*S, t);
    t.sess = S;
    t.slot = slot;
    srv->slot_trace[slot] = ti;
    srv->used_bytes += t.bytes;
  }
  return DM_OK;
}

static void retire(dm_server *srv, int ti) {
  Trace &t = srv->traces[ti];
  if (!t.sess) return;
  srv->slot_trace[t.slot] = -1;
  srv->used_bytes -= t.bytes;
  srv->pool.push_back(t.sess);
  t.sess = nullptr;
  t.slot = -1;
  t.next_frame = -1;
}

// Dispatch order of a batch: every batched kernel maps blockIdx.y to the
// list's sessions in order, and workgroups are dispatched x fastest, so the
// sessions whose rays cost most (pixels times depth range over voxel size)
// go first and the launch's tail is made of cheaper tiles.  Each session's
// results do not depend on its place in the list.
static void order_by_cost(const dm_server *srv, std::vector<int> &act) {
  auto cost = [srv](int t) {
    const Trace &tr = srv->traces[t];
    return (double)tr.sensor.rows * tr.sensor.cols *
           ((double)tr.props.max_sensor_depth - tr.props.min_sensor_depth) / tr.props.voxel_size;
  };
  std::stable_sort(act.begin(), act.end(), [&](int a, int b) { return cost(a) > cost(b); });
}

static Launch server_launch(const dm_server *srv) {
  Launch L;
  L.views = srv->d_table;
  L.active = srv->d_list;
  L.nact = srv->cfg.max_sessions;
  L.max_pixels = (int)(srv->cfg.max_rows * srv->cfg.max_cols);
  L.max_tiles = (int)(((srv->cfg.max_cols + 15) / 16) * ((srv->cfg.max_rows + 15) / 16));
  L.max_rows = (int)srv->cfg.max_rows;
  L.max_cols = (int)srv->cfg.max_cols;
  L.ray_tw_log2 = srv->ray_tw_log2;
  L.integ_wg = srv->integ_wg;
  L.icp_mfma = srv->icp_mfma;
  L.stream = srv->stream;
  return L;
}

// The batched frame pipeline (same stages as the per-frame path).
static void enqueue_pipeline(const dm_server *srv, const int iters[3]) {
  const Launch L = server_launch(srv);
  launch_pose_begin(L);
  launch_source_maps(L);
  for (int l = kLevels - 1; l >= 0; --l)
    for (int it = 1; it <= iters[l]; ++it) {
      launch_icp_reduce(L, l, it, true);
    }
  launch_alloc(L);
  launch_integrate(L);
  launch_raycast(L);
}

static int ensure_graph(dm_server *srv, const int iters[3]) {
  if (!srv->cfg.use_graph) return DM_OK;
  if (srv->graph && memcmp(iters, srv->graph_iters, sizeof(srv->graph_iters)) == 0) return DM_OK;
  if (srv->graph) { hipGraphExecDestroy(srv->graph); srv->graph = nullptr; }
  hipGraph_t g = nullptr;
  DM_HIP(hipStreamBeginCapture(srv->stream, hipStreamCaptureModeThreadLocal));
  enqueue_pipeline(srv, iters);
  DM_HIP(hipStreamEndCapture(srv->stream, &g));
  DM_HIP(hipGraphInstantiate(&srv->graph, g, nullptr, nullptr, 0));
  hipGraphDestroy(g);
  memcpy(srv->graph_iters, iters, sizeof(srv->graph_iters));
  return DM_OK;
}

using Clock = std::chrono::steady_clock;
static double secs_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace dm

extern "C" {

int dm_server_create(const dm_server_config *cfg, dm_server **out) {
  if (!cfg || !out || cfg->max_sessions <= 0 || cfg->max_sessions > 1024 || cfg->max_rows == 0 ||
      cfg->max_cols == 0) {
    set_error("invalid server config");
    return DM_E_INVALID;
  }
  dm_server *s = new dm_server();
  s->cfg = *cfg;
  s->slot_trace.assign(cfg->max_sessions, -1);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&s->d_table, sizeof(DmView) * cfg->max_sessions);
  if (e == hipSuccess) e = hipMalloc(&s->d_list, sizeof(int) * 2 * cfg->max_sessions);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = hipHostMalloc((void **)&s->h_table[k], sizeof(DmView) * cfg->max_sessions, hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&s->h_list[k], sizeof(int) * 2 * cfg->max_sessions, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_copy[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    int r = hip_fail(e, "server setup");
    dm_server_destroy(s);
    return r;
  }
  *out = s;
  return DM_OK;
}

int dm_server_reset(dm_server *srv) {
  if (!srv) return DM_OK;
  for (size_t t = 0; t < srv->traces.size(); ++t) retire(srv, (int)t);
  for (Sess *S : srv->pool) { S->release(); delete S; }
  srv->pool.clear();
  return DM_OK;
}

int dm_server_destroy(dm_server *srv) {
  if (!srv) return DM_OK;
  if (srv->stream) hipStreamSynchronize(srv->stream);
  dm_server_reset(srv);
  for (Trace &t : srv->traces) {
    if (t.d_depth) hipFree(t.d_depth);
    if (t.d_rgb) hipFree(t.d_rgb);
  }
  if (srv->graph) hipGraphExecDestroy(srv->graph);
  if (srv->d_table) hipFree(srv->d_table);
  if (srv->d_list) hipFree(srv->d_list);
  for (int k = 0; k < 2; ++k) {
    if (srv->h_table[k]) hipHostFree(srv->h_table[k]);
    if (srv->h_list[k]) hipHostFree(srv->h_list[k]);
    if (srv->ev_copy[k]) hipEventDestroy(srv->ev_copy[k]);
  }
  if (srv->stream) hipStreamDestroy(srv->stream);
  delete srv;
  return DM_OK;
}

int dm_server_add_trace(dm_server *srv, const dm_tsdf_props *props, const dm_icp_props *icp, const dm_sensor *sensor,
                        const double init_pose[16], int32_t nframes, const float *host_depth,
                        const uint8_t *host_rgb, int32_t *trace_id) {
  if (!srv || !props || !icp || !sensor || !init_pose || nframes <= 0 || !host_depth) {
    set_error("invalid trace");
    return DM_E_INVALID;
  }
  int r = validate_props(*props);
  if (r != DM_OK) return r;
  if (sensor->rows > srv->cfg.max_rows || sensor->cols > srv->cfg.max_cols) {
    set_error("trace frame larger than the server's max frame");
    return DM_E_INVALID;
  }
  Trace t;
  t.props = *props;
  t.icp = *icp;
  t.sensor = *sensor;
  memcpy(t.init_pose, init_pose, sizeof t.init_pose);
  t.nframes = nframes;
  t.bytes = session_bytes(*props, sensor->rows, sensor->cols);
  const size_t P = (size_t)sensor->rows * sensor->cols;
  DM_HIP(hipMalloc(&t.d_depth, P * 4 * nframes));
  DM_HIP(hipMemcpy(t.d_depth, host_depth, P * 4 * nframes, hipMemcpyHostToDevice));
  if (host_rgb) {
    DM_HIP(hipMalloc(&t.d_rgb, P * 3 * nframes));
    DM_HIP(hipMemcpy(t.d_rgb, host_rgb, P * 3 * nframes, hipMemcpyHostToDevice));
  }
  srv->traces.push_back(t);
  if (trace_id) *trace_id = (int32_t)srv->traces.size() - 1;
  return DM_OK;
}

int dm_server_trace_bytes(dm_server *srv, int32_t trace_id, uint64_t *bytes) {
  if (!srv || trace_id < 0 || trace_id >= (int)srv->traces.size() || !bytes) {
    set_error("invalid trace id");
    return DM_E_INVALID;
  }
  *bytes = srv->traces[trace_id].bytes;
  return DM_OK;
}

int dm_server_run(dm_server *srv, int32_t mode, int32_t first_frame, int32_t nframes, const double *arrivals,
                  double *completions, double *poses, double *wall_seconds, int64_t *batches) {
  if (!srv || nframes <= 0 || first_frame < 0 || mode < 0 || mode > 2) {
    set_error("invalid run arguments");
    return DM_E_INVALID;
  }
  const int NT = (int)srv->traces.size();
  for (int t = 0; t < NT; ++t)
    if (first_frame + nframes > srv->traces[t].nframes) {
      set_error("trace has fewer frames than requested");
      return DM_E_INVALID;
    }
  // fresh sessions when starting from frame 0; otherwise every session must
  // be live and have stopped exactly at first_frame
  if (first_frame == 0)
    for (int t = 0; t < NT; ++t) retire(srv, t);
  else
    for (int t = 0; t < NT; ++t)
      if (!srv->traces[t].sess || srv->traces[t].next_frame != first_frame) {
        set_error("a run from frame " + std::to_string(first_frame) + " must continue where trace " +
                  std::to_string(t) + " stopped (" +
                  (srv->traces[t].sess ? "at frame " + std::to_string(srv->traces[t].next_frame)
                                       : std::string("no live session")) + ")");
        return DM_E_STATE;
      }
  for (int t = 0; t < NT; ++t) srv->traces[t].next_frame = -1;  // set again when the run completes
  std::vector<int> order(NT);
  for (int t = 0; t < NT; ++t) order[t] = t;
  int r = admit(srv, order);
  if (r != DM_OK) return r;
  std::vector<int> next(NT, 0);  // frames done in this call
  int64_t nbatch = 0;
  auto arrival = [&](int t, int f) { return arrivals ? arrivals[(size_t)t * nframes + f] : 0.0; };
  auto frame_ptrs = [&](int t, int f, const float *&d, const uint8_t *&c) {
    const Trace &tr = srv->traces[t];
    const size_t P = (size_t)tr.sensor.rows * tr.sensor.cols;
    d = tr.d_depth + P * (size_t)(first_frame + f);
    c = tr.d_rgb ? tr.d_rgb + P * 3 * (size_t)(first_frame + f) : nullptr;
  };
  int remaining = NT * nframes;
  // the clock (the origin of arrivals and completions, and of the wall time)
  // starts after the set-up: graph capture and the pose log's allocation
  Clock::time_point t0 = Clock::now();
  double wall = 0.0;

  if (mode == DM_RUN_BATCHED) {
    int iters[3] = {0, 0, 0};
    for (int t = 0; t < NT; ++t)
      for (int l = 0; l < 3; ++l) iters[l] = std::max(iters[l], srv->traces[t].icp.lvl_iterations[l]);
    r = ensure_graph(srv, iters);
    if (r != DM_OK) return r;
    double *d_log = nullptr;
    if (poses) DM_HIP(hipMalloc(&d_log, sizeof(double) * 16 * (size_t)NT * nframes));
    DM_HIP(hipStreamSynchronize(srv->stream));
    t0 = Clock::now();
    hipEvent_t ev_start;
    DM_HIP(hipEventCreate(&ev_start));
    DM_HIP(hipEventRecord(ev_start, srv->stream));
    struct Pending { hipEvent_t ev; std::vector<std::pair<int, int>> frames; };
    std::vector<Pending> inflight;
    int buf = 0;
    auto drain = [&](bool all) -> int {
      while (!inflight.empty()) {
        Pending &p = inflight.front();
        if (!all && hipEventQuery(p.ev) == hipErrorNotReady) break;
        DM_HIP(hipEventSynchronize(p.ev));
        float ms = 0.f;
        DM_HIP(hipEventElapsedTime(&ms, ev_start, p.ev));
        for (auto &fr : p.frames)
          if (completions) completions[(size_t)fr.first * nframes + fr.second] = ms * 1e-3;
        hipEventDestroy(p.ev);
        inflight.erase(inflight.begin());
      }
      return DM_OK;
    };
    std::vector<int> pose_pending;  // traces whose last frame pose must be read before retiring
    while (remaining > 0) {
      const double now = secs_since(t0);
      std::vector<int> act;
      for (int t = 0; t < NT; ++t) {
        Trace &tr = srv->traces[t];
        if (!tr.sess || next[t] >= nframes) continue;
        if (arrival(t, next[t]) <= now) act.push_back(t);
      }
      order_by_cost(srv, act);
      if (act.empty()) {
        // wait for the earliest arrival or for an admission opportunity
        r = drain(false);
        if (r != DM_OK) return r;
        double nxt = 1e30;
        for (int t = 0; t < NT; ++t)
          if (srv->traces[t].sess && next[t] < nframes) nxt = std::min(nxt, arrival(t, next[t]));
        if (nxt > now && nxt < 1e29) {
          const double dt = std::min(nxt - now, 0.0005);
          std::this_thread::sleep_for(std::chrono::duration<double>(dt));
        } else if (nxt >= 1e29) {
          // nothing admitted has work: wait for in-flight work, retire, admit
          r = drain(true);
          if (r != DM_OK) return r;
          std::vector<int> waiting;
          for (int t = 0; t < NT; ++t)
            if (!srv->traces[t].sess && next[t] < nframes) waiting.push_back(t);
          r = admit(srv, waiting);
          if (r != DM_OK) return r;
          bool any = false;
          for (int t = 0; t < NT; ++t) any |= srv->traces[t].sess && next[t] < nframes;
          if (!any) { set_error("a trace does not fit the HBM budget"); return DM_E_CAPACITY; }
        }
        continue;
      }
      // build the session table and list for this batch
      DM_HIP(hipEventSynchronize(srv->ev_copy[buf]));
      DmView *tab = srv->h_table[buf];
      int *lst = srv->h_list[buf];
      for (int s = 0; s < srv->cfg.max_sessions; ++s) lst[s] = -1;
      Pending p;
      int k = 0;
      for (int t : act) {
        Trace &tr = srv->traces[t];
        Sess &S = *tr.sess;
        const float *d;
        const uint8_t *c;
        frame_ptrs(t, next[t], d, c);
        S.h.depth = d;
        S.h.rgb = c;
        tab[tr.slot] = S.h;
        lst[k++] = tr.slot;
        p.frames.push_back({t, next[t]});
      }
      int *dst = lst + srv->cfg.max_sessions;
      for (int q = 0; q < (int)act.size(); ++q) dst[q] = act[q] * nframes + next[act[q]];
      DM_HIP(hipMemcpyAsync(srv->d_table, tab, sizeof(DmView) * srv->cfg.max_sessions, hipMemcpyHostToDevice,
                            srv->stream));
      DM_HIP(hipMemcpyAsync(srv->d_list, lst, sizeof(int) * 2 * srv->cfg.max_sessions, hipMemcpyHostToDevice,
                            srv->stream));
      DM_HIP(hipEventRecord(srv->ev_copy[buf], srv->stream));
      buf ^= 1;
      if (srv->graph) {
        DM_HIP(hipGraphLaunch(srv->graph, srv->stream));
      } else {
        enqueue_pipeline(srv, iters);
      }
      if (poses) launch_pose_log(server_launch(srv), srv->d_list + srv->cfg.max_sessions, d_log);
      DM_HIP(hipGetLastError());
      DM_HIP(hipEventCreate(&p.ev));
      DM_HIP(hipEventRecord(p.ev, srv->stream));
      ++nbatch;
      for (int t : act) {
        next[t]++;
        remaining--;
      }
      inflight.push_back(std::move(p));
      // bound the queue of launched batches; retire finished traces
      if (inflight.size() >= 3) {
        r = drain(false);
        if (r != DM_OK) return r;
        if (inflight.size() >= 3) {
          DM_HIP(hipEventSynchronize(inflight.front().ev));
          r = drain(false);
          if (r != DM_OK) return r;
        }
      }
      bool retired = false;
      for (int t = 0; t < NT; ++t) {
        Trace &tr = srv->traces[t];
        if (tr.sess && next[t] >= nframes && first_frame + nframes >= tr.nframes) {
          r = drain(true);
          if (r != DM_OK) return r;
          retire(srv, t);
          retired = true;
        }
      }
      if (retired) {
        std::vector<int> waiting;
        for (int t = 0; t < NT; ++t)
          if (!srv->traces[t].sess && next[t] < nframes) waiting.push_back(t);
        r = admit(srv, waiting);
        if (r != DM_OK) return r;
      }
    }
    r = drain(true);
    if (r != DM_OK) return r;
    DM_HIP(hipStreamSynchronize(srv->stream));
    wall = secs_since(t0);
    hipEventDestroy(ev_start);
    if (d_log) {
      DM_HIP(hipMemcpy(poses, d_log, sizeof(double) * 16 * (size_t)NT * nframes, hipMemcpyDeviceToHost));
      hipFree(d_log);
    }
  } else {
    // per-frame path: one frame at a time in arrival order
    const bool host_solve = mode == DM_RUN_PER_FRAME_HOST;
    std::vector<std::pair<double, std::pair<int, int>>> q;
    for (int t = 0; t < NT; ++t)
      for (int f = 0; f < nframes; ++f) q.push_back({arrival(t, f), {t, f}});
    std::stable_sort(q.begin(), q.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });
    for (auto &e : q) {
      const int t = e.second.first, f = e.second.second;
      Trace &tr = srv->traces[t];
      if (!tr.sess) {
        std::vector<int> w = {t};
        // free finished sessions until this one fits
        r = admit(srv, w);
        if (r != DM_OK) return r;
        if (!tr.sess) {
          for (int o = 0; o < NT && !tr.sess; ++o)
            if (o != t && srv->traces[o].sess && next[o] >= nframes) { retire(srv, o); r = admit(srv, w); if (r != DM_OK) return r; }
          if (!tr.sess) { set_error("a trace does not fit the HBM budget"); return DM_E_CAPACITY; }
        }
      }
      const double wait = e.first - secs_since(t0);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      Sess &S = *tr.sess;
      const float *d;
      const uint8_t *c;
      frame_ptrs(t, f, d, c);
      double pose[16];
      S.icp_mfma = srv->icp_mfma;  // the per-frame path runs the server's ICP variant too
      r = dm_session_process_frame_dev(reinterpret_cast<dm_session *>(&S), d, c, host_solve ? 1 : 0, pose);
      if (r != DM_OK) return r;
      const double done = secs_since(t0);
      if (completions) completions[(size_t)t * nframes + f] = done;
      if (poses) memcpy(poses + 16 * ((size_t)t * nframes + f), pose, sizeof pose);
      next[t]++;
      ++nbatch;
      if (next[t] >= nframes && first_frame + nframes >= tr.nframes) retire(srv, t);
    }
  }
  for (int t = 0; t < NT; ++t)
    if (srv->traces[t].sess) srv->traces[t].next_frame = first_frame + next[t];
  if (mode != DM_RUN_BATCHED) wall = secs_since(t0);
  if (wall_seconds) *wall_seconds = wall;
  if (batches) *batches = nbatch;
  return DM_OK;
}

// Time each stage of the batched pipeline with HIP events on the server's
// stream, over the admitted sessions at trace frame `frame` (the sessions'
// state is advanced by the timed launches).  stage_ms[k] = mean duration of
// one launch of stage k: 0 source maps, 1 ICP reduction (level 0),
// 2 ICP solve, 3 allocation (all six kernels), 4 integration, 5 raycast,
// 6 ray traversal kernel alone.  work (summed over sessions, per launch):
// [0] pixels, [1] visible blocks, [2] block visits of the traversal, [3] new
// blocks of the frame, [4] sessions, [5] allocated blocks, [6] distinct blocks
// visited by the traversal.
// Put every admitted session that has frame `frame` into the batch table at
// that frame and run one full frame, so stages can be timed on a real state.
static int stage_setup(dm_server *srv, int32_t frame, std::vector<int> &act, int64_t &pixels) {
  act.clear();
  for (int t = 0; t < (int)srv->traces.size(); ++t)
    if (srv->traces[t].sess && frame < srv->traces[t].nframes) act.push_back(t);
  if (act.empty()) { set_error("no admitted session at this frame"); return DM_E_STATE; }
  order_by_cost(srv, act);
  for (int t : act) srv->traces[t].next_frame = -1;  // the timed launches leave no continuable state
  DM_HIP(hipStreamSynchronize(srv->stream));
  DmView *tab = srv->h_table[0];
  int *lst = srv->h_list[0];
  for (int s = 0; s < srv->cfg.max_sessions; ++s) lst[s] = -1;
  pixels = 0;
  for (size_t q = 0; q < act.size(); ++q) {
    Trace &tr = srv->traces[act[q]];
    const size_t P = (size_t)tr.sensor.rows * tr.sensor.cols;
    tr.sess->h.depth = tr.d_depth + P * (size_t)frame;
    tr.sess->h.rgb = tr.d_rgb ? tr.d_rgb + P * 3 * (size_t)frame : nullptr;
    tab[tr.slot] = tr.sess->h;
    lst[q] = tr.slot;
    pixels += (int64_t)P;
  }
  DM_HIP(hipMemcpyAsync(srv->d_table, tab, sizeof(DmView) * srv->cfg.max_sessions, hipMemcpyHostToDevice, srv->stream));
  DM_HIP(hipMemcpyAsync(srv->d_list, lst, sizeof(int) * srv->cfg.max_sessions, hipMemcpyHostToDevice, srv->stream));
  Launch L = server_launch(srv);
  launch_pose_begin(L);
  launch_source_maps(L);
  launch_alloc(L);
  launch_integrate(L);
  launch_raycast(L);
  DM_HIP(hipStreamSynchronize(srv->stream));
  return DM_OK;
}

int dm_server_profile(dm_server *srv, int32_t frame, int32_t reps, double *stage_ms, int64_t *work) {
  if (!srv || reps <= 0 || !stage_ms) { set_error("invalid profile arguments"); return DM_E_INVALID; }
  std::vector<int> act;
  int64_t pixels = 0;
  DM_TRY(stage_setup(srv, frame, act, pixels));
  Launch L = server_launch(srv);
  int64_t new_blocks = 0;  // blocks inserted by the frame's allocation (stage_setup)
  for (int t : act) {
    int32_t c[C_COUNT];
    DM_HIP(hipMemcpy(c, srv->traces[t].sess->h.ctr, sizeof c, hipMemcpyDeviceToHost));
    new_blocks += c[C_NNEW];
  }
  hipEvent_t e0, e1;
  DM_HIP(hipEventCreate(&e0));
  DM_HIP(hipEventCreate(&e1));
  auto reset_ctr = [&]() -> int {
    for (int t : act) {
      Sess &S = *srv->traces[t].sess;
      DM_HIP(hipMemsetAsync(S.h.ctr + C_NVIS, 0, 3 * sizeof(int32_t), srv->stream));
    }
    return DM_OK;
  };
  // the ICP stages time one full level-0 iteration: the per-level stop flag
  // (which the step sets once the level converges and which makes further
  // launches of the level return at once) is cleared before each repetition
  auto clear_stop = [&]() -> int {
    for (int t : act) {
      Sess &S = *srv->traces[t].sess;
      DM_HIP(hipMemsetAsync(S.h.ctr + C_ICP_STOP, 0, sizeof(int32_t), srv->stream));
    }
    return DM_OK;
  };
  for (int st = 0; st < 7; ++st) {
    int r = reset_ctr();
    if (r != DM_OK) return r;
    double total_ms = 0.0;
    const bool icp = st == 1 || st == 2;
    for (int k = 0; k < reps; ++k) {
      if (icp || k == 0) {
        if (icp && (r = clear_stop()) != DM_OK) return r;
        DM_HIP(hipEventRecord(e0, srv->stream));
      }
      switch (st) {
        case 0: launch_source_maps(L); break;
        case 1: launch_icp_reduce(L, 0, 1, true); break;
        case 2: launch_icp_reduce(L, 0, 1, false); break;
        case 3: launch_alloc(L); break;
        case 4: launch_integrate(L); break;
        case 5: launch_raycast(L); break;
        case 6: launch_alloc_rays(L); break;
      }
      if (icp || k == reps - 1) {
        DM_HIP(hipEventRecord(e1, srv->stream));
        DM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        DM_HIP(hipEventElapsedTime(&ms, e0, e1));
        total_ms += ms;
      }
    }
    stage_ms[st] = total_ms / reps;
    if (work && (st == 4 || st == 6)) {
      int64_t sum_vis = 0, sum_visit = 0, sum_alloc = 0;
      for (int t : act) {
        int32_t c[C_COUNT];
        DM_HIP(hipMemcpy(c, srv->traces[t].sess->h.ctr, sizeof c, hipMemcpyDeviceToHost));
        sum_vis += c[C_NVIS]; sum_visit += c[C_NVISIT]; sum_alloc += c[C_NUM_ALLOC] - c[C_NUM_FREE];
      }
      if (st == 4) { work[1] = sum_vis / reps; work[5] = sum_alloc; }
      if (st == 6) { work[2] = sum_visit / reps; }
    }
  }
  if (work) {
    work[0] = pixels; work[3] = new_blocks; work[4] = (int64_t)act.size();
    // distinct blocks of the traversal: one counting pass marks the pool
    // index of every block its rays find (the frame's blocks are all
    // allocated by now; keys dropped for lack of room are not counted)
    for (int t : act) {
      Sess &S = *srv->traces[t].sess;
      DM_HIP(hipMemsetAsync(S.h.visited, 0, ((size_t)S.h.num_blocks + 31) / 32 * 4, srv->stream));
    }
    launch_alloc_rays(L, true);
    DM_HIP(hipStreamSynchronize(srv->stream));
    int64_t uniq = 0;
    for (int t : act) {
      Sess &S = *srv->traces[t].sess;
      std::vector<uint32_t> bits(((size_t)S.h.num_blocks + 31) / 32);
      DM_HIP(hipMemcpy(bits.data(), S.h.visited, bits.size() * 4, hipMemcpyDeviceToHost));
      for (uint32_t w : bits) uniq += __builtin_popcount(w);
    }
    work[6] = uniq;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  DM_HIP(hipGetLastError());
  return DM_OK;
}

// Launch-shape autotuner: times the raycast with tiles of 8x32, 16x16, 32x8
// and 64x4 pixels and the integration with 1024, 2048 and 4096 workgroups per
// session on the state of `frame`, keeps the fastest of each, and captures
// the serving graph again with them.  The sessions' volumes are advanced by
// the timing runs: call dm_server_reset before serving.
int dm_server_autotune(dm_server *srv, int32_t frame, int32_t reps, int32_t choice[3], double *ms) {
  if (!srv || reps <= 0) { set_error("invalid autotune arguments"); return DM_E_INVALID; }
  std::vector<int> act;
  int64_t pixels = 0;
  DM_TRY(stage_setup(srv, frame, act, pixels));
  hipEvent_t e0, e1;
  DM_HIP(hipEventCreate(&e0));
  DM_HIP(hipEventCreate(&e1));
  auto timed = [&](const Launch &L, bool ray, float &out) -> int {
    ray ? launch_raycast(L) : launch_integrate(L);  // warm-up
    DM_HIP(hipEventRecord(e0, srv->stream));
    for (int k = 0; k < reps; ++k) ray ? launch_raycast(L) : launch_integrate(L);
    DM_HIP(hipEventRecord(e1, srv->stream));
    DM_HIP(hipEventSynchronize(e1));
    DM_HIP(hipEventElapsedTime(&out, e0, e1));
    out /= (float)reps;
    return DM_OK;
  };
  const int tws[4] = {3, 4, 5, 6};
  const int wgs[3] = {1024, 2048, 4096};
  int best_tw = srv->ray_tw_log2, best_wg = srv->integ_wg;
  float best = 3.0e38f;
  for (int k = 0; k < 4; ++k) {
    Launch L = server_launch(srv);
    L.ray_tw_log2 = tws[k];
    float t = 0.f;
    int rc = timed(L, true, t);
    if (rc != DM_OK) { hipEventDestroy(e0); hipEventDestroy(e1); return rc; }
    if (ms) ms[k] = t;
    if (t < best) { best = t; best_tw = tws[k]; }
  }
  best = 3.0e38f;
  for (int k = 0; k < 3; ++k) {
    Launch L = server_launch(srv);
    L.integ_wg = wgs[k];
    float t = 0.f;
    int rc = timed(L, false, t);
    if (rc != DM_OK) { hipEventDestroy(e0); hipEventDestroy(e1); return rc; }
    if (ms) ms[4 + k] = t;
    if (t < best) { best = t; best_wg = wgs[k]; }
  }
  // the ICP linearisation at each level, shuffle tree and matrix cores
  // (reduction only; the per-level stop flag is cleared before each launch
  // so that converged sessions do not return at once).  Reported, not
  // switched: the two sum in different orders, so the choice would change
  // the poses (dm_server_set_shapes selects it).
  for (int var = 0; var < 2; ++var) {
    for (int level = 0; level < kLevels; ++level) {
      Launch L = server_launch(srv);
      L.icp_mfma = var;
      float t = 0.f;
      for (int k = 0; k <= reps; ++k) {  // k = 0: warm-up
        for (int a : act) {
          Sess &S = *srv->traces[a].sess;
          DM_HIP(hipMemsetAsync(S.h.ctr + C_ICP_STOP, 0, sizeof(int32_t), srv->stream));
        }
        DM_HIP(hipEventRecord(e0, srv->stream));
        launch_icp_reduce(L, level, 1, false);
        DM_HIP(hipEventRecord(e1, srv->stream));
        DM_HIP(hipEventSynchronize(e1));
        float one = 0.f;
        DM_HIP(hipEventElapsedTime(&one, e0, e1));
        if (k > 0) t += one;
      }
      t /= (float)reps;
      if (ms) ms[7 + 3 * var + level] = t;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  srv->ray_tw_log2 = best_tw;
  srv->integ_wg = best_wg;
  if (srv->graph) { hipGraphExecDestroy(srv->graph); srv->graph = nullptr; }
  if (choice) { choice[0] = 1 << best_tw; choice[1] = best_wg; choice[2] = srv->icp_mfma; }
  DM_HIP(hipGetLastError());
  return DM_OK;
}

int dm_server_set_shapes(dm_server *srv, const int32_t shapes[3]) {
  if (!srv || !shapes) { set_error("invalid arguments"); return DM_E_INVALID; }
  const int tw = shapes[0], wg = shapes[1], icp = shapes[2];
  int tw_log2 = -1;
  for (int l = 3; l <= 6; ++l)
    if (tw == (1 << l)) tw_log2 = l;
  if (tw_log2 < 0 || wg < 1 || wg > 65535 || (icp != 0 && icp != 1)) {
    set_error("shapes: tile width 8, 16, 32 or 64; integration workgroups 1..65535; ICP variant 0 or 1");
    return DM_E_INVALID;
  }
  srv->ray_tw_log2 = tw_log2;
  srv->integ_wg = wg;
  srv->icp_mfma = icp;
  if (srv->graph) { hipGraphExecDestroy(srv->graph); srv->graph = nullptr; }
  return DM_OK;
}

}  // extern "C"
