/*
 * dynamap_session.h -- per-frame pipeline, tracking and serving entry points
 * of libdynamap_hip.so.  See dynamap.h for the conventions.
 */
#ifndef DYNAMAP_SESSION_H
#define DYNAMAP_SESSION_H
#include "dynamap.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Session: builder + solver::gnSolver                                 */
/* ------------------------------------------------------------------ */
typedef struct dm_session dm_session;

/* builder::init, include/builder/builder.cu:11-49: a volume, the live-frame
 * and model maps for one sensor, and the pose of the first frame. */
DM_API int dm_session_create(const dm_tsdf_props *props, const dm_icp_props *icp, const dm_sensor *sensor,
                             const double init_pose[16], dm_session **out);
DM_API int dm_session_destroy(dm_session *s);
/* the session's tsdfVolume (builder's volume_, builder.h:81) */
DM_API dm_volume *dm_session_volume(dm_session *s);
/* builder::buildModelLS, builder.cu:70-139: track the frame against the
 * model (gnSolver::solve, skipped for the first frame), integrate it at the
 * tracked pose (IntegrateScan) and raycast the model for the next frame
 * (GenerateDepth/GenerateRgb).  solve_on_host = 1 solves each 6x6 step on
 * the host after reading the reduction back (the reference's orchestration,
 * gnSolver.cu:379-415); 0 keeps the whole frame on the device.  pose_out
 * (may be NULL) receives the frame's camera-to-world pose. */
DM_API int dm_session_process_frame(dm_session *s, const float *host_depth, const uint8_t *host_rgb,
                                    int32_t solve_on_host, double pose_out[16]);
DM_API int dm_session_process_frame_dev(dm_session *s, const float *dev_depth, const uint8_t *dev_rgb,
                                        int32_t solve_on_host, double pose_out[16]);
DM_API int dm_session_get_pose(dm_session *s, double pose[16]);
/* Bilateral filter of the frame before tracking (pyramid::bilateralFilter,
 * pyramid.cu:259-282, as dm_depth_bilateral): the tracking maps are built
 * from the filtered depth, integration uses the raw depth.  radius 0 turns
 * it off (the default). */
DM_API int dm_session_set_depth_filter(dm_session *s, int32_t radius, float sigma_depth, float sigma_space);

/* rgbdImage::getPointCloudXYZ + getNormalsfromVertices (rgbdImage.cu:39-49,
 * 158-167) on pyramid level `level` (pyramid::downPyramidDepth,
 * pyramid.cu:119-143) of the frame: vertex/normal rows*cols*3 floats,
 * intensity rows*cols floats, level l of (rows >> l) x (cols >> l). */
DM_API int dm_session_source_maps(dm_session *s, const float *host_depth, const uint8_t *host_rgb, int32_t level,
                                  float *host_vertex, float *host_normal, float *host_intensity);
/* Load model maps of one level from host arrays (in place of a raycast), for
 * tracking in isolation. */
DM_API int dm_session_set_model_maps(dm_session *s, int32_t level, const float *host_vertex,
                                     const float *host_normal, const float *host_intensity);
/* Select the kernel that linearises ICP for this session (no reference
 * counterpart; gnSolver.cu:25-196 has one): 0 the shuffle-tree reduction
 * (default), 1 the FP32 matrix-core F^T F contraction.  Both give the same
 * system up to summation order. */
DM_API int dm_session_set_icp_reduce(dm_session *s, int32_t variant);
/* buildNLSKernel, gnSolver.cu:25-196: the normal equations of one
 * linearisation at `level` for increment Tinc, using the source maps of the
 * last dm_session_source_maps call and the current model maps.
 * A = sum J J^T (row-major 6x6), b = sum J r, sums = {#geometric residuals,
 * sum r^2, #photometric residuals}. */
DM_API int dm_session_icp_system(dm_session *s, int32_t level, const double Tinc[16], double A[36], double b[6],
                                 double sums[3]);
/* gnSolver::solve, gnSolver.cu:317-438: the multi-level increment (source
 * camera -> model camera) of the staged source frame against the model maps. */
DM_API int dm_session_icp_track(dm_session *s, int32_t solve_on_host, double Tinc_out[16]);
/* gnSolver::solve's `increment` (gnSolver.cu:353-430, returned through
 * prev_increment_): the summed twist (omega, t) of the session's last
 * tracking, dm_session_icp_track or a processed frame, zero where the
 * reference's guard (NaN, infinite, or a sum >= max_increment) reset it. */
DM_API int dm_session_icp_increment(dm_session *s, double twist[6]);

/* ------------------------------------------------------------------ */
/* Serving loop (no reference counterpart: the reference processes one  */
/* frame at a time, src/dynaMap.cpp:218-301)                            */
/* ------------------------------------------------------------------ */
typedef struct dm_server dm_server;

typedef struct dm_server_config {
  int32_t max_sessions;       /* session slots resident at once */
  uint32_t max_rows, max_cols;
  uint64_t hbm_budget_bytes;  /* admission budget for session state */
  int32_t use_graph;          /* capture the batched frame pipeline in a HIP graph */
} dm_server_config;

DM_API int dm_server_create(const dm_server_config *cfg, dm_server **out);
DM_API int dm_server_destroy(dm_server *srv);
/* Register a reconstruction request: parameters, first pose and its frames
 * (nframes x rows x cols depth, optional rgb), copied to device memory. */
DM_API int dm_server_add_trace(dm_server *srv, const dm_tsdf_props *props, const dm_icp_props *icp,
                               const dm_sensor *sensor, const double init_pose[16], int32_t nframes,
                               const float *host_depth, const uint8_t *host_rgb, int32_t *trace_id);
/* Device bytes of session state one trace needs (admission cost). */
DM_API int dm_server_trace_bytes(dm_server *srv, int32_t trace_id, uint64_t *bytes);

enum {
  DM_RUN_BATCHED = 0,        /* continuous batching across sessions */
  DM_RUN_PER_FRAME_HOST = 1, /* one frame at a time, host 6x6 solves */
  DM_RUN_PER_FRAME_DEV = 2   /* one frame at a time, device solves */
};

/* Serve frames [first_frame, first_frame + nframes) of every trace, starting
 * each trace from a fresh session when first_frame == 0.  A run from
 * first_frame > 0 continues the live sessions and needs every trace to have
 * stopped at first_frame in the previous run (DM_E_STATE otherwise, e.g.
 * after a trace's last frame retired its session, or after
 * dm_server_profile / dm_server_autotune advanced the volumes).  arrivals: per trace, per frame arrival time in
 * seconds relative to the call (trace-major, nframes per trace), or NULL
 * when every frame is available at once (closed loop).  Outputs (may be
 * NULL): completion time of every frame (seconds, same layout), the pose of
 * every frame (16 doubles each), the wall time of the call, the number of
 * batches launched. */
DM_API int dm_server_run(dm_server *srv, int32_t mode, int32_t first_frame, int32_t nframes,
                         const double *arrivals, double *completions, double *poses, double *wall_seconds,
                         int64_t *batches);
/* Time each stage of the batched pipeline (HIP events on the server's
 * stream) over the admitted sessions at trace frame `frame`; advances their
 * state.  stage_ms[7]: source maps, ICP reduction (level 0), ICP solve,
 * allocation (all kernels), integration, raycast, ray traversal alone.
 * work[7]: pixels, visible blocks, traversal block visits, blocks inserted
 * by the frame, sessions, allocated blocks, distinct blocks visited by the
 * traversal. */
DM_API int dm_server_profile(dm_server *srv, int32_t frame, int32_t reps, double *stage_ms, int64_t *work);
/* Retire every session (free their state). */
DM_API int dm_server_reset(dm_server *srv);
/* Launch-shape autotuner (no reference counterpart): times the raycast tile
 * shapes (8x32, 16x16, 32x8, 64x4 pixels) and the integration grid (1024,
 * 2048, 4096 workgroups per session) on the admitted sessions' state at
 * `frame`, and keeps the fastest of each for the serving graph (neither
 * changes a result).  It also times the ICP linearisation (shuffle-tree
 * reduction and FP32 matrix-core F^T F, at the three pyramid levels) but
 * keeps the configured variant, which dm_server_set_shapes selects: the two
 * sum in different orders, so the poses differ at the rounding level.
 * choice (may be NULL) = {tile width, integration workgroups, ICP variant
 * (0 shuffle tree, 1 matrix cores)}; ms (may be NULL) = the 13 timings: 4
 * raycast, 3 integration, ICP levels 0-2 shuffle tree, ICP levels 0-2
 * matrix cores.  Advances the sessions' volumes: dm_server_reset before
 * serving. */
DM_API int dm_server_autotune(dm_server *srv, int32_t frame, int32_t reps, int32_t choice[3], double *ms);
/* Set the serving graph's launch shapes directly (no reference counterpart):
 * shapes = {raycast tile width 8/16/32/64, integration workgroups per
 * session 1..65535, ICP variant 0 (shuffle tree) or 1 (matrix cores)}, the
 * values dm_server_autotune chooses.  Poses depend on the ICP variant (the
 * two sum in different orders). */
DM_API int dm_server_set_shapes(dm_server *srv, const int32_t shapes[3]);

#ifdef __cplusplus
}
#endif
#endif /* DYNAMAP_SESSION_H */
