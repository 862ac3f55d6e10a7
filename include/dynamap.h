/*
 * dynamap.h -- C ABI of libdynamap_hip.so, the MI355X (gfx950) implementation
 * of DynaMap's dense-mapping path.
 *
 * The reference exposes this path as C++ classes (no FFI).  Each entry point
 * below replaces one reference method; the citation gives the reference file
 * and line (paths relative to the reference repository root).  The C++
 * classes in include/dynamap/dynamap.hpp keep the reference class names and
 * method shapes on top of these functions.
 *
 * Conventions
 *   - Poses are row-major 4x4 camera-to-world matrices (the reference's
 *     float4x4, include/math/utils_matrix.h:829-834, camera-to-world as used
 *     by tsdf.cu:356).  float[16] where the reference takes float4x4,
 *     double[16] where it takes Eigen double/float matrices.
 *   - Depth is float32 metres, colour is interleaved uint8 r,g,b (uchar3).
 *   - Pointers named host_* are host memory; the library stages them.
 *     Functions with the _dev suffix take device pointers.
 *   - Every function returns DM_OK (0) or a negative DM_E_* code;
 *     dm_last_error() describes the last failure of the calling thread.
 *   - Objects are not thread safe; distinct objects may be used from
 *     distinct threads (each owns its HIP stream).
 */
#ifndef DYNAMAP_H
#define DYNAMAP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DM_API __attribute__((visibility("default")))

enum {
  DM_OK = 0,
  DM_E_INVALID = -1,  /* bad argument */
  DM_E_HIP = -2,      /* HIP runtime error (no device, out of memory, ...) */
  DM_E_CAPACITY = -3, /* a fixed capacity (sessions, HBM budget, triangles) would be exceeded */
  DM_E_STATE = -4     /* call not valid in the object's current state */
};

/* rgbdSensor, include/image/sensor/rgbdSensor.h:13-19 */
typedef struct dm_sensor {
  uint32_t rows, cols;
  float fx, fy, cx, cy;
  float depth_scaler;
} dm_sensor;

/* tsdfProperties, include/geometry/tsdf/tsdf.h:19-35 (entriesNum and
 * voxelNumPerAxes are derived; blockSize must be 8; num_blocks < 2^23, so
 * that a voxel index fits 32 bits with one all-zero block past the pool) */
typedef struct dm_tsdf_props {
  float voxel_size;
  int32_t num_buckets, bucket_size, num_blocks, block_size;
  int32_t max_sdf_weight;
  float truncation_distance, max_sensor_depth, min_sensor_depth;
} dm_tsdf_props;

/* solver::Properties, include/solver/gnSolver/gnSolver.h:47-55.  Levels are
 * 0 = full resolution, 1 = half, 2 = quarter; they run 2 -> 1 -> 0.
 * min/max_depth gate the correspondences (the reference hard-codes
 * DEPTH_MIN/DEPTH_MAX, include/image/pyramid/pyramid.h:15-16). */
typedef struct dm_icp_props {
  int32_t lvl_iterations[3];
  float min_depth, max_depth;
  float dist_max;        /* correspondence distance gate, metres */
  float photo_weight;    /* weight of the photometric term (0 = geometric only) */
  float regularization;  /* Levenberg damping factor (reference 0.02) */
  float max_increment;   /* reset guard on the summed twist (reference 0.3) */
  float min_increment;   /* a level ends once its step changes the summed twist by at
                            most this (reference 0.001, gnSolver.cu:414; 0 = every
                            iteration runs) */
} dm_icp_props;

DM_API const char *dm_last_error(void);
DM_API int dm_device_count(int32_t *n);

/* ------------------------------------------------------------------ */
/* Volume: tsdf::tsdfVolume / tsdf::HashTable                          */
/* ------------------------------------------------------------------ */
typedef struct dm_volume dm_volume;

/* tsdfVolume::init -> HashTable::Init, tsdf.cu:18-22, hashTable.cu:41-70.
 * max_rows/max_cols size the frame staging buffers. */
DM_API int dm_volume_create(const dm_tsdf_props *props, uint32_t max_rows, uint32_t max_cols, dm_volume **out);
/* tsdfVolume::Free / HashTable::Free, tsdf.cu:23-25, hashTable.cu:72-78 */
DM_API int dm_volume_destroy(dm_volume *vol);
/* tsdfVolume::IntegrateScan, tsdf.cu:454-469 (block allocation then
 * integration).  host_rgb may be NULL (colour 0). */
DM_API int dm_volume_integrate(dm_volume *vol, const float *host_depth, const uint8_t *host_rgb,
                               const dm_sensor *sensor, const float pose[16]);
/* AllocateFromDepthKernel only, tsdf.cu:345-453 */
DM_API int dm_volume_allocate(dm_volume *vol, const float *host_depth, const dm_sensor *sensor,
                              const float pose[16]);
/* tsdfVolume::GenerateDepth / GenerateRgb, tsdf.cu:527-550, plus the model
 * vertex and normal maps used by tracking.  Any output may be NULL.
 * vertex/normal: rows*cols*3 floats in camera coordinates; normal is NaN
 * where undefined; depth 0 = no surface. */
DM_API int dm_volume_raycast(dm_volume *vol, const float pose[16], const dm_sensor *sensor, float *host_depth,
                             float *host_vertex, float *host_normal, uint8_t *host_rgb);
/* HashTable::GetNumAllocatedBlocks, hashTable.cu:165-167 (live blocks) */
DM_API int dm_volume_num_allocated(dm_volume *vol, int32_t *n);
/* HashTable::GetNumEntries, hashTable.cu:169-171 */
DM_API int dm_volume_num_entries(dm_volume *vol, int64_t *n);
/* HashTable::GetHashEntry for every entry, hashTable.cu:173-175:
 * out[4*i..4*i+3] = (x, y, z, pointer), pointer -1 = free (kFreeEntry,
 * hashTable.h:9).  out holds num_entries*4 int32. */
DM_API int dm_volume_get_entries(dm_volume *vol, int32_t *out);
/* All allocated blocks in pool order: block coordinates (3 int32), and
 * optionally sdf (512 float), rgb (512*3 uint8), weight (512 uint8) per
 * block, voxel index x*64 + y*8 + z (geom_Types.h:116-119). */
DM_API int dm_volume_get_blocks(dm_volume *vol, int32_t cap, int32_t *xyz, float *sdf, uint8_t *rgb,
                                uint8_t *weight, int32_t *n_out);
/* HashTable::DeleteBlock (hashTable.cu:112-134) for n blocks (xyz: n x 3
 * block coordinates; absent or repeated ones are skipped): each block's
 * voxels are reset, its pool index goes back on the free stack (Heap::Append,
 * heap.cu:19-22) that later insertions take first, and its bucket's remaining
 * entries close up in their order (occupied slots stay a prefix).  deleted
 * (may be NULL) receives the number removed. */
DM_API int dm_volume_delete_blocks(dm_volume *vol, const int32_t *xyz, int64_t n, int64_t *deleted);
/* Pool recycling for long-running reconstructions (no reference
 * counterpart; the reference never calls DeleteBlock): delete every
 * allocated block whose centre ((x + 1/2) * 8 * voxel_size, per axis) lies
 * farther than radius metres from center. */
DM_API int dm_volume_recycle(dm_volume *vol, const double center[3], double radius, int64_t *deleted);
/* Capacity counters (the reference drops silently, hashTable.cu:109):
 * out[0] allocated blocks, [1] dropped (bucket full), [2] dropped (pool
 * exhausted), [3] block visits outside the key range, [4] new keys lost to a
 * full per-frame key set. */
DM_API int dm_volume_counters(dm_volume *vol, int64_t out[5]);
/* Snapshot / restore of the reconstruction state (no reference counterpart;
 * resumable sessions): the hash table, the allocated blocks and their flags,
 * and for a session its pose state and model maps, in one host buffer.
 * dm_volume_snapshot with buf NULL reports the size.  dm_volume_restore
 * needs a volume (or dm_session_volume of a session) created with the same
 * properties and frame size; processing then continues exactly as it would
 * have in the snapshotted one. */
DM_API int dm_volume_snapshot(dm_volume *vol, void *buf, uint64_t cap, uint64_t *size);
DM_API int dm_volume_restore(dm_volume *vol, const void *buf, uint64_t size);

/* ------------------------------------------------------------------ */
/* Depth preprocessing: pyramid::bilateralFilter / downPyramidDepth    */
/* ------------------------------------------------------------------ */
/* pyramid::bilateralFilter (pyramid.cu:200-282): weight
 * exp(-dd^2/(2 sigma_depth^2)) * exp(-|q-p|^2/(2 sigma_space^2)) over the
 * (2 radius + 1)^2 window, pixels without depth skipped and kept at 0.
 * radius 0..7; host buffers of rows*cols metres. */
DM_API int dm_depth_bilateral(const float *host_in, float *host_out, uint32_t rows, uint32_t cols, int32_t radius,
                              float sigma_depth, float sigma_space);
/* pyramid::downPyramidDepth (pyramid.cu:119-143), as a Gaussian pyramid:
 * level l+1 pixel (u, v) = binomial 5x5 average around (2u, 2v) of the
 * pixels within 3 sigma_depth of the centre's depth.  host_out receives
 * levels 0..levels-1 back to back ((rows >> l) x (cols >> l) each). */
DM_API int dm_depth_pyramid(const float *host_in, uint32_t rows, uint32_t cols, float sigma_depth, int32_t levels,
                            float *host_out);

/* ------------------------------------------------------------------ */
/* Mesh: geometry::MeshExtractor / geometry::Mesh                      */
/* ------------------------------------------------------------------ */
/* The marching-cubes triangle table (meshExtractor/lookupTables.h triTable):
 * one word per cube index, edge k of the triangle list in bits 4k..4k+3,
 * 0xF ends the list. */
DM_API int dm_mc_tri_table(uint64_t out[256]);
/* MeshExtractor::ExtractMesh, meshExtractor.cu:364-402, tsdf.cu:551-560.
 * Every voxel of every allocated block is a cell (cell centre = voxel
 * centre, corners at +-voxel/2); with box_lo/box_hi (both or neither) only
 * cells whose centre lies in the box.  Triangles are ordered by hash entry
 * (bucket, slot), then cell (x-major within the block), then table order.
 * *ntri receives the total; the first min(total, cap) triangles are written:
 * pos and col 9 floats per triangle (3 vertices; col = rgb / 255 of the cell's
 * centre voxel, meshExtractor.cu:193), edge_ids 3 per triangle (the grid edge
 * of each vertex: axis << 60 | (x + 2^19) << 40 | (y + 2^19) << 20 | (z + 2^19)
 * of its lower end, in corner units where corner c lies at (c - 1/2) * voxel).
 * Equal ids are the same vertex, so welding is a merge by id.  Any output
 * may be NULL; cap 0 counts only. */
DM_API int dm_mesh_extract(dm_volume *vol, const double box_lo[3], const double box_hi[3], int64_t cap, float *pos,
                           float *col, uint64_t *edge_ids, int64_t *ntri);
/* dm_mesh_extract without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * triangles, order and values as dm_mesh_extract; the outputs are DEVICE
 * pointers (pos, col 9 floats and edge_ids 3 per triangle, any may be NULL)
 * on the volume's device.  The kernels' scratch is one fixed array per
 * device (about 29 MiB, held until the process ends); the volume may have at
 * most 2^20 hash buckets and 2^21 allocated blocks (DM_E_CAPACITY
 * otherwise).  Calls that overlap on one device share that array and race:
 * serialise them.  Calls no hipMalloc or hipFree; it synchronises the
 * volume's stream. */
DM_API int dm_mesh_extract_noalloc(dm_volume *vol, const double box_lo[3], const double box_hi[3], int64_t cap,
                                   float *dev_pos, float *dev_col, uint64_t *dev_edge_ids, int64_t *ntri);

#ifdef __cplusplus
}
#endif
#endif /* DYNAMAP_H */
