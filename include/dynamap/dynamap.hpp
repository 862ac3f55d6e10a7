// dynamap.hpp -- C++ classes with the reference's names and method shapes,
// forwarding to the C ABI of libdynamap_hip.so (include/dynamap.h,
// include/dynamap_session.h).  Header only; link with -ldynamap_hip.
//
// Reference classes mirrored (paths relative to the reference root):
//   DynaMap::tsdf::tsdfVolume      include/geometry/tsdf/tsdf.h:37-81
//   DynaMap::tsdf::HashTable       include/dataStructure/hashTable/hashTable.h:32-141
//   DynaMap::solver::gnSolver      include/solver/gnSolver/gnSolver.h:57-100
//   DynaMap::builder               include/builder/builder.h:49-79
//   DynaMap::geometry::MeshExtractor / Mesh
//                                  include/geometry/mesh/meshExtractor/meshExtractor.h:20-59,
//                                  include/geometry/mesh/mesh.h:25-66
// Differences: device memory is owned by the objects (the reference uses
// managed memory the caller frees); outputs are returned in std::vector;
// Eigen matrices are replaced by row-major double[16] poses; failures throw
// DynaMap::Error carrying the DM_E_* code.
#pragma once
#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dynamap.h"
#include "../dynamap_session.h"

namespace DynaMap {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int code) {
  if (code != DM_OK) throw Error(code, std::string("dynamap: ") + dm_last_error());
}

using rgbdSensor = dm_sensor;  // include/image/sensor/rgbdSensor.h:13-19

// row-major 4x4 (include/math/utils_matrix.h:829-834)
struct float4x4 {
  float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float4x4() = default;
  explicit float4x4(const float v[16]) {
    for (int k = 0; k < 16; ++k) m[k] = v[k];
  }
  explicit float4x4(const double v[16]) {
    for (int k = 0; k < 16; ++k) m[k] = (float)v[k];
  }
};

struct float3 {
  float x, y, z;
};

// rgbdImage (include/image/sensor/rgbdImage.h) over host buffers
struct rgbdImage {
  const float *depth = nullptr;   // rows*cols metres
  const uint8_t *rgb = nullptr;   // rows*cols*3, may be null
  rgbdSensor sensor{};
};

namespace tsdf {

using tsdfProperties = dm_tsdf_props;  // tsdf.h:19-35

// HashEntry (hashTable.h:20-25): block position and pool pointer, -1 = free
struct HashEntry {
  int x, y, z;
  int pointer;
};

class tsdfVolume {
 public:
  tsdfVolume() = default;
  tsdfVolume(const tsdfVolume &) = delete;
  tsdfVolume &operator=(const tsdfVolume &) = delete;
  ~tsdfVolume() { Free(); }

  // tsdfVolume::init -> HashTable::Init (tsdf.cu:18-22)
  void init(const tsdfProperties &p, uint32_t max_rows = 480, uint32_t max_cols = 640) {
    Free();
    Properties = p;
    check(dm_volume_create(&p, max_rows, max_cols, &v_));
    owned_ = true;
  }
  void Free() {
    if (v_ && owned_) dm_volume_destroy(v_);
    v_ = nullptr;
  }
  // tsdf.cu:454-469
  void IntegrateScan(const rgbdImage &image, const float4x4 &camera_pose) {
    check(dm_volume_integrate(v_, image.depth, image.rgb, &image.sensor, camera_pose.m));
  }
  // tsdf.cu:527-538 (returns an owning buffer instead of managed memory)
  std::vector<float> GenerateDepth(const float4x4 &camera_pose, const rgbdSensor &sensor) {
    std::vector<float> d((size_t)sensor.rows * sensor.cols);
    check(dm_volume_raycast(v_, camera_pose.m, &sensor, d.data(), nullptr, nullptr, nullptr));
    return d;
  }
  // tsdf.cu:539-550
  std::vector<uint8_t> GenerateRgb(const float4x4 &camera_pose, const rgbdSensor &sensor) {
    std::vector<uint8_t> c((size_t)sensor.rows * sensor.cols * 3);
    check(dm_volume_raycast(v_, camera_pose.m, &sensor, nullptr, nullptr, nullptr, c.data()));
    return c;
  }
  // HashTable::GetNumAllocatedBlocks (hashTable.cu:165-167)
  int GetNumAllocatedBlocks() {
    int32_t n = 0;
    check(dm_volume_num_allocated(v_, &n));
    return n;
  }
  // HashTable::GetNumEntries (hashTable.cu:169-171)
  int GetNumEntries() {
    int64_t n = 0;
    check(dm_volume_num_entries(v_, &n));
    return (int)n;
  }
  // HashTable::GetHashEntry (hashTable.cu:173-175); reads the whole table
  HashEntry GetHashEntry(int i) {
    std::vector<int32_t> e((size_t)GetNumEntries() * 4);
    check(dm_volume_get_entries(v_, e.data()));
    return HashEntry{e[4 * i], e[4 * i + 1], e[4 * i + 2], e[4 * i + 3]};
  }
  tsdfProperties GetProperties() const { return Properties; }
  dm_volume *handle() { return v_; }

  tsdfProperties Properties{};

 private:
  dm_volume *v_ = nullptr;
  bool owned_ = false;
};

}  // namespace tsdf

namespace solver {

// solver::Properties (gnSolver.h:47-55); levels 0 = full resolution
struct Properties {
  int lvlIterationNum[3] = {10, 5, 4};
  float minDepth = 0.3f, maxDepth = 5.0f;
  float distMax = 0.1f;
  float photoWeight = 0.0f;
  float regularization = 0.02f;
  float maxIncrement = 0.3f;
  dm_icp_props c() const {
    dm_icp_props p{};
    for (int l = 0; l < 3; ++l) p.lvl_iterations[l] = lvlIterationNum[l];
    p.min_depth = minDepth; p.max_depth = maxDepth; p.dist_max = distMax;
    p.photo_weight = photoWeight; p.regularization = regularization; p.max_increment = maxIncrement;
    return p;
  }
};

}  // namespace solver

// builder (builder.h:49-79): volume + gnSolver per frame.
class builder {
 public:
  builder() = default;
  builder(const builder &) = delete;
  builder &operator=(const builder &) = delete;
  ~builder() {
    if (s_) dm_session_destroy(s_);
  }
  // builder::init (builder.cu:11-49)
  void init(const tsdf::tsdfProperties &tp, const solver::Properties &op, const rgbdSensor &sensor,
            const double init_pose[16] = nullptr) {
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    dm_icp_props ip = op.c();
    check(dm_session_create(&tp, &ip, &sensor, init_pose ? init_pose : I, &s_));
  }
  // builder::buildModelLS (builder.cu:70-139): track, integrate, raycast
  std::array<double, 16> buildModelLS(const rgbdImage &frame, bool solve_on_host = false) {
    std::array<double, 16> pose{};
    check(dm_session_process_frame(s_, frame.depth, frame.rgb, solve_on_host ? 1 : 0, pose.data()));
    return pose;
  }
  std::array<double, 16> currentPose() {
    std::array<double, 16> pose{};
    check(dm_session_get_pose(s_, pose.data()));
    return pose;
  }
  dm_volume *volume() { return dm_session_volume(s_); }
  dm_session *handle() { return s_; }

 private:
  dm_session *s_ = nullptr;
};

}  // namespace DynaMap
