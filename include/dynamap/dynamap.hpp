// dynamap.hpp -- C++ classes with the reference's names and method shapes,
// forwarding to the C ABI of libdynamap_hip.so (include/dynamap.h,
// include/dynamap_session.h).  Header only; link with -ldynamap_hip.
//
// Reference classes mirrored (paths relative to the reference root):
//   DynaMap::tsdf::tsdfVolume      include/geometry/tsdf/tsdf.h:37-81
//   DynaMap::tsdf::HashTable       include/dataStructure/hashTable/hashTable.h:32-141
//   DynaMap::solver::gnSolver      include/solver/gnSolver/gnSolver.h:57-100
//   DynaMap::builder               include/builder/builder.h:49-79
//   DynaMap::geometry::defGraph    include/geometry/defGraph/defGraph.h:45-87
//   DynaMap::blender::dqBlender    include/blender/dqBlender/dqBlender.h:19-83
//   DynaMap::solver::nonRigidICP   include/solver/nonRigidICP/nonRigidICP.h:59-183
//   DynaMap::geometry::MeshExtractor / Mesh (extraction and OBJ output)
//                                  include/geometry/mesh/meshExtractor/meshExtractor.h:20-59,
//                                  include/geometry/mesh/mesh.h:25-66
// Differences: device memory is owned by the objects (the reference uses
// managed memory the caller frees); outputs are returned in std::vector;
// Eigen matrices are replaced by row-major double[16] poses; failures throw
// DynaMap::Error carrying the DM_E_* code.
#pragma once
#include <algorithm>
#include <array>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dynamap.h"
#include "../dynamap_nonrigid.h"
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

namespace geometry {

// Vertex / Triangle (include/geometry/geom_Types.h): position and colour
struct Vertex {
  float3 position;
  float3 color;
};
struct Triangle {
  Vertex v0, v1, v2;
};

// Mesh (mesh.h:25-66) as a host triangle soup.  edge_ids[3 * i + j] names
// the grid edge vertex j of triangle i lies on; equal ids are one vertex.
class Mesh {
 public:
  std::vector<Triangle> triangles_;
  std::vector<uint64_t> edge_ids_;
  unsigned int num_triangles_ = 0;  // found by the extraction
  unsigned int max_triangles_ = 0;  // held at most
  // Mesh::SaveToFile (mesh.cu:38-72): "v x y z r g b" per triangle vertex,
  // then one face "f 3i+1 3i+2 3i+3" per triangle; returns num_triangles_
  int SaveToFile(const std::string &filename) const {
    std::ofstream f(filename);
    if (!f) throw Error(DM_E_INVALID, "cannot open " + filename);
    for (const Triangle &t : triangles_)
      for (const Vertex *v : {&t.v0, &t.v1, &t.v2})
        f << "v " << v->position.x << " " << v->position.y << " " << v->position.z << " " << v->color.x << " "
          << v->color.y << " " << v->color.z << "\n";
    for (size_t i = 1; i <= triangles_.size() * 3; i += 3) f << "f " << i << " " << i + 1 << " " << i + 2 << "\n";
    return (int)num_triangles_;
  }
};

}  // namespace geometry

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
  // whole reconstruction state (no reference counterpart: resumable sessions)
  std::vector<uint8_t> Snapshot() {
    uint64_t n = 0;
    check(dm_volume_snapshot(v_, nullptr, 0, &n));
    std::vector<uint8_t> buf(n);
    check(dm_volume_snapshot(v_, buf.data(), n, &n));
    return buf;
  }
  void Restore(const std::vector<uint8_t> &buf) { check(dm_volume_restore(v_, buf.data(), buf.size())); }
  // tsdfVolume::ExtractMesh (tsdf.cu:551-560): at most 200 000 triangles
  geometry::Mesh ExtractMesh(const float3 &lower_corner, const float3 &upper_corner);
  // every allocated block
  geometry::Mesh ExtractMesh();

  tsdfProperties Properties{};

 private:
  dm_volume *v_ = nullptr;
  bool owned_ = false;
};

}  // namespace tsdf

namespace geometry {

// MeshExtractor (meshExtractor.h:20-59): marching cubes on the GPU over the
// allocated blocks; with a box, over the cells whose centre lies in it.
class MeshExtractor {
 public:
  // MeshExtractor::Init (meshExtractor.cu:11-15); the voxel size is the volume's
  void Init(unsigned int max_triangles, float voxel_size = 0.f) {
    (void)voxel_size;
    max_triangles_ = max_triangles;
  }
  void Free() { mesh_ = Mesh(); }
  // MeshExtractor::ExtractMesh (meshExtractor.cu:391-402)
  void ExtractMesh(tsdf::tsdfVolume *volume, float3 lower_corner, float3 upper_corner) {
    const double lo[3] = {lower_corner.x, lower_corner.y, lower_corner.z};
    const double hi[3] = {upper_corner.x, upper_corner.y, upper_corner.z};
    run(volume, lo, hi);
  }
  void ExtractMesh(tsdf::tsdfVolume *volume) { run(volume, nullptr, nullptr); }
  Mesh GetMesh() const { return mesh_; }

 private:
  void run(tsdf::tsdfVolume *volume, const double *lo, const double *hi) {
    int64_t total = 0;
    check(dm_mesh_extract(volume->handle(), lo, hi, 0, nullptr, nullptr, nullptr, &total));
    const int64_t m = std::min<int64_t>(total, max_triangles_);
    std::vector<float> pos((size_t)m * 9), col((size_t)m * 9);
    mesh_ = Mesh();
    mesh_.edge_ids_.resize((size_t)m * 3);
    check(dm_mesh_extract(volume->handle(), lo, hi, m, pos.data(), col.data(), mesh_.edge_ids_.data(), &total));
    mesh_.triangles_.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
      Vertex *v[3] = {&mesh_.triangles_[i].v0, &mesh_.triangles_[i].v1, &mesh_.triangles_[i].v2};
      for (int j = 0; j < 3; ++j) {
        const float *p = &pos[(size_t)i * 9 + 3 * j], *c = &col[(size_t)i * 9 + 3 * j];
        v[j]->position = float3{p[0], p[1], p[2]};
        v[j]->color = float3{c[0], c[1], c[2]};
      }
    }
    mesh_.num_triangles_ = (unsigned int)total;
    mesh_.max_triangles_ = max_triangles_;
  }
  unsigned int max_triangles_ = 200000;
  Mesh mesh_;
};

}  // namespace geometry

namespace tsdf {
inline geometry::Mesh tsdfVolume::ExtractMesh(const float3 &lower_corner, const float3 &upper_corner) {
  geometry::MeshExtractor x;
  x.Init(200000, Properties.voxel_size);
  x.ExtractMesh(this, lower_corner, upper_corner);
  return x.GetMesh();
}
inline geometry::Mesh tsdfVolume::ExtractMesh() {
  geometry::MeshExtractor x;
  x.Init(200000, Properties.voxel_size);
  x.ExtractMesh(this);
  return x.GetMesh();
}
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
  float minIncrement = 0.001f;  // src/dynaMap.cpp:156
  dm_icp_props c() const {
    dm_icp_props p{};
    for (int l = 0; l < 3; ++l) p.lvl_iterations[l] = lvlIterationNum[l];
    p.min_depth = minDepth; p.max_depth = maxDepth; p.dist_max = distMax;
    p.photo_weight = photoWeight; p.regularization = regularization; p.max_increment = maxIncrement;
    p.min_increment = minIncrement;
    return p;
  }
};

// gnSolver (gnSolver.h:57-100): multi-level projective ICP of a source frame
// against target maps, on the GPU.  The target is given as a depth frame
// (its vertex/normal maps become the model) or left as the last raycast of
// a session.
class gnSolver {
 public:
  gnSolver() = default;
  gnSolver(const gnSolver &) = delete;
  gnSolver &operator=(const gnSolver &) = delete;
  ~gnSolver() {
    if (s_) dm_session_destroy(s_);
  }
  // gnSolver::init (gnSolver.cu:287-315)
  void init(const Properties &prop, const rgbdSensor &sensor, const tsdf::tsdfProperties &tp) {
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    dm_icp_props ip = prop.c();
    check(dm_session_create(&tp, &ip, &sensor, I, &s_));
    rows_ = sensor.rows;
    cols_ = sensor.cols;
  }
  // gnSolver::solve (gnSolver.cu:317-438): increment source camera -> target
  // camera; solve_on_host keeps the reference's host 6x6 steps
  std::array<double, 16> solve(const rgbdImage &target, const rgbdImage &source, bool solve_on_host = false) {
    for (int l = 0; l < 3; ++l) {
      const size_t P = (size_t)(rows_ >> l) * (cols_ >> l);
      std::vector<float> v(P * 3), n(P * 3), in(P);
      check(dm_session_source_maps(s_, target.depth, target.rgb, l, v.data(), n.data(), in.data()));
      check(dm_session_set_model_maps(s_, l, v.data(), n.data(), in.data()));
    }
    std::vector<float> tmp((size_t)rows_ * cols_ * 3);
    check(dm_session_source_maps(s_, source.depth, source.rgb, 0, tmp.data(), nullptr, nullptr));
    std::array<double, 16> T{};
    check(dm_session_icp_track(s_, solve_on_host ? 1 : 0, T.data()));
    return T;
  }

 private:
  dm_session *s_ = nullptr;
  uint32_t rows_ = 0, cols_ = 0;
};

// nonRigidICP properties (the build's formulation; the reference's
// solver::Properties carries the iteration counts and regularisation)
struct NonRigidProperties {
  int gnIterations = 5, pcgIterations = 40;
  double pcgTol = 1e-6, lambdaReg = 10.0, damping = 1e-3, distMax = 0.05, normalCos = 0.5;
  double minDepth = 0.1, maxDepth = 5.0;
  dm_nonrigid_props c() const {
    return dm_nonrigid_props{gnIterations, pcgIterations, pcgTol, lambdaReg, damping, distMax, normalCos, minDepth,
                             maxDepth};
  }
};

}  // namespace solver

namespace geometry {

// MeshSTD (meshSTD.h): vertices with normals
struct MeshSTD {
  std::vector<float3> vertices, normals;
};

// defGraph (defGraph.h:45-87): nodes sampled from the canonical mesh, their
// neighbours and weights, node transforms; on the GPU.
class defGraph {
 public:
  defGraph() = default;
  defGraph(const defGraph &) = delete;
  defGraph &operator=(const defGraph &) = delete;
  ~defGraph() { Free(); }
  // defGraph::init + defGraphUpdateNodes (defGraph.cu:15-69, :483-545)
  void init(const MeshSTD &srcMesh, float nodeRadius = 0.025f, float dgw = 0.05f, int maxNodes = 4096) {
    Free();
    dm_graph_props p{nodeRadius, dgw, maxNodes};
    check(dm_graph_create(&srcMesh.vertices[0].x, &srcMesh.normals[0].x, (int64_t)srcMesh.vertices.size(), &p, &g_));
    int32_t n = 0;
    check(dm_graph_num_nodes(g_, &n));
    nodeNum = n;
    verticesNum = srcMesh.vertices.size();
  }
  void defGraphUpdateNodes() {}  // neighbours and weights are built by init
  void Free() {
    if (g_) dm_graph_destroy(g_);
    g_ = nullptr;
  }
  // defGraph::updateNodesDQ (defGraph.cu:704-730): node transforms, nodeNum x 16
  void updateNodesDQ(const std::vector<double> &T) { check(dm_graph_set_transforms(g_, T.data())); }
  std::vector<double> nodeTransforms() {
    std::vector<double> T((size_t)nodeNum * 16);
    check(dm_graph_get_transforms(g_, T.data()));
    return T;
  }
  dm_graph *handle() { return g_; }
  int nodeNum = 0;
  size_t verticesNum = 0;

 private:
  dm_graph *g_ = nullptr;
};

}  // namespace geometry

namespace blender {

// dqBlender (dqBlender.h:19-83)
class dqBlender {
 public:
  void init(geometry::defGraph &graph) { graph_ = &graph; }
  // dqBlender::blendMesh (dqBlender.cu:187-218)
  void blendMesh(geometry::MeshSTD &dst, geometry::defGraph &graph) {
    dst.vertices.resize(graph.verticesNum);
    dst.normals.resize(graph.verticesNum);
    check(dm_blend(graph.handle(), &dst.vertices[0].x, &dst.normals[0].x));
  }
  void blendMesh(geometry::MeshSTD &dst) { blendMesh(dst, *graph_); }

 private:
  geometry::defGraph *graph_ = nullptr;
};

}  // namespace blender

namespace solver {

// nonRigidICP (nonRigidICP.h:59-183)
class nonRigidICP {
 public:
  void init(geometry::defGraph &, const NonRigidProperties &prop, int = 0) { prop_ = prop; }
  // nonRigidICP::solve (nonRigidICP.cu:557-704): Gauss-Newton on the node
  // transforms against one depth frame seen from pose (camera-to-world)
  std::array<double, 4> solve(geometry::defGraph &graph, blender::dqBlender &, const rgbdImage &target,
                              const double pose[16]) {
    std::array<double, 4> st{};
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_solve(graph.handle(), target.depth, &target.sensor, pose, &p, st.data()));
    return st;
  }

 private:
  NonRigidProperties prop_;
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
