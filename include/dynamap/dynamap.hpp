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
#include <cmath>
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
struct uchar3 {
  uint8_t x, y, z;
};
inline uchar3 make_uchar3(uint8_t x, uint8_t y, uint8_t z) { return uchar3{x, y, z}; }

// Eigen::Matrix4f / Eigen::Matrix<double, 6, 1> of the reference's solver
// signatures (gnSolver.h:72-76, nonRigidICP.h:112-122) as PODs: a row-major
// camera-to-world pose and a twist (omega, t)
using Matrix4d = std::array<double, 16>;
using Vector6d = std::array<double, 6>;
inline Matrix4d Identity4() { return Matrix4d{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; }

// rgbdImage (include/image/sensor/rgbdImage.h) over host buffers
struct rgbdImage {
  const float *depth = nullptr;   // rows*cols metres
  const uint8_t *rgb = nullptr;   // rows*cols*3, may be null
  rgbdSensor sensor{};
};

// pyramid (include/image/pyramid/pyramid.h): a frame the caller fills in
// place after init, as the reference's test programs do
// (liveFrame.depth[i] = ..., liveFrame.rgb[i] = make_uchar3(...))
class pyramid {
 public:
  rgbdSensor sensor{};
  float *depth = nullptr;
  uchar3 *rgb = nullptr;
  pyramid() = default;
  pyramid(const pyramid &o) { *this = o; }
  pyramid &operator=(const pyramid &o) {
    sensor = o.sensor;
    depth_ = o.depth_;
    rgb_ = o.rgb_;
    depth = depth_.empty() ? nullptr : depth_.data();
    rgb = rgb_.empty() ? nullptr : rgb_.data();
    return *this;
  }
  void init(const rgbdSensor &s) {
    sensor = s;
    depth_.assign((size_t)s.rows * s.cols, 0.f);
    rgb_.assign((size_t)s.rows * s.cols, uchar3{0, 0, 0});
    depth = depth_.data();
    rgb = rgb_.data();
  }
  rgbdImage image() const { return rgbdImage{depth, rgb ? &rgb->x : nullptr, sensor}; }

 private:
  std::vector<float> depth_;
  std::vector<uchar3> rgb_;
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

// PointCloudXYZ (geom_Types.h): a vertex map and its normals
struct PointCloudXYZ {
  std::vector<float3> points, normals;
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
  // HashTable::DeleteBlock (hashTable.cu:112-134): true when the block was allocated
  bool DeleteBlock(int x, int y, int z) {
    const int32_t xyz[3] = {x, y, z};
    int64_t n = 0;
    check(dm_volume_delete_blocks(v_, xyz, 1, &n));
    return n == 1;
  }
  // pool recycling: delete every block whose centre lies farther than radius from center
  int64_t Recycle(const float3 &center, double radius) {
    const double c[3] = {center.x, center.y, center.z};
    int64_t n = 0;
    check(dm_volume_recycle(v_, c, radius, &n));
    return n;
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

// solver::Properties (gnSolver.h:47-55).  lvlIterationNum[p] iterations run
// at the pyramid level of scale lvlScale[p] (1 = full resolution, 2 = half,
// 4 = quarter; any order of the three, as the reference's src/dynaMap.cpp:
// 150-155 gives {4, 2, 1}); levels run coarse to fine.  maxIteration and
// huberConstant are kept for source compatibility: the reference computes
// but never applies the robust weights (gnSolver.cu:116-119).
struct Properties {
  int lvlIterationNum[3] = {10, 5, 4};
  int lvlScale[3] = {1, 2, 4};
  int maxIteration = 20;
  float minIncrement = 0.001f;  // src/dynaMap.cpp:156
  float maxIncrement = 0.3f;
  float regularization = 0.02f;
  float huberConstant = 0.3f;
  // the build's gates and weights (the reference hard-codes DEPTH_MIN/MAX,
  // pyramid.h:15-16, and sets the photometric weight to 0, gnSolver.cu:180)
  float minDepth = 0.3f, maxDepth = 5.0f;
  float distMax = 0.1f;
  float photoWeight = 0.0f;
  dm_icp_props c() const {
    dm_icp_props p{};
    bool seen[3] = {false, false, false};
    for (int k = 0; k < 3; ++k) {
      const int l = lvlScale[k] == 1 ? 0 : lvlScale[k] == 2 ? 1 : lvlScale[k] == 4 ? 2 : -1;
      if (l < 0 || seen[l]) throw Error(DM_E_INVALID, "dynamap: lvlScale must hold 1, 2 and 4");
      seen[l] = true;
      p.lvl_iterations[l] = lvlIterationNum[k];
    }
    p.min_depth = minDepth; p.max_depth = maxDepth; p.dist_max = distMax;
    p.photo_weight = photoWeight; p.regularization = regularization; p.max_increment = maxIncrement;
    p.min_increment = minIncrement;
    return p;
  }
  // iterations at full resolution (the level the non-rigid solve runs on,
  // nonRigidICP.cu:580-583)
  int fullResolutionIterations() const {
    for (int k = 0; k < 3; ++k)
      if (lvlScale[k] == 1) return lvlIterationNum[k];
    return 0;
  }
};

inline Matrix4d mul4(const Matrix4d &A, const Matrix4d &B) {
  Matrix4d C{};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += A[4 * r + k] * B[4 * k + c];
      C[4 * r + c] = s;
    }
  return C;
}

// gnSolver (gnSolver.h:57-100): multi-level projective ICP of a source frame
// against target maps, on the GPU.
class gnSolver {
 public:
  gnSolver() = default;
  gnSolver(const gnSolver &) = delete;
  gnSolver &operator=(const gnSolver &) = delete;
  ~gnSolver() {
    if (s_) dm_session_destroy(s_);
  }
  // gnSolver::init (gnSolver.cu:287-315).  The tracker needs no voxel
  // volume: it is created with a minimal one.
  void init(const Properties &prop, const rgbdSensor &sensor) {
    const tsdf::tsdfProperties tp{0.01f, 1 << 10, 10, 1 << 10, 8, 64, 0.03f, prop.maxDepth, prop.minDepth};
    init(prop, sensor, tp);
  }
  void init(const Properties &prop, const rgbdSensor &sensor, const tsdf::tsdfProperties &tp) {
    if (s_) dm_session_destroy(s_);
    s_ = nullptr;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    dm_icp_props ip = prop.c();
    check(dm_session_create(&tp, &ip, &sensor, I, &s_));
    rows_ = sensor.rows;
    cols_ = sensor.cols;
  }
  // gnSolver::solve (gnSolver.cu:317-438), the reference's call shape
  // (builder.cu:94-99): registers sourceImage to targetImage, updates pose
  // with the increment (pose <- pose * T_inc, camera-to-world; T_inc maps
  // source-camera points into the target camera) and returns the summed
  // twist in prev_increment_.  targetCloud / sourceCloud receive the
  // full-resolution vertex and normal maps (camera coordinates; NaN normal =
  // none), as the reference's getPointCloudXYZ / getNormalsfromVertices fill
  // them.
  void solve(Matrix4d &pose, Vector6d &prev_increment_, const rgbdImage &targetImage, const rgbdImage &sourceImage,
             geometry::PointCloudXYZ &targetCloud, geometry::PointCloudXYZ &sourceCloud) {
    const Matrix4d Tinc = track(targetImage, sourceImage, &targetCloud, &sourceCloud);
    check(dm_session_icp_increment(s_, prev_increment_.data()));
    pose = mul4(pose, Tinc);
  }
  void solve(Matrix4d &pose, Vector6d &prev_increment_, pyramid &targetImage, pyramid &sourceImage,
             geometry::PointCloudXYZ &targetCloud, geometry::PointCloudXYZ &sourceCloud) {
    solve(pose, prev_increment_, targetImage.image(), sourceImage.image(), targetCloud, sourceCloud);
  }
  // the increment alone (source camera -> target camera); solve_on_host
  // keeps the reference's host 6x6 steps
  Matrix4d solve(const rgbdImage &target, const rgbdImage &source, bool solve_on_host = false) {
    return track(target, source, nullptr, nullptr, solve_on_host);
  }

 private:
  Matrix4d track(const rgbdImage &target, const rgbdImage &source, geometry::PointCloudXYZ *tc,
                 geometry::PointCloudXYZ *sc, bool solve_on_host = false) {
    for (int l = 0; l < 3; ++l) {
      const size_t P = (size_t)(rows_ >> l) * (cols_ >> l);
      std::vector<float> v(P * 3), n(P * 3), in(P);
      check(dm_session_source_maps(s_, target.depth, target.rgb, l, v.data(), n.data(), in.data()));
      check(dm_session_set_model_maps(s_, l, v.data(), n.data(), in.data()));
      if (l == 0 && tc) to_cloud(v, n, *tc);
    }
    const size_t P = (size_t)rows_ * cols_;
    std::vector<float> v(P * 3), n(sc ? P * 3 : 0);
    check(dm_session_source_maps(s_, source.depth, source.rgb, 0, v.data(), sc ? n.data() : nullptr, nullptr));
    if (sc) to_cloud(v, n, *sc);
    Matrix4d T{};
    check(dm_session_icp_track(s_, solve_on_host ? 1 : 0, T.data()));
    return T;
  }
  static void to_cloud(const std::vector<float> &v, const std::vector<float> &n, geometry::PointCloudXYZ &c) {
    c.points.resize(v.size() / 3);
    c.normals.resize(v.size() / 3);
    for (size_t i = 0; i < c.points.size(); ++i) {
      c.points[i] = float3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
      c.normals[i] = float3{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
    }
  }
  dm_session *s_ = nullptr;
  uint32_t rows_ = 0, cols_ = 0;
};

// nonRigidICP properties of the build's formulation (Gauss-Newton
// iterations, PCG, ARAP weight, association gates)
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

// defGraph (defGraph.h:45-87): nodes of the canonical mesh, their
// neighbours and weights, node transforms; on the GPU.
class defGraph {
 public:
  static constexpr int NODE_NUM = 273;  // include/DynaMap.h:22
  static constexpr int KNN = 4;         // include/DynaMap.h:18
  defGraph() = default;
  defGraph(const defGraph &) = delete;
  defGraph &operator=(const defGraph &) = delete;
  ~defGraph() { Free(); }
  // defGraph::init(MeshSTD&, int ObservationNum, int mode) (defGraph.cu:15-69)
  // with sampleDownMeshHost (:673-702): min(nodes, ObservationNum) nodes
  // (nodes = NODE_NUM unless given), mode 1 the first mesh vertices in
  // order, mode 0 distinct random vertices below ObservationNum (seeded by
  // `seed`; the reference seeds rand() with the time).  Weights
  // exp(-d^2/dgw^2) with dgw = weightScale (include/DynaMap.h:24: 0.015).
  void init(const MeshSTD &srcMesh, int ObservationNum, int mode = 0, int nodes = NODE_NUM) {
    Free();
    const int nv = (int)srcMesh.vertices.size();
    const int range = std::min(ObservationNum, nv);
    const int n = std::min(nodes, range);
    std::vector<int32_t> ids((size_t)std::max(n, 0));
    if (mode == 1) {
      for (int k = 0; k < n; ++k) ids[k] = k;
    } else {
      // a partial Fisher-Yates shuffle of [0, range) driven by splitmix64
      std::vector<int32_t> pool((size_t)std::max(range, 0));
      for (int k = 0; k < range; ++k) pool[k] = k;
      uint64_t x = seed;
      for (int k = 0; k < n; ++k) {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const int j = k + (int)(z % (uint64_t)(range - k));
        std::swap(pool[k], pool[j]);
        ids[k] = pool[k];
      }
    }
    check(dm_graph_create_nodes(&srcMesh.vertices[0].x, &srcMesh.normals[0].x, (int64_t)nv, ids.data(), n,
                                weightScale, &g_));
    finish(srcMesh);
  }
  // the build's own sampling: per grid cell of edge nodeRadius the vertex
  // closest to the cell centre
  void init(const MeshSTD &srcMesh, float nodeRadius, float dgw, int maxNodes = 4096) {
    Free();
    dm_graph_props p{nodeRadius, dgw, maxNodes};
    check(dm_graph_create(&srcMesh.vertices[0].x, &srcMesh.normals[0].x, (int64_t)srcMesh.vertices.size(), &p, &g_));
    finish(srcMesh);
  }
  // defGraph::defGraphUpdateNodes (defGraph.cu:483-545): node neighbours and
  // weights (and the vertices' nearest nodes) from the canonical positions
  void defGraphUpdateNodes() { check(dm_graph_update_nodes(g_)); }
  // defGraph::updateActiveNeighbourNodes (defGraph.cu:243-302, :340-395):
  // the nearest nodes of the observations.  Observations are the canonical
  // vertices (a target mesh corresponds vertex for vertex; a frame is
  // associated through the warped canonical mesh), so this recomputes the
  // vertices' neighbours.
  void updateActiveNeighbourNodes(const MeshSTD &targetMesh, const Matrix4d &) {
    if (targetMesh.vertices.size() != verticesNum)
      throw Error(DM_E_INVALID, "dynamap: the target mesh needs the canonical mesh's vertex count");
    defGraphUpdateNodes();
  }
  void updateActiveNeighbourNodes(const pyramid &, const rgbdSensor &, const Matrix4d &) { defGraphUpdateNodes(); }
  void Free() {
    if (g_) dm_graph_destroy(g_);
    g_ = nullptr;
  }
  // defGraph::updateNodesDQ (defGraph.cu:704-730): per node 6 floats (roll,
  // pitch, yaw, tx, ty, tz); the rotation of Quaternion(EulerAngles)
  // (Quaternion.h:83-96: q = q_z(yaw) q_y(pitch) q_x(roll)), the translation t
  void updateNodesDQ(const float *dqVector) {
    std::vector<double> T((size_t)nodeNum * 16);
    for (int i = 0; i < nodeNum; ++i) {
      const double roll = dqVector[6 * i], pitch = dqVector[6 * i + 1], yaw = dqVector[6 * i + 2];
      const double cr = std::cos(roll * 0.5), sr = std::sin(roll * 0.5), cp = std::cos(pitch * 0.5),
                   sp = std::sin(pitch * 0.5), cy = std::cos(yaw * 0.5), sy = std::sin(yaw * 0.5);
      const double w = cy * cp * cr + sy * sp * sr, x = cy * cp * sr - sy * sp * cr, y = sy * cp * sr + cy * sp * cr,
                   z = sy * cp * cr - cy * sp * sr;
      double *M = &T[16 * (size_t)i];
      M[0] = 1 - 2 * (y * y + z * z); M[1] = 2 * (x * y - w * z);     M[2] = 2 * (x * z + w * y);
      M[4] = 2 * (x * y + w * z);     M[5] = 1 - 2 * (x * x + z * z); M[6] = 2 * (y * z - w * x);
      M[8] = 2 * (x * z - w * y);     M[9] = 2 * (y * z + w * x);     M[10] = 1 - 2 * (x * x + y * y);
      M[3] = dqVector[6 * i + 3]; M[7] = dqVector[6 * i + 4]; M[11] = dqVector[6 * i + 5];
      M[12] = M[13] = M[14] = 0.0; M[15] = 1.0;
    }
    updateNodesDQ(T);
  }
  // node transforms directly, nodeNum x 16 row-major
  void updateNodesDQ(const std::vector<double> &T) { check(dm_graph_set_transforms(g_, T.data())); }
  std::vector<double> nodeTransforms() {
    std::vector<double> T((size_t)nodeNum * 16);
    check(dm_graph_get_transforms(g_, T.data()));
    return T;
  }
  // defGraphNode::vertexID and nIds (defGraph.h:35-43)
  std::vector<int32_t> nodeVertexIds() {
    std::vector<int32_t> v((size_t)nodeNum);
    check(dm_graph_get_nodes(g_, v.data(), nullptr, nullptr));
    return v;
  }
  std::vector<int32_t> nodeNeighbours() {
    std::vector<int32_t> v((size_t)nodeNum * KNN);
    check(dm_graph_get_nodes(g_, nullptr, v.data(), nullptr));
    return v;
  }
  dm_graph *handle() { return g_; }
  int nodeNum = 0;
  int nNum = KNN;
  size_t verticesNum = 0;
  float weightScale = 0.015f;
  uint64_t seed = 0;

 private:
  void finish(const MeshSTD &m) {
    int32_t n = 0;
    check(dm_graph_num_nodes(g_, &n));
    nodeNum = n;
    verticesNum = m.vertices.size();
  }
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
  void blendMesh(geometry::MeshSTD &dst) {
    if (!graph_) throw Error(DM_E_STATE, "dynamap: dqBlender::init was not called");
    blendMesh(dst, *graph_);
  }

 private:
  geometry::defGraph *graph_ = nullptr;
};

}  // namespace blender

namespace solver {

// nonRigidICP (nonRigidICP.h:59-183)
class nonRigidICP {
 public:
  // nonRigidICP::init(warpField, Prop, problemSize) (nonRigidICP.cu:18-40):
  // the Gauss-Newton iterations are those of the full-resolution level
  // (the level the reference's solve runs), the frame buffers are sized for
  // problemSize observations
  void init(geometry::defGraph &warpField, const Properties &prop, int problemSize) {
    NonRigidProperties p = prop_;
    p.gnIterations = prop.fullResolutionIterations();
    init(warpField, p, problemSize);
  }
  void init(geometry::defGraph &warpField, const NonRigidProperties &prop, int problemSize = 0) {
    prop_ = prop;
    check(dm_nonrigid_reserve(warpField.handle(), problemSize, prop.pcgIterations));
  }
  // nonRigidICP::solve(warpField, blender, targetImage, sourceImage, pose,
  // sensor) (nonRigidICP.cu:557-704): the warped canonical mesh against the
  // target (live) frame seen from pose (camera-to-world); a pixel takes part
  // only where the source (predicted) image's depth is in range too
  std::array<double, 4> solve(geometry::defGraph &warpField, blender::dqBlender &, pyramid &targetImage,
                              pyramid &sourceImage, Matrix4d pose, rgbdSensor sensor) {
    std::array<double, 4> st{};
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_solve_pair(warpField.handle(), targetImage.depth, sourceImage.depth, &sensor, pose.data(), &p,
                                 st.data()));
    return st;
  }
  // nonRigidICP::solve(warpField, blender, targetMesh, pose)
  // (nonRigidICP.cu:706-852): target vertex v (camera coordinates) matches
  // canonical vertex v
  std::array<double, 4> solve(geometry::defGraph &warpField, blender::dqBlender &, const geometry::MeshSTD &targetMesh,
                              Matrix4d pose) {
    if (targetMesh.vertices.size() != warpField.verticesNum || targetMesh.normals.size() != warpField.verticesNum)
      throw Error(DM_E_INVALID, "dynamap: the target mesh needs the canonical mesh's vertex count");
    std::array<double, 4> st{};
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_solve_mesh(warpField.handle(), &targetMesh.vertices[0].x, &targetMesh.normals[0].x, pose.data(),
                                 &p, st.data()));
    return st;
  }
  // against one depth frame (no source image)
  std::array<double, 4> solve(geometry::defGraph &graph, blender::dqBlender &, const rgbdImage &target,
                              const double pose[16]) {
    std::array<double, 4> st{};
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_solve(graph.handle(), target.depth, &target.sensor, pose, &p, st.data()));
    return st;
  }
  // no reference counterpart: one solve per graph (distinct graphs), graph i
  // against depths[i] (rows * cols metres) seen from poses[i], all enqueued
  // before any is waited for; each result equals the graph's own solve
  std::vector<std::array<double, 4>> solveMany(const std::vector<geometry::defGraph *> &graphs,
                                               const std::vector<const float *> &depths, const rgbdSensor &sensor,
                                               const std::vector<Matrix4d> &poses) {
    const size_t n = graphs.size();
    if (depths.size() != n || poses.size() != n)
      throw Error(DM_E_INVALID, "dynamap: one depth frame and one pose per graph");
    std::vector<dm_graph *> h(n);
    std::vector<double> P(16 * n);
    for (size_t i = 0; i < n; ++i) {
      h[i] = graphs[i]->handle();
      std::copy(poses[i].begin(), poses[i].end(), P.begin() + 16 * i);
    }
    std::vector<std::array<double, 4>> st(n);
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_solve_many(h.data(), (int32_t)n, depths.data(), &sensor, P.data(), &p, st[0].data()));
    return st;
  }
  const NonRigidProperties &properties() const { return prop_; }

 private:
  NonRigidProperties prop_;
};

// no reference counterpart: a fixed set of distinct graphs whose solves share
// launches (graph = blockIdx.y); destroy it before its graphs
class nonRigidBatch {
 public:
  nonRigidBatch(const std::vector<geometry::defGraph *> &graphs, const nonRigidICP &solver)
      : prop_(solver.properties()), n_(graphs.size()) {
    std::vector<dm_graph *> h(n_);
    for (size_t i = 0; i < n_; ++i) h[i] = graphs[i]->handle();
    check(dm_nonrigid_batch_create(h.data(), (int32_t)n_, &b_));
  }
  nonRigidBatch(const nonRigidBatch &) = delete;
  nonRigidBatch &operator=(const nonRigidBatch &) = delete;
  ~nonRigidBatch() {
    if (b_) dm_nonrigid_batch_destroy(b_);
  }
  std::vector<std::array<double, 4>> solve(const std::vector<const float *> &depths, const rgbdSensor &sensor,
                                           const std::vector<Matrix4d> &poses) {
    if (depths.size() != n_ || poses.size() != n_)
      throw Error(DM_E_INVALID, "dynamap: one depth frame and one pose per graph");
    std::vector<double> P(16 * n_);
    for (size_t i = 0; i < n_; ++i) std::copy(poses[i].begin(), poses[i].end(), P.begin() + 16 * i);
    std::vector<std::array<double, 4>> st(n_);
    dm_nonrigid_props p = prop_.c();
    check(dm_nonrigid_batch_solve(b_, depths.data(), &sensor, P.data(), &p, st[0].data()));
    return st;
  }

 private:
  NonRigidProperties prop_;
  size_t n_ = 0;
  dm_nonrigid_batch *b_ = nullptr;
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
