// Drives the reference-shaped C++ classes through the C ABI, as the
// reference's own callers do: builder::buildModelLS's gnSolver call
// (include/builder/builder.cu:94-99) and the non-rigid test program
// (src/test/testDynaMap.cpp:100-190).  Synthetic walls and patches are
// integrated, raycast, tracked and deformed.  Exit code 0 = pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dynamap/dynamap.hpp"

int main() {
  using namespace DynaMap;
  rgbdSensor s{120, 160, 150.f, 150.f, 79.5f, 59.5f, 1.f};
  tsdf::tsdfProperties p{0.01f, 1 << 14, 10, 1 << 14, 8, 64, 0.03f, 5.0f, 0.1f};
  std::vector<float> depth(120 * 160, 1.0f);
  rgbdImage img{depth.data(), nullptr, s};
  tsdf::tsdfVolume vol;
  vol.init(p, 120, 160);
  float4x4 I;
  vol.IntegrateScan(img, I);
  const int nb = vol.GetNumAllocatedBlocks();
  std::vector<float> d = vol.GenerateDepth(I, s);
  const float centre = d[60 * 160 + 80];
  std::printf("blocks %d centre depth %.5f\n", nb, centre);
  if (nb <= 0 || std::fabs(centre - 1.0f) > 2e-3f) return 1;
  // mesh of the wall in a box (tsdfVolume::ExtractMesh): vertices at z = 1
  geometry::Mesh mesh = vol.ExtractMesh(float3{-0.2f, -0.2f, 0.5f}, float3{0.2f, 0.2f, 1.5f});
  std::printf("triangles %u\n", mesh.num_triangles_);
  if (mesh.num_triangles_ < 100 || mesh.triangles_.size() != mesh.num_triangles_) return 5;
  for (const auto &t : mesh.triangles_)
    if (std::fabs(t.v0.position.z - 1.0f) > 2e-3f || std::fabs(t.v2.position.z - 1.0f) > 2e-3f) return 6;
  geometry::MeshExtractor mx;
  mx.Init(10);
  mx.ExtractMesh(&vol);
  if (mx.GetMesh().triangles_.size() != 10 || mx.GetMesh().num_triangles_ < mesh.num_triangles_) return 7;
  if (mesh.SaveToFile("/tmp/dm_wall_mesh.obj") != (int)mesh.num_triangles_) return 8;
  // builder: two identical frames of a single plane: the motion along the
  // plane normal (z) is constrained and must be zero; in-plane sliding is not
  // observable by point-to-plane ICP and is only kept small by the damping
  builder b;
  solver::Properties op;
  op.minDepth = 0.1f;
  b.init(p, op, s);
  b.buildModelLS(img);
  auto T = b.buildModelLS(img, true);
  std::printf("pose t = %.3g %.3g %.3g\n", T[3], T[7], T[11]);
  if (std::fabs(T[11]) > 1e-5 || std::fabs(T[3]) > 2e-3 || std::fabs(T[7]) > 2e-3) return 2;
  try {
    tsdf::tsdfProperties bad = p;
    bad.block_size = 4;
    tsdf::tsdfVolume v2;
    v2.init(bad);
    return 3;
  } catch (const Error &e) {
    if (e.code != DM_E_INVALID) return 4;
  }
  // gnSolver: a wall at 1.02 m registered to a wall at 1.00 m (the motion
  // along the normal is observable: tz = -0.02)
  {
    std::vector<float> near(120 * 160, 1.02f);
    solver::gnSolver g;
    solver::Properties gp;
    gp.minDepth = 0.1f;
    g.init(gp, s, p);
    auto Tinc = g.solve(img, rgbdImage{near.data(), nullptr, s});
    std::printf("gnSolver tz %.5f\n", Tinc[11]);
    if (std::fabs(Tinc[11] + 0.02) > 1e-3) return 9;
  }
  // gnSolver driven as builder::buildModelLS does (builder.cu:94-99), with
  // the reference's level order (src/dynaMap.cpp:150-158: scales 4, 2, 1):
  // the source frame sees the wall 2 cm further away, so the camera moved
  // back 2 cm: pose <- pose * T_inc gives tz = -0.02
  {
    solver::Properties OptimizerProp;
    OptimizerProp.lvlIterationNum[0] = 4;
    OptimizerProp.lvlIterationNum[1] = 5;
    OptimizerProp.lvlIterationNum[2] = 10;
    OptimizerProp.lvlScale[0] = 4;
    OptimizerProp.lvlScale[1] = 2;
    OptimizerProp.lvlScale[2] = 1;
    OptimizerProp.minIncrement = 0.001f;
    OptimizerProp.maxIncrement = 0.3f;
    OptimizerProp.regularization = 0.02f;
    OptimizerProp.minDepth = 0.1f;
    solver::gnSolver GNSolver;
    GNSolver.init(OptimizerProp, s);
    pyramid targetImage, sourceImage;
    targetImage.init(s);
    sourceImage.init(s);
    for (uint32_t i = 0; i < s.rows; i++)
      for (uint32_t j = 0; j < s.cols; j++) {
        targetImage.depth[i * s.cols + j] = 1.0f;
        sourceImage.depth[i * s.cols + j] = 1.02f;
        sourceImage.rgb[i * s.cols + j] = make_uchar3(10, 20, 30);
      }
    Matrix4d currentPose = Identity4();
    currentPose[3] = 0.5;  // the increment composes on the right of the camera pose
    Vector6d prev_increment_{};
    geometry::PointCloudXYZ targetPCL, sourcePCL;
    GNSolver.solve(currentPose, prev_increment_, targetImage, sourceImage, targetPCL, sourcePCL);
    const size_t c = 60 * 160 + 80;
    std::printf("builder-shaped gnSolver: t = %.4f %.4f %.4f, increment tz %.4f, cloud z %.3f / %.3f\n",
                currentPose[3], currentPose[7], currentPose[11], prev_increment_[5], targetPCL.points[c].z,
                sourcePCL.points[c].z);
    if (std::fabs(currentPose[11] + 0.02) > 1e-3 || std::fabs(currentPose[3] - 0.5) > 2e-3) return 12;
    if (std::fabs(prev_increment_[5] + 0.02) > 1e-3) return 13;
    if (targetPCL.points.size() != (size_t)s.rows * s.cols || std::fabs(targetPCL.points[c].z - 1.0f) > 1e-6f ||
        std::fabs(sourcePCL.points[c].z - 1.02f) > 1e-6f || std::fabs(targetPCL.normals[c].z) < 0.99f)
      return 14;
    // a copy of a frame keeps its own buffers (targetImage = sourceImage)
    targetImage = sourceImage;
    sourceImage.depth[c] = 3.0f;
    if (targetImage.depth[c] != 1.02f || targetImage.rgb[c].y != 20) return 15;
    try {
      solver::Properties bad = OptimizerProp;
      bad.lvlScale[0] = 3;
      solver::gnSolver g2;
      g2.init(bad, s);
      return 16;
    } catch (const Error &e) {
      if (e.code != DM_E_INVALID) return 17;
    }
  }
  // defGraph + dqBlender + nonRigidICP driven as src/test/testDynaMap.cpp:
  // 100-190 does: graph->init(mesh, rows * cols, 1), defGraphUpdateNodes,
  // Solver->init(graph, OptimizerProp, rows * cols), Solver->solve(graph,
  // blender, liveFrame, PredictedImage, pose, sensor), blendMesh.  A 20 x 20
  // vertex patch at z = 1 m; the live frame sees it 1 cm further away.
  {
    geometry::MeshSTD srcMesh;
    for (int j = 0; j < 20; ++j)
      for (int i = 0; i < 20; ++i) {
        srcMesh.vertices.push_back(float3{0.005f * (i - 10), 0.005f * (j - 10), 1.0f});
        srcMesh.normals.push_back(float3{0.f, 0.f, -1.f});
      }
    geometry::defGraph graph;
    graph.init(srcMesh, (int)(s.rows * s.cols), 1);
    graph.defGraphUpdateNodes();
    std::vector<int32_t> vid = graph.nodeVertexIds();
    if (graph.nodeNum != geometry::defGraph::NODE_NUM || vid[0] != 0 || vid[272] != 272) return 18;
    solver::Properties OptimizerProp;
    OptimizerProp.lvlIterationNum[0] = 5;  // full resolution
    solver::nonRigidICP Solver;
    Solver.init(graph, OptimizerProp, (int)(s.rows * s.cols));
    pyramid liveFrame, PredictedImage;
    liveFrame.init(s);
    PredictedImage.init(s);
    for (uint32_t i = 0; i < s.rows; i++)
      for (uint32_t j = 0; j < s.cols; j++) {
        liveFrame.depth[i * s.cols + j] = 1.01f;
        PredictedImage.depth[i * s.cols + j] = 1.0f;
      }
    Matrix4d pose = Identity4();
    blender::dqBlender blender;
    blender.init(graph);
    std::array<double, 4> st{};
    for (int f = 0; f < 2; ++f) st = Solver.solve(graph, blender, liveFrame, PredictedImage, pose, s);
    geometry::MeshSTD dstMesh;
    blender.blendMesh(dstMesh, graph);
    double worst = 0;
    for (const float3 &v : dstMesh.vertices) worst = std::max(worst, std::fabs((double)v.z - 1.01));
    std::printf("testDynaMap-shaped non-rigid: %d nodes, %.0f residuals, max |z - 1.01| %.2e\n", graph.nodeNum, st[0],
                worst);
    if (st[0] != 400 || worst > 1e-4) return 19;
    // the source image gates the pixels: an unobserved predicted image leaves no residual
    for (uint32_t k = 0; k < s.rows * s.cols; ++k) PredictedImage.depth[k] = 0.f;
    st = Solver.solve(graph, blender, liveFrame, PredictedImage, pose, s);
    if (st[0] != 0) return 20;
    // mesh target (nonRigidICP.cu:706-852): the patch moved back to z = 1.02
    geometry::MeshSTD targetMesh = srcMesh;
    for (float3 &v : targetMesh.vertices) v.z = 1.02f;
    graph.updateActiveNeighbourNodes(targetMesh, pose);
    for (int f = 0; f < 2; ++f) st = Solver.solve(graph, blender, targetMesh, pose);
    blender.blendMesh(dstMesh);
    worst = 0;
    for (const float3 &v : dstMesh.vertices) worst = std::max(worst, std::fabs((double)v.z - 1.02));
    std::printf("mesh-target non-rigid: max |z - 1.02| %.2e\n", worst);
    if (st[0] != 400 || worst > 1e-4) return 21;
    // several graphs at once (no reference counterpart): solveMany on their
    // own streams and a batch along blockIdx.y give the blended meshes of the
    // same solves one graph at a time, bit for bit
    {
      std::vector<geometry::defGraph> gs(6);
      for (auto &g : gs) {
        g.init(srcMesh, (int)(s.rows * s.cols), 1);
        g.defGraphUpdateNodes();
      }
      std::vector<float> near((size_t)s.rows * s.cols, 1.005f), far((size_t)s.rows * s.cols, 1.015f);
      const std::vector<const float *> depths = {near.data(), far.data()};
      const std::vector<Matrix4d> poses = {Identity4(), Identity4()};
      solver::nonRigidICP one;
      one.init(gs[0], OptimizerProp, (int)(s.rows * s.cols));
      rgbdImage a{near.data(), nullptr, s}, b{far.data(), nullptr, s};
      std::vector<geometry::defGraph *> many = {&gs[2], &gs[3]}, bat = {&gs[4], &gs[5]};
      solver::nonRigidBatch batch(bat, one);
      for (int f = 0; f < 2; ++f) {
        one.solve(gs[0], blender, a, poses[0].data());
        one.solve(gs[1], blender, b, poses[1].data());
        one.solveMany(many, depths, s, poses);
        batch.solve(depths, s, poses);
      }
      for (int k = 0; k < 2; ++k) {
        geometry::MeshSTD m0, m1, m2;
        blender.blendMesh(m0, gs[k]);
        blender.blendMesh(m1, gs[2 + k]);
        blender.blendMesh(m2, gs[4 + k]);
        for (size_t q = 0; q < m0.vertices.size(); ++q)
          if (std::memcmp(&m0.vertices[q], &m1.vertices[q], sizeof(float3)) != 0 ||
              std::memcmp(&m0.vertices[q], &m2.vertices[q], sizeof(float3)) != 0)
            return 28;
        if (std::fabs(m0.vertices[0].z - (k ? 1.015f : 1.005f)) > 1e-4f) return 29;
      }
      std::printf("solveMany and nonRigidBatch: blended meshes equal the single solves'\n");
    }
    // updateNodesDQ from (roll, pitch, yaw, tx, ty, tz): a common yaw of 0.1
    // rad and a 1 cm x shift move every vertex rigidly
    std::vector<float> dq((size_t)graph.nodeNum * 6, 0.f);
    for (int n = 0; n < graph.nodeNum; ++n) { dq[6 * n + 2] = 0.1f; dq[6 * n + 3] = 0.01f; }
    graph.updateNodesDQ(dq.data());
    blender.blendMesh(dstMesh);
    worst = 0;
    const double cz = std::cos(0.1), sz = std::sin(0.1);
    for (size_t k = 0; k < srcMesh.vertices.size(); ++k) {
      const float3 &a = srcMesh.vertices[k], &b = dstMesh.vertices[k];
      worst = std::max(worst, std::fabs(b.x - (cz * a.x - sz * a.y + 0.01)));
      worst = std::max(worst, std::fabs(b.y - (sz * a.x + cz * a.y)));
      worst = std::max(worst, std::fabs((double)b.z - a.z));
    }
    std::printf("updateNodesDQ(roll, pitch, yaw, t): max error %.2e\n", worst);
    if (worst > 2e-6) return 22;
    // mode 0: random distinct vertices below ObservationNum
    geometry::defGraph g0;
    g0.init(srcMesh, 300, 0);
    std::vector<int32_t> r = g0.nodeVertexIds();
    std::vector<int32_t> sorted = r;
    std::sort(sorted.begin(), sorted.end());
    if (g0.nodeNum != 273 || std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end() || sorted.back() >= 300 ||
        std::is_sorted(r.begin(), r.end()))
      return 23;
  }
  // defGraph + dqBlender: a common node translation moves every vertex
  {
    geometry::MeshSTD m;
    for (int j = 0; j < 30; ++j)
      for (int i = 0; i < 40; ++i) {
        m.vertices.push_back(float3{0.01f * i, 0.01f * j, 1.0f});
        m.normals.push_back(float3{0.f, 0.f, -1.f});
      }
    geometry::defGraph graph;
    graph.init(m, 0.05f, 0.08f, 1024);
    std::vector<double> T((size_t)graph.nodeNum * 16, 0.0);
    for (int n = 0; n < graph.nodeNum; ++n) {
      for (int k = 0; k < 4; ++k) T[16 * n + 5 * k] = 1.0;
      T[16 * n + 3] = 0.01;
    }
    graph.updateNodesDQ(T);
    blender::dqBlender b;
    b.init(graph);
    geometry::MeshSTD out;
    b.blendMesh(out, graph);
    double worst = 0;
    for (size_t k = 0; k < m.vertices.size(); ++k)
      worst = std::max(worst, (double)std::fabs(out.vertices[k].x - m.vertices[k].x - 0.01f));
    std::printf("defGraph nodes %d max blend error %.2e\n", graph.nodeNum, worst);
    if (graph.nodeNum < 5 || worst > 1e-6) return 10;
  }
  // snapshot / restore of the volume
  {
    const int before = vol.GetNumAllocatedBlocks();
    std::vector<uint8_t> snap = vol.Snapshot();
    vol.Restore(snap);
    if (vol.GetNumAllocatedBlocks() != before || snap.size() < 4096) return 11;
  }
  // HashTable::DeleteBlock: an allocated block goes, a second delete finds nothing
  {
    // block (0, 0, 12) (1 cm voxels, 8 cm blocks) holds the wall at z = 1 m under the centre pixel
    const int before = vol.GetNumAllocatedBlocks();
    if (!vol.DeleteBlock(0, 0, 12) || vol.DeleteBlock(0, 0, 12)) return 25;
    if (vol.GetNumAllocatedBlocks() != before - 1) return 26;
    if (vol.Recycle(float3{0.f, 0.f, 1.f}, 0.0) != before - 1 || vol.GetNumAllocatedBlocks() != 0) return 27;
  }
  std::printf("ok\n");
  return 0;
}
