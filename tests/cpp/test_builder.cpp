// Drives the reference-shaped C++ classes (tsdfVolume, builder) through the
// C ABI, as the reference's own test programs do (src/test/testDynaMap.cpp):
// a synthetic wall is integrated, raycast and tracked.  Exit code 0 = pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
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
  std::printf("ok\n");
  return 0;
}
