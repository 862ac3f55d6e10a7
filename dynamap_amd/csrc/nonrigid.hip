// Deformation graph, dual-quaternion blending and the non-rigid
// Gauss-Newton step.
//
// Reference (paths relative to the reference root):
//   include/geometry/defGraph/defGraph.cu      defGraph::init, defGraphUpdateNodes,
//                                              updateNodeDistnacesKernel, sortNeighbourNodesKernel,
//                                              updateNodeWeightsKernel, updateNodesDQ
//   include/blender/dqBlender/dqBlender.cu     blendVertex :162-186, blendMesh :187-218
//   include/math/dualQuat.h, Quaternion.h      transformPosition :293-302, rotate :278-282
//   include/solver/nonRigidICP/nonRigidICP.cu  solve :557-704, getRegResiduals :494-515
//
// Layout (one dm_graph): canonical vertices / normals (float3), per-vertex
// 4 nearest nodes (int32) with raw and normalised weights, per-node vertex
// id, 4 nearest nodes and weights, node transforms T_i (FP64 4x4) and their
// unit dual quaternions (FP32 x 8).  The normal equations are block-sparse:
// the 6x6 blocks (a, b >= a) that some data residual or ARAP edge couples,
// sorted by (a, b), each with the list of its contributions (vertex and
// neighbour slots, or ARAP edge and role) in a stable order fixed when the
// graph is built, so every sum has a fixed order and the solve is
// deterministic.  A row index (block, transposed?) sorted by column serves
// the symmetric SpMV.
//
// Per Gauss-Newton iteration (all on the graph's stream, no host round
// trip): blend -> per-vertex data linearisation (association in FP64) ->
// per-edge ARAP terms -> block assembly (one wavefront per block, FP64
// accumulation, FP32 storage; diagonal blocks also produce the rhs, the
// damping and the FP64 inverse used as block-Jacobi preconditioner) ->
// PCG with FP64 vectors and fixed-order FP64 dot products -> per-node
// T_i <- exp(xi_i) T_i.  The formulation (residuals, Jacobians, gates)
// is the one of the CPU restatement in oracle/dm_oracle_nr.c.
#include <hipcub/hipcub.hpp>
#include <string.h>
#include <algorithm>
#include <vector>
#include "../../include/dynamap_nonrigid.h"
#include "dm_device.h"
#include "dm_scan.h"
#include "dm_solve.h"

namespace dm {

constexpr int kNK = 4;              // neighbours per vertex and per node
constexpr int kMaxNodes = 1 << 15;  // node ids fit 15 bits of the block keys
constexpr int kNodeBits = 15;
constexpr uint32_t kArapTag = 0x80000000u;

struct dq8 { float rw, rx, ry, rz, dw, dx, dy, dz; };

// ---------------------------------------------------------------- graph build
__device__ __forceinline__ uint64_t cell_key(int cx, int cy, int cz) {
  const uint64_t m = (1u << 21) - 1;
  return (1ull << 63) | (((uint64_t)(uint32_t)(cx + (1 << 20)) & m) << 42) |
         (((uint64_t)(uint32_t)(cy + (1 << 20)) & m) << 21) | ((uint64_t)(uint32_t)(cz + (1 << 20)) & m);
}
__device__ __forceinline__ uint32_t cell_hash(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
  return (uint32_t)k;
}
__device__ __forceinline__ void vertex_cell(const float *p, float r, int c[3], float &d2) {
  d2 = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    c[k] = (int)floorf(p[k] / r);
    const float e = p[k] - ((float)c[k] + 0.5f) * r;
    d2 = d2 + e * e;
  }
}

// per occupied cell: min over (d2 to the cell centre, vertex index)
__global__ __launch_bounds__(256) void k_nr_cells(const float *__restrict__ verts, int64_t nv, float r,
                                                 uint64_t *__restrict__ keys, unsigned long long *__restrict__ vals,
                                                 uint32_t mask) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  int c[3];
  float d2;
  vertex_cell(verts + 3 * v, r, c, d2);
  const uint64_t key = cell_key(c[0], c[1], c[2]);
  const unsigned long long val = ((unsigned long long)__float_as_uint(d2) << 32) | (unsigned long long)v;
  uint32_t s = cell_hash(key) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {  // the table has >= 2 nv slots: a free slot exists
    const unsigned long long prev = atomicCAS((unsigned long long *)&keys[s], 0ull, (unsigned long long)key);
    if (prev == 0ull || prev == key) {
      atomicMin(&vals[s], val);
      return;
    }
    s = (s + 1) & mask;
  }
}

__global__ __launch_bounds__(256) void k_nr_mark(const float *__restrict__ verts, int64_t nv, float r,
                                                const uint64_t *__restrict__ keys,
                                                const unsigned long long *__restrict__ vals, uint32_t mask,
                                                int *__restrict__ isnode) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  int c[3];
  float d2;
  vertex_cell(verts + 3 * v, r, c, d2);
  const uint64_t key = cell_key(c[0], c[1], c[2]);
  uint32_t s = cell_hash(key) & mask;
  int flag = 0;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    if (keys[s] == key) {
      flag = (vals[s] & 0xffffffffull) == (unsigned long long)v;
      break;
    }
    if (keys[s] == 0) break;
    s = (s + 1) & mask;
  }
  isnode[v] = flag;
}

__global__ __launch_bounds__(256) void k_nr_nodes(const int *__restrict__ isnode, const long long *__restrict__ rank,
                                                 int64_t nv, int *__restrict__ node_vid,
                                                 int *__restrict__ node_of_vertex) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  if (isnode[v]) {
    node_vid[rank[v]] = (int)v;
    node_of_vertex[v] = (int)rank[v];
  } else {
    node_of_vertex[v] = -1;
  }
}

__device__ __forceinline__ float clamp_dist(float d2) {
  const float d = sqrtf(d2);
  return d < 1e-4f ? 1e-4f : d;
}
__device__ __forceinline__ float weight_of(float d, float dgw) {
  return (float)exp(-(double)((d * d) / (dgw * dgw)));
}

// K nearest nodes of each query point (nodes streamed through LDS);
// distance ascending, ties to the lower node index; `self_is_node` excludes
// the query's own node (node queries).  Also the weights and, for vertex
// queries, the weights normalised in FP64.
constexpr int kKnnTile = 1024;
__global__ __launch_bounds__(256) void k_nr_knn(const float *__restrict__ verts, const int *__restrict__ node_vid,
                                               int nn, const int *__restrict__ query_vid, int64_t nq, float dgw,
                                               int *__restrict__ ids, float *__restrict__ w,
                                               float *__restrict__ wn) {
  __shared__ float tile[kKnnTile * 3];
  const int64_t qi = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool active = qi < nq;
  const int64_t qv = active ? (query_vid ? (int64_t)query_vid[qi] : qi) : 0;
  const int self = query_vid && active ? (int)qi : -1;
  const float q0 = verts[3 * qv], q1 = verts[3 * qv + 1], q2 = verts[3 * qv + 2];
  float bd[kNK];
  int bi[kNK];
#pragma unroll
  for (int k = 0; k < kNK; ++k) { bd[k] = INFINITY; bi[k] = -1; }
  for (int base = 0; base < nn; base += kKnnTile) {
    const int m = min(kKnnTile, nn - base);
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += 256) {
      const int64_t nvid = node_vid[base + t];
      tile[3 * t] = verts[3 * nvid];
      tile[3 * t + 1] = verts[3 * nvid + 1];
      tile[3 * t + 2] = verts[3 * nvid + 2];
    }
    __syncthreads();
    if (!active) continue;
    for (int t = 0; t < m; ++t) {
      const int j = base + t;
      const float dx = q0 - tile[3 * t], dy = q1 - tile[3 * t + 1], dz = q2 - tile[3 * t + 2];
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < bd[kNK - 1]) || j == self) continue;
      // insertion (strict <: equal distances keep the lower index first)
      int k = kNK - 1;
      while (k > 0 && d2 < bd[k - 1]) { bd[k] = bd[k - 1]; bi[k] = bi[k - 1]; --k; }
      bd[k] = d2;
      bi[k] = j;
    }
  }
  if (!active) return;
  double s = 0.0;
  float ww[kNK];
#pragma unroll
  for (int k = 0; k < kNK; ++k) {
    ww[k] = weight_of(clamp_dist(bd[k]), dgw);
    s += (double)ww[k];
    ids[kNK * qi + k] = bi[k];
    w[kNK * qi + k] = ww[k];
  }
  if (wn)
#pragma unroll
    for (int k = 0; k < kNK; ++k) wn[kNK * qi + k] = (float)((double)ww[k] / s);
}

// number of block contributions per vertex (10, or 1 for a node vertex)
__global__ __launch_bounds__(256) void k_nr_contrib_count(const int *__restrict__ node_of_vertex, int64_t nv,
                                                         int *__restrict__ cnt) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v < nv) cnt[v] = node_of_vertex[v] >= 0 ? 1 : 10;
}

__device__ __forceinline__ uint32_t pair_key(int a, int b) {
  return a <= b ? ((uint32_t)a << kNodeBits) | (uint32_t)b : ((uint32_t)b << kNodeBits) | (uint32_t)a;
}

// contributions: data (v, slot a <= slot b) then ARAP (edge, role)
__global__ __launch_bounds__(256) void k_nr_contrib_emit(const int *__restrict__ node_of_vertex,
                                                        const int *__restrict__ vnn, const long long *__restrict__ off,
                                                        int64_t nv, const int *__restrict__ nbr, int nn,
                                                        long long arap_base, uint32_t *__restrict__ keys,
                                                        uint32_t *__restrict__ codes) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < nv) {
    long long o = off[t];
    const int nd = node_of_vertex[t];
    if (nd >= 0) {
      keys[o] = pair_key(nd, nd);
      codes[o] = (uint32_t)(t * 16);
      return;
    }
    for (int a = 0; a < kNK; ++a)
      for (int b = a; b < kNK; ++b) {
        keys[o] = pair_key(vnn[kNK * t + a], vnn[kNK * t + b]);
        codes[o] = (uint32_t)(t * 16 + a * 4 + b);
        ++o;
      }
    return;
  }
  const int64_t e = t - nv;  // ARAP edge e = i * K + k
  if (e >= (int64_t)nn * kNK) return;
  const int i = (int)(e / kNK), j = nbr[e];
  const long long o = arap_base + 3 * e;
  keys[o] = pair_key(i, i);      codes[o] = kArapTag | (uint32_t)(e * 4 + 0);
  keys[o + 1] = pair_key(j, j);  codes[o + 1] = kArapTag | (uint32_t)(e * 4 + 1);
  keys[o + 2] = pair_key(i, j);  codes[o + 2] = kArapTag | (uint32_t)(e * 4 + 2);
}

// Per contribution, the weights of the block's two nodes at the vertex (the
// normalised blend weights of slots a and b; 1 for a vertex that is itself
// a node); ARAP contributions get 0 (unused).
__global__ __launch_bounds__(256) void k_nr_contrib_weights(const uint32_t *__restrict__ codes, long long nc,
                                                           const int *__restrict__ node_of_vertex,
                                                           const float *__restrict__ vnn_wn, float2 *__restrict__ cw) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const uint32_t code = codes[i];
  float2 w = make_float2(0.f, 0.f);
  if (!(code & kArapTag)) {
    const int64_t v = code >> 4;
    if (node_of_vertex[v] >= 0) w = make_float2(1.f, 1.f);
    else w = make_float2(vnn_wn[kNK * v + ((code >> 2) & 3)], vnn_wn[kNK * v + (code & 3)]);
  }
  cw[i] = w;
}

__global__ __launch_bounds__(256) void k_flag_heads(const uint32_t *__restrict__ keys, long long n, uint32_t shift,
                                                   int *__restrict__ flag) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? 1 : 0;
}

// block table from the sorted contribution keys
__global__ __launch_bounds__(256) void k_nr_blocks(const uint32_t *__restrict__ keys, const int *__restrict__ flag,
                                                  const long long *__restrict__ rank, long long n,
                                                  int *__restrict__ blk_a, int *__restrict__ blk_b,
                                                  int *__restrict__ cstart) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const long long k = rank[i];
  blk_a[k] = (int)(keys[i] >> kNodeBits);
  blk_b[k] = (int)(keys[i] & ((1u << kNodeBits) - 1));
  cstart[k] = (int)i;
}

// row entries of the symmetric SpMV: (row, col) -> block | transposed << 31
__global__ __launch_bounds__(256) void k_nr_row_emit(const int *__restrict__ blk_a, const int *__restrict__ blk_b,
                                                    int nb, const long long *__restrict__ off,
                                                    uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nb) return;
  const int a = blk_a[k], b = blk_b[k];
  long long o = off[k];
  keys[o] = ((uint32_t)a << kNodeBits) | (uint32_t)b;
  vals[o] = (uint32_t)k;
  if (a != b) {
    keys[o + 1] = ((uint32_t)b << kNodeBits) | (uint32_t)a;
    vals[o + 1] = (uint32_t)k | 0x80000000u;
  }
}

__global__ __launch_bounds__(256) void k_nr_row_start(const uint32_t *__restrict__ keys, long long n,
                                                     int *__restrict__ row_start, int nn) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i > n) return;
  if (i == n) { row_start[nn] = (int)n; return; }
  const int r = (int)(keys[i] >> kNodeBits);
  const int rp = i == 0 ? -1 : (int)(keys[i - 1] >> kNodeBits);
  for (int q = rp + 1; q <= r; ++q) row_start[q] = (int)i;
}

// SpMV row entries: the block (and transpose bit) with the column node (the
// sorted key's low bits) beside it, so that one load gives both
__global__ __launch_bounds__(256) void k_nr_row_ec(const uint32_t *__restrict__ keys,
                                                  const uint32_t *__restrict__ ent, long long n,
                                                  uint2 *__restrict__ row_ec) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) row_ec[i] = make_uint2(ent[i], keys[i] & ((1u << kNodeBits) - 1u));
}

// ---------------------------------------------------------------- dual quaternions
// unit quaternion of the rotation of T (FP64, Shepperd), w >= 0, and the
// dual part 0.5 (0, t) q; rounded to float
__device__ __forceinline__ dq8 dq_of_T(const double *T) {
  const double m00 = T[0], m01 = T[1], m02 = T[2], m10 = T[4], m11 = T[5], m12 = T[6], m20 = T[8], m21 = T[9],
               m22 = T[10];
  const double tr = m00 + m11 + m22;
  double w, x, y, z;
  if (tr > 0) {
    const double s = sqrt(tr + 1.0) * 2.0;
    w = 0.25 * s; x = (m21 - m12) / s; y = (m02 - m20) / s; z = (m10 - m01) / s;
  } else if (m00 > m11 && m00 > m22) {
    const double s = sqrt(1.0 + m00 - m11 - m22) * 2.0;
    w = (m21 - m12) / s; x = 0.25 * s; y = (m01 + m10) / s; z = (m02 + m20) / s;
  } else if (m11 > m22) {
    const double s = sqrt(1.0 + m11 - m00 - m22) * 2.0;
    w = (m02 - m20) / s; x = (m01 + m10) / s; y = 0.25 * s; z = (m12 + m21) / s;
  } else {
    const double s = sqrt(1.0 + m22 - m00 - m11) * 2.0;
    w = (m10 - m01) / s; x = (m02 + m20) / s; y = (m12 + m21) / s; z = 0.25 * s;
  }
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  const double n = sqrt(w * w + x * x + y * y + z * z);
  const double qw = w / n, qx = x / n, qy = y / n, qz = z / n;
  const double tx = T[3], ty = T[7], tz = T[11];
  const double dw = -(tx * qx + ty * qy + tz * qz);
  const double dx = qw * tx + (ty * qz - tz * qy);
  const double dy = qw * ty + (tz * qx - tx * qz);
  const double dz = qw * tz + (tx * qy - ty * qx);
  return dq8{(float)qw, (float)qx, (float)qy, (float)qz, (float)(0.5 * dw), (float)(0.5 * dx), (float)(0.5 * dy),
             (float)(0.5 * dz)};
}

__global__ __launch_bounds__(256) void k_nr_node_dq(const double *__restrict__ T, int nn, dq8 *__restrict__ dq) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nn) dq[i] = dq_of_T(T + 16 * (size_t)i);
}

// Quaternion::rotate (Quaternion.h:278-282): v + 2 qv x (qv x v + w v)
__device__ __forceinline__ void q_rotate(float w, float x, float y, float z, const float v[3], float o[3]) {
  const float qx = x * 2.f, qy = y * 2.f, qz = z * 2.f;
  const float cx = y * v[2] - z * v[1], cy = z * v[0] - x * v[2], cz = x * v[1] - y * v[0];
  const float ax = cx + v[0] * w, ay = cy + v[1] * w, az = cz + v[2] * w;
  o[0] = v[0] + (qy * az - qz * ay);
  o[1] = v[1] + (qz * ax - qx * az);
  o[2] = v[2] + (qx * ay - qy * ax);
}

// dqBlender::blendVertex (dqBlender.cu:162-186) with Kavan's sign rule and
// one normalisation; a node vertex takes its node's dq (K)
// Blended position and (when nrm) normal of canonical vertex v.
__device__ __forceinline__ void blend_vertex(const float *__restrict__ verts, const float *__restrict__ normals,
                                             const int *__restrict__ node_of_vertex, const int *__restrict__ vnn,
                                             const float *__restrict__ vnn_w, const dq8 *__restrict__ dq, int64_t v,
                                             float pos[3], float *nrm) {
  dq8 b;
  const int nd = node_of_vertex[v];
  if (nd >= 0) {
    b = dq[nd];
  } else {
    b = dq8{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const dq8 r0 = dq[vnn[kNK * v]];
#pragma unroll
    for (int k = 0; k < kNK; ++k) {
      const dq8 q = dq[vnn[kNK * v + k]];
      float w = vnn_w[kNK * v + k];
      const float dt = r0.rw * q.rw + r0.rx * q.rx + r0.ry * q.ry + r0.rz * q.rz;
      if (dt < 0.f) w = -w;
      b.rw = b.rw + w * q.rw; b.rx = b.rx + w * q.rx; b.ry = b.ry + w * q.ry; b.rz = b.rz + w * q.rz;
      b.dw = b.dw + w * q.dw; b.dx = b.dx + w * q.dx; b.dy = b.dy + w * q.dy; b.dz = b.dz + w * q.dz;
    }
    const float n = sqrtf(b.rw * b.rw + b.rx * b.rx + b.ry * b.ry + b.rz * b.rz);
    b.rw = b.rw / n; b.rx = b.rx / n; b.ry = b.ry / n; b.rz = b.rz / n;
    b.dw = b.dw / n; b.dx = b.dx / n; b.dy = b.dy / n; b.dz = b.dz / n;
  }
  const float p[3] = {verts[3 * v], verts[3 * v + 1], verts[3 * v + 2]};
  float rp[3];
  q_rotate(b.rw, b.rx, b.ry, b.rz, p, rp);
  const float tx = ((b.dx * b.rw - b.rx * b.dw) + (b.ry * b.dz - b.rz * b.dy)) * 2.f;
  const float ty = ((b.dy * b.rw - b.ry * b.dw) + (b.rz * b.dx - b.rx * b.dz)) * 2.f;
  const float tz = ((b.dz * b.rw - b.rz * b.dw) + (b.rx * b.dy - b.ry * b.dx)) * 2.f;
  pos[0] = rp[0] + tx;
  pos[1] = rp[1] + ty;
  pos[2] = rp[2] + tz;
  if (nrm) {
    const float n[3] = {normals[3 * v], normals[3 * v + 1], normals[3 * v + 2]};
    q_rotate(b.rw, b.rx, b.ry, b.rz, n, nrm);
  }
}
__device__ __forceinline__ void d_nr_blend(const float *__restrict__ verts, const float *__restrict__ normals,
                                           int64_t nv, const int *__restrict__ node_of_vertex,
                                           const int *__restrict__ vnn, const float *__restrict__ vnn_w,
                                           const dq8 *__restrict__ dq, float *__restrict__ out_v,
                                           float *__restrict__ out_n) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  float pos[3], nrm[3];
  blend_vertex(verts, normals, node_of_vertex, vnn, vnn_w, dq, v, pos, out_n ? nrm : nullptr);
  out_v[3 * v] = pos[0];
  out_v[3 * v + 1] = pos[1];
  out_v[3 * v + 2] = pos[2];
  if (out_n) { out_n[3 * v] = nrm[0]; out_n[3 * v + 1] = nrm[1]; out_n[3 * v + 2] = nrm[2]; }
}
__global__ __launch_bounds__(256) void k_nr_blend(const float *__restrict__ verts, const float *__restrict__ normals,
                                                 int64_t nv, const int *__restrict__ node_of_vertex,
                                                 const int *__restrict__ vnn, const float *__restrict__ vnn_w,
                                                 const dq8 *__restrict__ dq, float *__restrict__ out_v,
                                                 float *__restrict__ out_n) { d_nr_blend(verts, normals, nv, node_of_vertex, vnn, vnn_w, dq, out_v, out_n); }
// ---------------------------------------------------------------- Gauss-Newton
struct NrFrame {
  int rows, cols;
  double fx, fy, cx, cy;
  double pose[16];  // camera-to-world
  double min_depth, max_depth, dist_max, normal_cos, lambda, damping, pcg_tol;
  int use_src;      // gate the depth target by the source (predicted) depth's range
};

// target vertex (camera) per pixel, components rounded to float (w = 1 valid)
__device__ __forceinline__ void d_nr_target_vert(const float *__restrict__ depth, const NrFrame *Fp,
                                                 float4 *__restrict__ tv) {
  const NrFrame &F = *Fp;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= F.rows * F.cols) return;
  const float d = depth[i];
  if (!(d > 0.0f)) { tv[i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
  const int vv = i / F.cols, uu = i - vv * F.cols;
  const double dd = (double)d;
  tv[i] = make_float4((float)((((double)uu - F.cx) * dd) / F.fx), (float)((((double)vv - F.cy) * dd) / F.fy), d, 1.f);
}
__global__ __launch_bounds__(256) void k_nr_target_vert(const float *__restrict__ depth, const NrFrame *Fp,
                                                       float4 *__restrict__ tv) { d_nr_target_vert(depth, Fp, tv); }
struct b_nr_target_vert {
  const float *depth;
  const NrFrame *Fp;
  float4 *tv;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_target_vert for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_target_vert(const b_nr_target_vert *__restrict__ t) {
  const b_nr_target_vert &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_target_vert(a.depth, a.Fp, a.tv);
}

// target normal (camera): cross(V[i+2]-V[i-2], V[i+W]-V[i-W]) in FP64 with
// out-of-range neighbours replaced by the pixel itself; w = 1 valid
__device__ __forceinline__ void d_nr_target_norm(const float4 *__restrict__ tv, const NrFrame *Fp,
                                                 float4 *__restrict__ tn) {
  const NrFrame &F = *Fp;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int size = F.rows * F.cols;
  if (i >= size) return;
  const int idx[5] = {i, i - 2 < 0 ? i : i - 2, i + 2 >= size ? i : i + 2, i - F.cols <= 0 ? i : i - F.cols,
                      i + F.cols >= size ? i : i + F.cols};
  float4 V[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    V[k] = tv[idx[k]];
    if (V[k].w == 0.f) { tn[i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
  }
  const double ax = (double)V[2].x - (double)V[1].x, ay = (double)V[2].y - (double)V[1].y,
               az = (double)V[2].z - (double)V[1].z;
  const double bx = (double)V[4].x - (double)V[3].x, by = (double)V[4].y - (double)V[3].y,
               bz = (double)V[4].z - (double)V[3].z;
  const double c0 = ay * bz - az * by, c1 = az * bx - ax * bz, c2 = ax * by - ay * bx;
  const double nn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  if (!(nn > 0)) { tn[i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
  tn[i] = make_float4((float)(c0 / nn), (float)(c1 / nn), (float)(c2 / nn), 1.f);
}
__global__ __launch_bounds__(256) void k_nr_target_norm(const float4 *__restrict__ tv, const NrFrame *Fp,
                                                       float4 *__restrict__ tn) { d_nr_target_norm(tv, Fp, tn); }
struct b_nr_target_norm {
  const float4 *tv;
  const NrFrame *Fp;
  float4 *tn;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_target_norm for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_target_norm(const b_nr_target_norm *__restrict__ t) {
  const b_nr_target_norm &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_target_norm(a.tv, a.Fp, a.tn);
}

// per-vertex data term: g = [p' x n, n] (world), residual r = n . (p' - q);
// out[8 * v] = g0..g5, r, valid
__device__ __forceinline__ void nr_data_record(const double p[3], const double nt[3], const double q[3],
                                               const double *pose, float *o) {
  double nW[3], qW[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    nW[r] = pose[4 * r] * nt[0] + pose[4 * r + 1] * nt[1] + pose[4 * r + 2] * nt[2];
    qW[r] = pose[4 * r] * q[0] + pose[4 * r + 1] * q[1] + pose[4 * r + 2] * q[2] + pose[4 * r + 3];
  }
  const double res = nW[0] * (p[0] - qW[0]) + nW[1] * (p[1] - qW[1]) + nW[2] * (p[2] - qW[2]);
  o[0] = (float)(p[1] * nW[2] - p[2] * nW[1]);
  o[1] = (float)(p[2] * nW[0] - p[0] * nW[2]);
  o[2] = (float)(p[0] * nW[1] - p[1] * nW[0]);
  o[3] = (float)nW[0]; o[4] = (float)nW[1]; o[5] = (float)nW[2];
  o[6] = (float)res;
  o[7] = 1.f;
}

// source depth (sdepth) gates a pixel when F.use_src: the reference's range
// test on the source image (nonRigidICP.cu:291-292)
// Data-term record of vertex v at blended position pv and normal pn.
__device__ __forceinline__ void data_vertex(const float pv[3], const float pn[3], int64_t v,
                                            const float4 *__restrict__ tv, const float4 *__restrict__ tn,
                                            const float *__restrict__ depth, const float *__restrict__ sdepth,
                                            const NrFrame *Fp, float *__restrict__ out) {
  const NrFrame &F = *Fp;
  float *o = out + 8 * v;
  o[7] = 0.f;
  const double *pose = F.pose;
  const double p[3] = {pv[0], pv[1], pv[2]};
  const double nw[3] = {pn[0], pn[1], pn[2]};
  double pc[3], ncam[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    pc[r] = pose[r] * (p[0] - pose[3]) + pose[4 + r] * (p[1] - pose[7]) + pose[8 + r] * (p[2] - pose[11]);
    ncam[r] = pose[r] * nw[0] + pose[4 + r] * nw[1] + pose[8 + r] * nw[2];
  }
  if (!(pc[2] > 0)) return;
  const double uf = F.fx * pc[0] / pc[2] + F.cx, vf = F.fy * pc[1] / pc[2] + F.cy;
  if (!(uf > -1.0 && vf > -1.0 && uf < (double)F.cols && vf < (double)F.rows)) return;
  const int u = (int)floor(uf + 0.5), vv = (int)floor(vf + 0.5);
  if (u < 0 || vv < 0 || u >= F.cols || vv >= F.rows) return;
  const int pix = vv * F.cols + u;
  const float d = depth[pix];
  if (!(d >= (float)F.min_depth && d <= (float)F.max_depth)) return;
  if (F.use_src) {
    const float ds = sdepth[pix];
    if (!(ds >= (float)F.min_depth && ds <= (float)F.max_depth)) return;
  }
  const float4 qv = tv[pix], qn = tn[pix];
  if (qn.w == 0.f) return;
  const double q[3] = {qv.x, qv.y, qv.z}, nt[3] = {qn.x, qn.y, qn.z};
  const double dx = pc[0] - q[0], dy = pc[1] - q[1], dz = pc[2] - q[2];
  if (dx * dx + dy * dy + dz * dz > F.dist_max * F.dist_max) return;
  if (fabs(ncam[0] * nt[0] + ncam[1] * nt[1] + ncam[2] * nt[2]) < F.normal_cos) return;
  nr_data_record(p, nt, q, pose, o);
}
// mesh target (nonRigidICP.cu:354-376): target vertex v and its normal, in
// camera coordinates, correspond to canonical vertex v; same gates and
// record as the depth target
__device__ __forceinline__ void data_mesh_vertex(const float pv[3], const float pn[3], int64_t v,
                                                 const float *__restrict__ mv, const float *__restrict__ mn,
                                                 const NrFrame *Fp, float *__restrict__ out) {
  const NrFrame &F = *Fp;
  float *o = out + 8 * v;
  o[7] = 0.f;
  const double *pose = F.pose;
  const double p[3] = {pv[0], pv[1], pv[2]};
  const double nw[3] = {pn[0], pn[1], pn[2]};
  double pc[3], ncam[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    pc[r] = pose[r] * (p[0] - pose[3]) + pose[4 + r] * (p[1] - pose[7]) + pose[8 + r] * (p[2] - pose[11]);
    ncam[r] = pose[r] * nw[0] + pose[4 + r] * nw[1] + pose[8 + r] * nw[2];
  }
  const double q[3] = {mv[3 * v], mv[3 * v + 1], mv[3 * v + 2]}, nt[3] = {mn[3 * v], mn[3 * v + 1], mn[3 * v + 2]};
  const double n2 = nt[0] * nt[0] + nt[1] * nt[1] + nt[2] * nt[2];
  if (!(n2 > 0) || !(q[0] == q[0] && q[1] == q[1] && q[2] == q[2])) return;
  const double dx = pc[0] - q[0], dy = pc[1] - q[1], dz = pc[2] - q[2];
  if (dx * dx + dy * dy + dz * dz > F.dist_max * F.dist_max) return;
  if (fabs(ncam[0] * nt[0] + ncam[1] * nt[1] + ncam[2] * nt[2]) < F.normal_cos) return;
  nr_data_record(p, nt, q, pose, o);
}
// ARAP edge e = i * K + k (j = k-th neighbour of i): a = T_i g_j, b = T_j g_j,
// w_ij = max(w_i[k], w_j[k]) (nonRigidICP.cu:494-515); out[8 e] = a, b, w
__device__ __forceinline__ void arap_edge(const double *__restrict__ T, const float *__restrict__ verts,
                                          const int *__restrict__ node_vid, const int *__restrict__ nbr,
                                          const float *__restrict__ nbr_w, int e, double *__restrict__ out) {
  const int i = e / kNK, k = e - i * kNK, j = nbr[e];
  const double *Ti = T + 16 * (size_t)i, *Tj = T + 16 * (size_t)j;
  const int64_t gv = node_vid[j];
  const double g0 = verts[3 * gv], g1 = verts[3 * gv + 1], g2 = verts[3 * gv + 2];
  double *o = out + 8 * (size_t)e;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    o[r] = Ti[4 * r] * g0 + Ti[4 * r + 1] * g1 + Ti[4 * r + 2] * g2 + Ti[4 * r + 3];
    o[3 + r] = Tj[4 * r] * g0 + Tj[4 * r + 1] * g1 + Tj[4 * r + 2] * g2 + Tj[4 * r + 3];
  }
  double wij = (double)nbr_w[e];
  const double wj = (double)nbr_w[kNK * j + k];
  if (wj > wij) wij = wj;
  o[6] = wij;
  o[7] = 0.0;
}
// The linearisation's per-vertex and per-edge terms in one launch:
// workgroups 0 .. vwg - 1 blend vertex v and write its data-term record
// (depth target, or with kMesh the vertex-for-vertex mesh target), the rest
// write the ARAP record of edge e.
struct NrTerms {
  const float *verts, *normals;
  const int *node_of_vertex, *vnn;
  const float *vnn_w;
  const dq8 *dq;
  int64_t nv;
  const float4 *tv, *tn;        // depth target maps
  const float *depth, *sdepth;  // depth target, source gate
  const float *mv, *mn;         // mesh target
  const NrFrame *Fp;
  float *vdata;
  const double *T;
  const int *node_vid, *nbr;
  const float *nbr_w;
  int nn;
  double *edata;
  unsigned vwg;   // workgroups of the vertex part
  unsigned grid;  // all workgroups of this graph's launch
};
template <bool kMesh> __device__ __forceinline__ void d_nr_terms(const NrTerms &a) {
  if (blockIdx.x < a.vwg) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= a.nv) return;
    float pv[3], pn[3];
    blend_vertex(a.verts, a.normals, a.node_of_vertex, a.vnn, a.vnn_w, a.dq, v, pv, pn);
    if constexpr (kMesh) data_mesh_vertex(pv, pn, v, a.mv, a.mn, a.Fp, a.vdata);
    else data_vertex(pv, pn, v, a.tv, a.tn, a.depth, a.sdepth, a.Fp, a.vdata);
  } else {
    const int e = (int)(blockIdx.x - a.vwg) * 256 + threadIdx.x;
    if (e >= a.nn * kNK) return;
    arap_edge(a.T, a.verts, a.node_vid, a.nbr, a.nbr_w, e, a.edata);
  }
}
template <bool kMesh> __global__ __launch_bounds__(256) void k_nr_terms(NrTerms a) { d_nr_terms<kMesh>(a); }
// k_nr_terms (depth target) for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_terms(const NrTerms *__restrict__ t) {
  const NrTerms &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_terms<false>(a);
}

// ARAP Jacobian rows for component c: Ji = w [row c of -[a]x, e_c],
// Jj = -w [row c of -[b]x, e_c]
__device__ __forceinline__ void arap_rows(const double *ed, int c, double Ji[6], double Jj[6]) {
  const double *A = ed, *B = ed + 3;
  const double w = ed[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) { Ji[q] = 0.0; Jj[q] = 0.0; }
  if (c == 0) { Ji[1] = A[2]; Ji[2] = -A[1]; Jj[1] = B[2]; Jj[2] = -B[1]; }
  if (c == 1) { Ji[0] = -A[2]; Ji[2] = A[0]; Jj[0] = -B[2]; Jj[2] = B[0]; }
  if (c == 2) { Ji[0] = A[1]; Ji[1] = -A[0]; Jj[0] = B[1]; Jj[1] = -B[0]; }
  Ji[3 + c] = 1.0;
  Jj[3 + c] = 1.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) { Ji[q] *= w; Jj[q] *= -w; }
}

// kAsmLanes lanes per block (4 blocks per wavefront): H_ab (and for a == b
// the rhs, the damping and the FP64 block itself, inverted by k_nr_invdiag).
// Lanes take contributions strided by kAsmLanes; the lane sums are combined
// by the fixed DPP tree of row_sum within the block's lanes (one row of 16).
// 16 lanes give each lane several independent gathers in flight and a
// 4-level tree (a wavefront per block took about 153 us per assembly at
// 4048 nodes, these about 65 us).
// Launched at 3 waves per SIMD (168 VGPRs, a few spilled) rather than the
// compiler's 2 (176 VGPRs): more blocks in flight.
constexpr int kAsmLanes = 16;
__device__ __forceinline__ void d_nr_assemble(const int *__restrict__ blk_a, const int *__restrict__ blk_b,
                                              const int *__restrict__ cstart, const uint32_t *__restrict__ codes,
                                              const float2 *__restrict__ cw, int nb, const float *__restrict__ vdata,
                                              const int *__restrict__ nbr, const double *__restrict__ edata,
                                              const NrFrame *Fp, float *__restrict__ H, double *__restrict__ rhs,
                                              double *__restrict__ Hd) {
  const int blk = blockIdx.x * (256 / kAsmLanes) + threadIdx.x / kAsmLanes;
  const int lane = threadIdx.x % kAsmLanes;
  if (blk >= nb) return;  // uniform per block's lanes (the tree stays within them)
  const NrFrame &F = *Fp;
  const int a = blk_a[blk], b = blk_b[blk];
  const bool diag = a == b;
  double h[36], g[6];
#pragma unroll
  for (int q = 0; q < 36; ++q) h[q] = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) g[q] = 0.0;
  const double lam = F.lambda;
  // one contribution: a vertex's data term (weights wa, wb of the block's two
  // nodes at the vertex, from cw) or an ARAP edge's rows
  auto add_one = [&](uint32_t code, float2 w, float4 d0, float4 d1) {
    if (!(code & kArapTag)) {
      if (d1.w == 0.f) return;
      const double wa = (double)w.x, wb = (double)w.y;
      const double gv[6] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y};
      const double s = wa * wb;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int q = 0; q < 6; ++q) h[6 * r + q] += s * gv[r] * gv[q];
      if (diag) {
        const double rr = (double)d1.z;
#pragma unroll
        for (int r = 0; r < 6; ++r) g[r] -= wa * gv[r] * rr;
      }
    } else {
      const int e = (int)((code & ~kArapTag) >> 2), role = code & 3;
      const double *ed = edata + 8 * (size_t)e;
      const int i = e / kNK, j = nbr[e];
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) {
        double Ji[6], Jj[6];
        arap_rows(ed, c3, Ji, Jj);
        const double rc = ed[6] * (ed[c3] - ed[3 + c3]);
        // the block is (min, max): transpose the (i, j) product when i > j
        // (values selected, not pointers, so the rows stay in registers)
        const bool ui = role == 0 || (role == 2 && i <= j), wi = role == 0 || (role == 2 && i > j);
        double U[6], W[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          U[r] = ui ? Ji[r] : Jj[r];
          W[r] = wi ? Ji[r] : Jj[r];
        }
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
          for (int q = 0; q < 6; ++q) h[6 * r + q] += lam * U[r] * W[q];
        if (role < 2)
#pragma unroll
          for (int r = 0; r < 6; ++r) g[r] -= lam * U[r] * rc;
      }
    }
  };
  // contributions c, c + kAsmLanes, ... in order, two at a time so that both
  // vertices' rows are loaded together
  const int ce = cstart[blk + 1];
  const float4 *V4 = reinterpret_cast<const float4 *>(vdata);
  for (int c = cstart[blk] + lane; c < ce; c += 2 * kAsmLanes) {
    const int c1 = c + kAsmLanes;
    const bool has1 = c1 < ce;
    const uint32_t k0 = codes[c], k1 = has1 ? codes[c1] : kArapTag;
    const float2 w0 = cw[c], w1 = has1 ? cw[c1] : make_float2(0.f, 0.f);
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 a0 = z4, a1 = z4, b0 = z4, b1 = z4;
    if (!(k0 & kArapTag)) { const size_t v = k0 >> 4; a0 = V4[2 * v]; a1 = V4[2 * v + 1]; }
    if (!(k1 & kArapTag)) { const size_t v = k1 >> 4; b0 = V4[2 * v]; b1 = V4[2 * v + 1]; }
    add_one(k0, w0, a0, a1);
    if (has1) add_one(k1, w1, b0, b1);
  }
  static_assert(kAsmLanes == 16, "a block's lanes are one DPP row");
  auto lanes_sum = [](double x) { return row_sum(x); };
#pragma unroll
  for (int q = 0; q < 36; ++q) h[q] = lanes_sum(h[q]);
#pragma unroll
  for (int q = 0; q < 6; ++q) g[q] = lanes_sum(g[q]);  // zero off the diagonal
  if (lane != 0) return;
  if (diag)
#pragma unroll
    for (int r = 0; r < 6; ++r) h[7 * r] += F.damping;
  float *Hb = H + 36 * (size_t)blk;
#pragma unroll
  for (int q = 0; q < 36; ++q) Hb[q] = (float)h[q];
  if (!diag) return;
#pragma unroll
  for (int r = 0; r < 6; ++r) rhs[6 * a + r] = g[r];
  double *D = Hd + 36 * (size_t)a;
#pragma unroll
  for (int q = 0; q < 36; ++q) D[q] = h[q];
}
#define DM_ASM_BOUNDS __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
__global__ DM_ASM_BOUNDS void k_nr_assemble(const int *__restrict__ blk_a, const int *__restrict__ blk_b,
                                            const int *__restrict__ cstart, const uint32_t *__restrict__ codes,
                                            const float2 *__restrict__ cw, int nb, const float *__restrict__ vdata,
                                            const int *__restrict__ nbr, const double *__restrict__ edata,
                                            const NrFrame *Fp, float *__restrict__ H, double *__restrict__ rhs,
                                            double *__restrict__ Hd) {
  d_nr_assemble(blk_a, blk_b, cstart, codes, cw, nb, vdata, nbr, edata, Fp, H, rhs, Hd);
}
struct b_nr_assemble {
  const int *blk_a;
  const int *blk_b;
  const int *cstart;
  const uint32_t *codes;
  const float2 *cw;
  int nb;
  const float *vdata;
  const int *nbr;
  const double *edata;
  const NrFrame *Fp;
  float *H;
  double *rhs;
  double *Hd;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_assemble for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ DM_ASM_BOUNDS void kb_nr_assemble(const b_nr_assemble *__restrict__ t) {
  const b_nr_assemble &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_assemble(a.blk_a, a.blk_b, a.cstart, a.codes, a.cw, a.nb, a.vdata, a.nbr, a.edata, a.Fp, a.H, a.rhs, a.Hd);
}

// Block-Jacobi preconditioner: the FP64 inverse of each diagonal block (one
// thread per node; LDL^T once, then the 6 unit columns; a block that is not
// positive definite gets 0), and the PCG's start vectors of the node: x = 0,
// r = rhs, u = M r (row sums in column order).
__device__ __forceinline__ void d_nr_invdiag(const double *__restrict__ Hd, int nn, double *__restrict__ Minv,
                                             const double *__restrict__ rhs, double *__restrict__ x,
                                             double *__restrict__ r, double *__restrict__ u) {
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (a >= nn) return;
  double h[36], L[36], D[6], b[6], uo[6] = {0, 0, 0, 0, 0, 0};
  for (int q = 0; q < 36; ++q) h[q] = Hd[36 * (size_t)a + q];
  for (int q = 0; q < 6; ++q) b[q] = rhs[6 * (size_t)a + q];
  const bool ok = ldlt6_factor(h, L, D) == 0;
  double *M = Minv + 36 * (size_t)a;
  for (int c = 0; c < 6; ++c) {
    double e6[6] = {0, 0, 0, 0, 0, 0}, col[6] = {0, 0, 0, 0, 0, 0};
    e6[c] = 1.0;
    if (ok) ldlt6_solve(L, D, e6, col);
    for (int k = 0; k < 6; ++k) {
      M[6 * k + c] = col[k];
      uo[k] += col[k] * b[c];
    }
  }
  for (int k = 0; k < 6; ++k) {
    x[6 * (size_t)a + k] = 0.0;
    r[6 * (size_t)a + k] = b[k];
    u[6 * (size_t)a + k] = uo[k];
  }
}
__global__ __launch_bounds__(256) void k_nr_invdiag(const double *__restrict__ Hd, int nn, double *__restrict__ Minv,
                                                   const double *__restrict__ rhs, double *__restrict__ x,
                                                   double *__restrict__ r, double *__restrict__ u) {
  d_nr_invdiag(Hd, nn, Minv, rhs, x, r, u);
}
struct b_nr_invdiag {
  const double *Hd;
  int nn;
  double *Minv;
  const double *rhs;
  double *x, *r, *u;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_invdiag for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_invdiag(const b_nr_invdiag *__restrict__ t) {
  const b_nr_invdiag &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_invdiag(a.Hd, a.nn, a.Minv, a.rhs, a.x, a.r, a.u);
}

// ---------------------------------------------------------------- PCG
// Block-Jacobi PCG in the pipelined form of Ghysels and Vanroose ("Hiding
// global synchronization latency in the preconditioned Conjugate Gradient
// algorithm", Parallel Computing 40 (2014), algorithm 3).  Besides x, r,
// u = M r and w = A u it carries p, s = A p, q = M s and t = A q by
// recurrences, and m = M w, so that an iteration's SpMV (n = A m) needs
// nothing of the iteration's dot products.  One kernel per iteration adds
// the previous kernel's partials of r.u and w.u (the same fixed order in
// every workgroup), computes n = A m for its nodes, updates the eight
// vectors of its nodes' rows and writes m and the partials for the next
// iteration.  Mathematically the standard PCG iteration (the same x_k), with
// one kernel boundary per iteration.
constexpr int kSpmvNodes = 16;  // wavefronts (nodes) per workgroup

struct PcgState {
  double *x, *r, *u, *w;  // 6 N each: iterate, residual, u = M r, w = A u
  double *p, *s, *q, *t;  // direction p, s = A p, q = M s, t = A q
  double *m0, *m1;        // m = M w: iteration k reads m_(k & 1), writes m_((k + 1) & 1)
  double *part;           // partials, two per workgroup: two halves of (r.u [2 nwg], w.u [2 nwg]): the
                          // start kernel writes half 0, iteration k reads half k & 1 and writes
                          // half (k + 1) & 1, so no workgroup overwrites a partial another may still read
  double *rz;             // rz[k] = r_k . u_k (k = 0 .. max)
  double *al;             // al[k]: step length of iteration k
  int *flag;              // [0] converged, [1] iterations run
  int nwg;                // workgroups of the start and iteration kernels (kSpmvNodes nodes each)
};

// Row `lane` (< 6) of (A v) at the wavefront's node, 0 in the other lanes.
// Lane j takes the node's block entries j, j + 64, ... whole (a 6x6 block,
// transposed for the lower triangle, times the column node's 6 values), so
// a node's entries are one round of loads; the six row sums are then added
// over the lanes by the fixed tree of wave_sum.
__device__ __forceinline__ double spmv_node(const float *__restrict__ H, const int *__restrict__ row_start,
                                            const uint2 *__restrict__ row_ec,
                                            const double *__restrict__ v, int node, int nn, int lane) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (node < nn) {
    const int k1 = row_start[node + 1];  // node is uniform in the wavefront: scalar loads
    for (int k = row_start[node] + lane; k < k1; k += 64) {
      const uint2 ec = row_ec[k];
      const int blk = (int)(ec.x & 0x7fffffffu);
      const bool tr = ec.x >> 31;
      const float *B = H + 36 * (size_t)blk;
      const double2 *V2 = reinterpret_cast<const double2 *>(v + 6 * (size_t)ec.y);
      double vc[6];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double2 d = V2[q];
        vc[2 * q] = d.x; vc[2 * q + 1] = d.y;
      }
      const float4 *B4 = reinterpret_cast<const float4 *>(B);
      float b[36];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        const float4 f = B4[q];
        b[4 * q] = f.x; b[4 * q + 1] = f.y; b[4 * q + 2] = f.z; b[4 * q + 3] = f.w;
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double sv = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) sv += (double)(tr ? b[6 * c + r] : b[6 * r + c]) * vc[c];
        acc[r] += sv;
      }
    }
  }
  double out = 0.0;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const double t = wave_sum(acc[r]);
    if (lane == r) out = t;
  }
  return node < nn ? out : 0.0;
}

// The partials of r.u and w.u into half `half` of the partial array, one
// per update wavefront (0 and 1) of each workgroup: the wavefront's lane
// values (0 in lanes without a row) by the fixed tree of wave_sum.  No
// barrier: each wavefront writes its own entries.
__device__ __forceinline__ void pcg_partials(const PcgState &S, double ru, double wu, int half) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (wv < 2) {
    ru = wave_sum(ru);
    wu = wave_sum(wu);
    const size_t np = 2 * (size_t)S.nwg, k = 2 * (size_t)blockIdx.x + wv;
    if (lane == 0) S.part[(size_t)half * 2 * np + k] = ru;
    if (lane == 1) S.part[(size_t)half * 2 * np + np + k] = wu;
  }
}

// Rows of the workgroup (6 kSpmvNodes from row 6 kSpmvNodes blockIdx.x):
// wavefront 0 takes rows 0 .. 59 (nodes 0 .. 9) in lanes 0 .. 59, wavefront
// 1 rows 60 .. 95 (nodes 10 .. 15) in lanes 0 .. 35, so that each node's six
// rows are in one wavefront; -1 for a lane without a row.
constexpr int kPcgRows = 6 * kSpmvNodes;
static_assert(kPcgRows == 96, "two update wavefronts of 60 and 36 rows");
__device__ __forceinline__ int pcg_row() {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  return wv == 0 ? (lane < 60 ? lane : -1) : wv == 1 ? (lane < 36 ? 60 + lane : -1) : -1;
}

// Start: w = A u, m = M w, partials of r_0.u_0 and w_0.u_0 (x = 0, r = rhs and
// u = M r are written by the preconditioner kernel).
__device__ __forceinline__ void d_pcg_start(PcgState S, const float *__restrict__ H,
                                            const int *__restrict__ row_start, const uint2 *__restrict__ row_ec,
                                            const double *__restrict__ Minv, int nn) {
  __shared__ double nl[kPcgRows];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = pcg_row();
  const size_t e = (size_t)kPcgRows * blockIdx.x + i;
  const bool own = i >= 0 && e < 6 * (size_t)nn;
  double mrow[6] = {0, 0, 0, 0, 0, 0}, rv = 0.0, uv = 0.0;
  if (own) {
    rv = S.r[e];
    uv = S.u[e];
#pragma unroll
    for (int c = 0; c < 6; ++c) mrow[c] = Minv[6 * e + c];
  }
  const int node = blockIdx.x * kSpmvNodes + __builtin_amdgcn_readfirstlane(wv);
  const double wn = spmv_node(H, row_start, row_ec, S.u, node, nn, lane);
  if (lane < 6) nl[6 * wv + lane] = wn;
  __syncthreads();
  double w = 0.0;
  if (own) { w = nl[i]; S.w[e] = w; }
  double mn = 0.0;
  if (own) {
    const int r0 = i - (i % 6);
#pragma unroll
    for (int c = 0; c < 6; ++c) mn += mrow[c] * nl[r0 + c];
    S.m0[e] = mn;
  }
  pcg_partials(S, own ? rv * uv : 0.0, own ? w * uv : 0.0, 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) { S.flag[0] = 0; S.flag[1] = 0; }
}
__global__ __launch_bounds__(64 * kSpmvNodes) void k_pcg_start(PcgState S, const float *__restrict__ H,
                                                              const int *__restrict__ row_start,
                                                              const uint2 *__restrict__ row_ec,
                                                              const double *__restrict__ Minv, int nn) {
  d_pcg_start(S, H, row_start, row_ec, Minv, nn);
}
struct b_pcg_start {
  PcgState S;
  const float *H;
  const int *row_start;
  const uint2 *row_ec;
  const double *Minv;
  int nn;
  unsigned grid;  // workgroups of this graph's launch
};
// k_pcg_start for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(64 * kSpmvNodes) void kb_pcg_start(const b_pcg_start *__restrict__ t) {
  const b_pcg_start &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_pcg_start(a.S, a.H, a.row_start, a.row_ec, a.Minv, a.nn);
}

// Iteration k: gamma = r.u and delta = w.u from the partials (wavefronts 2
// and 3 side by side, lane l adding partials l, l + 64, ... and then the xor
// tree: the same order in every workgroup); stop when gamma <= tol^2 gamma_0;
// beta = gamma / gamma_{k-1}, alpha = gamma / (delta - beta gamma /
// alpha_{k-1}) (k = 0: beta = 0, alpha = gamma / delta); n = A m;
// t = n + beta t; q = m + beta q; s = w + beta s; p = u + beta p;
// x += alpha p; r -= alpha s; u -= alpha q; w -= alpha t; m = M w; partials
// of r.u and w.u.  The partials, the rows' operands and the SpMV's loads
// are all issued before the first wait.
__device__ __forceinline__ void d_pcg_iter(PcgState S, const float *__restrict__ H,
                                           const int *__restrict__ row_start, const uint2 *__restrict__ row_ec,
                                           const double *__restrict__ Minv, int nn, int it, const NrFrame *Fp) {
  __shared__ double nl[kPcgRows], wl[kPcgRows], l2[2], sc[5];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // wavefronts 2 and 3: lane l adds partials l, l + 64, ... in order; the
  // first 8 are loaded here and added after the SpMV's loads are issued.
  // Lanes 0..4 of wavefront 2 also read the tolerance, r_0.u_0, the previous
  // iteration's gamma and alpha and the converged flag, with vector loads
  // (per-lane addresses) that overlap the SpMV rather than scalar loads that
  // the kernel's first scalar wait would include.
  const bool pw = wv == 2 || wv == 3;
  const int np = 2 * S.nwg;  // partials per quantity (two per workgroup)
  const double *part = S.part + (size_t)(it & 1) * 2 * np + (wv == 3 ? np : 0);
  double pv8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, scv = 0.0;
  if (pw) {
#pragma unroll
    for (int j = 0; j < 8; ++j) pv8[j] = part[min(64 * j + lane, np - 1)];  // in bounds; lanes past np skip it below
    const int km = it > 0 ? it - 1 : 0;
    const double *q = lane == 0 ? &Fp->pcg_tol
                      : lane == 1 ? S.rz
                      : lane == 2 ? S.rz + km
                      : lane == 3 ? S.al + km
                                  : reinterpret_cast<const double *>(S.flag);  // flag[0] in the low word
    if (wv == 2 && lane < 5) scv = *q;
  }
  const double *m_in = (it & 1) ? S.m1 : S.m0;
  double *m_out = (it & 1) ? S.m0 : S.m1;
  const int i = pcg_row();
  const size_t e = (size_t)kPcgRows * blockIdx.x + i;
  const bool own = i >= 0 && e < 6 * (size_t)nn;
  double mrow[6] = {0, 0, 0, 0, 0, 0}, tv = 0.0, qv = 0.0, sv = 0.0, pv = 0.0, xv = 0.0, rv = 0.0, uv = 0.0,
         wo = 0.0, mv = 0.0;
  if (own) {
    if (it > 0) { tv = S.t[e]; qv = S.q[e]; sv = S.s[e]; pv = S.p[e]; }
    xv = S.x[e]; rv = S.r[e]; uv = S.u[e]; wo = S.w[e]; mv = m_in[e];
#pragma unroll
    for (int c = 0; c < 6; ++c) mrow[c] = Minv[6 * e + c];
  }
  const int node = blockIdx.x * kSpmvNodes + __builtin_amdgcn_readfirstlane(wv);
  // (computed also after convergence, so that the flag's load is not on the SpMV's path)
  const double n = spmv_node(H, row_start, row_ec, m_in, node, nn, lane);
  if (lane < 6) nl[6 * wv + lane] = n;
  if (pw) {
    double pacc = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (64 * j + lane < np) pacc += pv8[j];
    for (int k = 512 + lane; k < np; k += 64) pacc += part[k];
    pacc = wave_sum(pacc);
    if (lane == 0) l2[wv - 2] = pacc;
    if (wv == 2 && lane < 5) sc[lane] = scv;
  }
  __syncthreads();
  if (__double2loint(sc[4]) != 0) return;  // converged at an earlier iteration
  const double tol = sc[0], rz0 = sc[1], rzp = sc[2], alp = sc[3];
  const double gamma = l2[0], delta = l2[1];
  const bool stop = it == 0 ? !(gamma > 0.0) : (!(gamma > tol * tol * rz0) || !(gamma > 0.0));
  double beta = 0.0, alpha = gamma / delta;
  if (it > 0) {
    beta = gamma / rzp;
    alpha = gamma / (delta - beta * gamma / alp);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.rz[it] = gamma;
    S.al[it] = alpha;
    if (stop) S.flag[0] = 1;
    else S.flag[1] = it + 1;
  }
  if (stop) return;
  double wn = 0.0;
  if (own) {
    const double nv = nl[i];
    tv = it == 0 ? nv : nv + beta * tv;
    qv = it == 0 ? mv : mv + beta * qv;
    sv = it == 0 ? wo : wo + beta * sv;
    pv = it == 0 ? uv : uv + beta * pv;
    xv = xv + alpha * pv;
    rv = rv - alpha * sv;
    uv = uv - alpha * qv;
    wn = wo - alpha * tv;
    S.t[e] = tv; S.q[e] = qv; S.s[e] = sv; S.p[e] = pv;
    S.x[e] = xv; S.r[e] = rv; S.u[e] = uv; S.w[e] = wn;
    wl[i] = wn;
  }
  // m = M w: the node's six w are in this wavefront (written to LDS just
  // above by the same wavefront, whose LDS operations complete in order)
  if (own) {
    const int r0 = i - (i % 6);
    double mn = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) mn += mrow[c] * wl[r0 + c];
    m_out[e] = mn;
  }
  pcg_partials(S, own ? rv * uv : 0.0, own ? wn * uv : 0.0, (it + 1) & 1);
}
__global__ __launch_bounds__(64 * kSpmvNodes) void k_pcg_iter(PcgState S, const float *__restrict__ H,
                                                             const int *__restrict__ row_start,
                                                             const uint2 *__restrict__ row_ec,
                                                             const double *__restrict__ Minv, int nn, int it,
                                                             const NrFrame *Fp) {
  d_pcg_iter(S, H, row_start, row_ec, Minv, nn, it, Fp);
}
struct b_pcg_iter {
  PcgState S;
  const float *H;
  const int *row_start;
  const uint2 *row_ec;
  const double *Minv;
  int nn;
  const NrFrame *Fp;
  unsigned grid;  // workgroups of this graph's launch
};
// k_pcg_iter for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(64 * kSpmvNodes) void kb_pcg_iter(const b_pcg_iter *__restrict__ t, int it) {
  const b_pcg_iter &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_pcg_iter(a.S, a.H, a.row_start, a.row_ec, a.Minv, a.nn, it, a.Fp);
}

// T_i <- exp(x_i) T_i, then the node's dual quaternion
__device__ __forceinline__ void d_nr_update(const double *__restrict__ x, double *__restrict__ T, int nn,
                                            dq8 *__restrict__ dq) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  double E[16], Tn[16];
  se3_exp(x + 6 * (size_t)i, E);
  mat4_mul(E, T + 16 * (size_t)i, Tn);
  for (int k = 0; k < 16; ++k) T[16 * (size_t)i + k] = Tn[k];
  dq[i] = dq_of_T(Tn);
}
__global__ __launch_bounds__(256) void k_nr_update(const double *__restrict__ x, double *__restrict__ T, int nn,
                                                  dq8 *__restrict__ dq) { d_nr_update(x, T, nn, dq); }
struct b_nr_update {
  const double *x;
  double *T;
  int nn;
  dq8 *dq;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_update for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_update(const b_nr_update *__restrict__ t) {
  const b_nr_update &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_update(a.x, a.T, a.nn, a.dq);
}

// energies at the last linearisation: [0] residual count, [1] data, [2] ARAP.
// kStatsWg workgroups write fixed-order partials (grid-stride over vertices
// and edges); k_nr_stats_sum adds them in order.
constexpr int kStatsWg = 128;
__device__ __forceinline__ void d_nr_stats(const float *__restrict__ vdata, int64_t nv,
                                           const double *__restrict__ edata, int ne, const NrFrame *Fp,
                                           double *__restrict__ part) {
  const double lam = Fp->lambda;
  __shared__ double lds[3][256 / 64];
  double c = 0.0, ed = 0.0, er = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += stride) {
    const float *d = vdata + 8 * v;
    if (d[7] != 0.f) { c += 1.0; ed += (double)d[6] * (double)d[6]; }
  }
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < ne; e += stride) {
    const double *q = edata + 8 * (size_t)e;
    for (int k = 0; k < 3; ++k) {
      const double rc = q[6] * (q[k] - q[3 + k]);
      er += lam * rc * rc;
    }
  }
  c = wave_sum(c); ed = wave_sum(ed); er = wave_sum(er);
  if ((threadIdx.x & 63) == 0) { lds[0][threadIdx.x >> 6] = c; lds[1][threadIdx.x >> 6] = ed; lds[2][threadIdx.x >> 6] = er; }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int k = 0; k < 256 / 64; ++k) s += lds[threadIdx.x][k];
    part[3 * blockIdx.x + threadIdx.x] = s;
  }
}
__global__ __launch_bounds__(256) void k_nr_stats(const float *__restrict__ vdata, int64_t nv,
                                                 const double *__restrict__ edata, int ne, const NrFrame *Fp,
                                                 double *__restrict__ part) { d_nr_stats(vdata, nv, edata, ne, Fp, part); }
struct b_nr_stats {
  const float *vdata;
  int64_t nv;
  const double *edata;
  int ne;
  const NrFrame *Fp;
  double *part;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_stats for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ __launch_bounds__(256) void kb_nr_stats(const b_nr_stats *__restrict__ t) {
  const b_nr_stats &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_stats(a.vdata, a.nv, a.edata, a.ne, a.Fp, a.part);
}

__device__ __forceinline__ void d_nr_stats_sum(const double *__restrict__ part, double *__restrict__ out) {
  if (threadIdx.x >= 3) return;
  double s = 0.0;
  for (int b = 0; b < kStatsWg; ++b) s += part[3 * b + threadIdx.x];
  out[threadIdx.x] = s;
}
__global__ void k_nr_stats_sum(const double *__restrict__ part, double *__restrict__ out) { d_nr_stats_sum(part, out); }
struct b_nr_stats_sum {
  const double *part;
  double *out;
  unsigned grid;  // workgroups of this graph's launch
};
// k_nr_stats_sum for a batch of graphs: graph blockIdx.y, its arguments in t[blockIdx.y]
__global__ void kb_nr_stats_sum(const b_nr_stats_sum *__restrict__ t) {
  const b_nr_stats_sum &a = t[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  d_nr_stats_sum(a.part, a.out);
}

__global__ void k_set_identity(double *T, int nn) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nn) return;
  for (int k = 0; k < 16; ++k) T[16 * (size_t)i + k] = (k % 5 == 0) ? 1.0 : 0.0;
}

}  // namespace dm

// ---------------------------------------------------------------- host object
using namespace dm;

struct dm_graph {
  hipStream_t stream = nullptr;
  int64_t nv = 0;
  int nn = 0, nb = 0;
  long long ncontrib = 0, nrow = 0;
  dm_graph_props props{};
  std::vector<void *> allocs;
  // canonical mesh and graph
  float *verts = nullptr, *normals = nullptr;
  int *node_vid = nullptr, *node_of_vertex = nullptr, *nbr = nullptr, *vnn = nullptr;
  float *nbr_w = nullptr, *vnn_w = nullptr, *vnn_wn = nullptr;
  double *T = nullptr;
  dq8 *dq = nullptr;
  // block pattern
  int *blk_a = nullptr, *blk_b = nullptr, *cstart = nullptr, *row_start = nullptr;
  uint32_t *codes = nullptr, *row_ent = nullptr;
  float2 *cw = nullptr;  // per contribution: the weights of the block's two nodes at the vertex
  uint2 *row_ec = nullptr;  // row entries for the SpMV: (row_ent, column node)
  // per-iteration buffers
  float *pv = nullptr, *pn = nullptr, *vdata = nullptr, *H = nullptr;
  double *edata = nullptr, *rhs = nullptr, *Minv = nullptr, *Hd = nullptr, *stats = nullptr, *stats_part = nullptr;
  PcgState pcg{};
  int pcg_cap = 0;
  // frame staging
  float *depth = nullptr, *sdepth = nullptr;
  float4 *tv = nullptr, *tn = nullptr;
  size_t pixels_cap = 0;
  float *mesh_v = nullptr, *mesh_n = nullptr;  // mesh target (nv x 3 each), allocated on first use
  NrFrame *d_frame = nullptr, *h_frame = nullptr;  // device record and its pinned host copy
  double *h_stats = nullptr;  // pinned: the last solve's 3 energies and 2 PCG flags (as doubles' storage)
  // the captured solve and what it was captured for
  hipGraphExec_t exec = nullptr;
  int key_gn = -1, key_pcg = -1, key_rows = -1, key_cols = -1, key_mode = -1;
  int buf_gen = 0;  // counts frame-buffer reallocations (batches capture again when it moves)

  template <class T> int dalloc(T **p, size_t n) {
    void *q = nullptr;
    DM_HIP(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 16)));
    allocs.push_back(q);
    *p = (T *)q;
    return DM_OK;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    if (h_frame) (void)hipHostFree(h_frame);
    h_frame = nullptr;
    if (h_stats) (void)hipHostFree(h_stats);
    h_stats = nullptr;
    for (void *p : allocs) (void)hipFree(p);
    allocs.clear();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};


static unsigned grid_of(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static int sort_pairs(const uint32_t *ki, uint32_t *ko, const uint32_t *vi, uint32_t *vo, long long n, int bits,
                      hipStream_t st) {
  size_t tb = 0;
  DM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ki, ko, vi, vo, (int)n, 0, bits, st));
  DevBuf tmp;
  DM_TRY(dev_alloc(tmp, tb));
  DM_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, ki, ko, vi, vo, (int)n, 0, bits, st));
  DM_HIP(hipStreamSynchronize(st));  // tmp is released on return
  return DM_OK;
}

static int graph_finish(dm_graph *g);

static int graph_build(dm_graph *g, const float *h_verts, const float *h_normals) {
  hipStream_t st = g->stream;
  const int64_t nv = g->nv;
  DM_TRY(g->dalloc(&g->verts, (size_t)nv * 3));
  DM_TRY(g->dalloc(&g->normals, (size_t)nv * 3));
  DM_HIP(hipMemcpyAsync(g->verts, h_verts, (size_t)nv * 12, hipMemcpyHostToDevice, st));
  DM_HIP(hipMemcpyAsync(g->normals, h_normals, (size_t)nv * 12, hipMemcpyHostToDevice, st));
  // node sampling
  uint32_t tsize = 1;
  while (tsize < 2 * (uint64_t)nv) tsize <<= 1;
  DevBuf hkeys, hvals, isnode, rank, wsum;
  DM_TRY(dev_alloc(hkeys, (size_t)tsize * 8));
  DM_TRY(dev_alloc(hvals, (size_t)tsize * 8));
  DM_TRY(dev_alloc(isnode, (size_t)nv * 4));
  DM_TRY(dev_alloc(rank, (size_t)nv * 8));
  DM_HIP(hipMemsetAsync(hkeys.p, 0, (size_t)tsize * 8, st));
  DM_HIP(hipMemsetAsync(hvals.p, 0xff, (size_t)tsize * 8, st));
  const float r = g->props.node_radius;
  hipLaunchKernelGGL(k_nr_cells, dim3(grid_of(nv)), dim3(256), 0, st, g->verts, nv, r, hkeys.as<uint64_t>(),
                     hvals.as<unsigned long long>(), tsize - 1);
  hipLaunchKernelGGL(k_nr_mark, dim3(grid_of(nv)), dim3(256), 0, st, g->verts, nv, r, hkeys.as<uint64_t>(),
                     hvals.as<unsigned long long>(), tsize - 1, isnode.as<int>());
  DM_HIP(hipGetLastError());
  long long *d_nn = nullptr, nn = 0;
  DM_TRY(scan_ints(isnode.as<int>(), (int)nv, rank.as<long long>(), wsum, st, &d_nn));
  DM_HIP(hipMemcpyAsync(&nn, d_nn, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  if (nn < kNK + 1) { set_error("deformation graph: fewer than 5 nodes"); return DM_E_INVALID; }
  if (nn > g->props.max_nodes) {
    set_error("deformation graph: " + std::to_string(nn) + " nodes exceed max_nodes");
    return DM_E_CAPACITY;
  }
  g->nn = (int)nn;
  DM_TRY(g->dalloc(&g->node_vid, nn));
  DM_TRY(g->dalloc(&g->node_of_vertex, nv));
  hipLaunchKernelGGL(k_nr_nodes, dim3(grid_of(nv)), dim3(256), 0, st, isnode.as<int>(), rank.as<long long>(), nv,
                     g->node_vid, g->node_of_vertex);
  return graph_finish(g);
}

// defGraph::init with the nodes given (defGraph.cu:15-69, sampleDownMeshHost
// :673-702): node k is vertex ids[k]
static int graph_build_nodes(dm_graph *g, const float *h_verts, const float *h_normals, const int32_t *ids, int nn) {
  hipStream_t st = g->stream;
  const int64_t nv = g->nv;
  std::vector<int32_t> nov((size_t)nv, -1);
  for (int k = 0; k < nn; ++k) {
    if (ids[k] < 0 || ids[k] >= nv || nov[ids[k]] >= 0) {
      set_error("deformation graph: node vertex id out of range or repeated");
      return DM_E_INVALID;
    }
    nov[ids[k]] = k;
  }
  DM_TRY(g->dalloc(&g->verts, (size_t)nv * 3));
  DM_TRY(g->dalloc(&g->normals, (size_t)nv * 3));
  DM_HIP(hipMemcpyAsync(g->verts, h_verts, (size_t)nv * 12, hipMemcpyHostToDevice, st));
  DM_HIP(hipMemcpyAsync(g->normals, h_normals, (size_t)nv * 12, hipMemcpyHostToDevice, st));
  g->nn = nn;
  DM_TRY(g->dalloc(&g->node_vid, nn));
  DM_TRY(g->dalloc(&g->node_of_vertex, nv));
  DM_HIP(hipMemcpyAsync(g->node_vid, ids, (size_t)nn * 4, hipMemcpyHostToDevice, st));
  DM_HIP(hipMemcpyAsync(g->node_of_vertex, nov.data(), (size_t)nv * 4, hipMemcpyHostToDevice, st));
  DM_HIP(hipStreamSynchronize(st));  // nov is released on return
  return graph_finish(g);
}

// defGraphUpdateNodes (defGraph.cu:483-545) and the vertex neighbours
// (updateActiveNeighbourNodes, :243-302): the K nearest nodes of every node
// and of every vertex, with weights exp(-d^2/dgw^2)
static int graph_links(dm_graph *g) {
  hipStream_t st = g->stream;
  const int nn = g->nn;
  const int64_t nv = g->nv;
  hipLaunchKernelGGL(k_nr_knn, dim3(grid_of(nn)), dim3(256), 0, st, g->verts, g->node_vid, nn, g->node_vid,
                     (int64_t)nn, g->props.dgw, g->nbr, g->nbr_w, (float *)nullptr);
  hipLaunchKernelGGL(k_nr_knn, dim3(grid_of(nv)), dim3(256), 0, st, g->verts, g->node_vid, nn,
                     (const int *)nullptr, nv, g->props.dgw, g->vnn, g->vnn_w, g->vnn_wn);
  DM_HIP(hipGetLastError());
  return DM_OK;
}

// neighbours, the block pattern of the normal equations, node state and the
// per-iteration buffers of a graph whose nodes are set
static int graph_finish(dm_graph *g) {
  hipStream_t st = g->stream;
  const int64_t nv = g->nv;
  const long long nn = g->nn;
  DM_TRY(g->dalloc(&g->nbr, nn * kNK));
  DM_TRY(g->dalloc(&g->nbr_w, nn * kNK));
  DM_TRY(g->dalloc(&g->vnn, (size_t)nv * kNK));
  DM_TRY(g->dalloc(&g->vnn_w, (size_t)nv * kNK));
  DM_TRY(g->dalloc(&g->vnn_wn, (size_t)nv * kNK));
  DM_TRY(graph_links(g));
  // block pattern: contributions sorted by block key (stable: vertex order, then edges)
  DevBuf ccnt, coff, wsum2;
  DM_TRY(dev_alloc(ccnt, (size_t)nv * 4));
  DM_TRY(dev_alloc(coff, (size_t)nv * 8));
  hipLaunchKernelGGL(k_nr_contrib_count, dim3(grid_of(nv)), dim3(256), 0, st, g->node_of_vertex, nv, ccnt.as<int>());
  long long *d_nd = nullptr, ndata = 0;
  DM_TRY(scan_ints(ccnt.as<int>(), (int)nv, coff.as<long long>(), wsum2, st, &d_nd));
  DM_HIP(hipMemcpyAsync(&ndata, d_nd, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  const long long nc = ndata + 3LL * nn * kNK;
  if (nc >= (1LL << 31)) { set_error("deformation graph: too many block contributions"); return DM_E_CAPACITY; }
  g->ncontrib = nc;
  DevBuf k0, c0, k1, flag, frank, wsum3;
  DM_TRY(dev_alloc(k0, nc * 4));
  DM_TRY(dev_alloc(c0, nc * 4));
  DM_TRY(dev_alloc(k1, nc * 4));
  DM_TRY(g->dalloc(&g->codes, nc));
  hipLaunchKernelGGL(k_nr_contrib_emit, dim3(grid_of(nv + nn * kNK)), dim3(256), 0, st, g->node_of_vertex, g->vnn,
                     coff.as<long long>(), nv, g->nbr, (int)nn, ndata, k0.as<uint32_t>(), c0.as<uint32_t>());
  DM_HIP(hipGetLastError());
  DM_TRY(sort_pairs(k0.as<uint32_t>(), k1.as<uint32_t>(), c0.as<uint32_t>(), g->codes, nc, 2 * kNodeBits, st));
  DM_TRY(g->dalloc(&g->cw, nc));
  hipLaunchKernelGGL(k_nr_contrib_weights, dim3(grid_of(nc)), dim3(256), 0, st, g->codes, nc, g->node_of_vertex,
                     g->vnn_wn, g->cw);
  DM_TRY(dev_alloc(flag, nc * 4));
  DM_TRY(dev_alloc(frank, nc * 8));
  hipLaunchKernelGGL(k_flag_heads, dim3(grid_of(nc)), dim3(256), 0, st, k1.as<uint32_t>(), nc, 0u, flag.as<int>());
  long long *d_nb = nullptr, nb = 0;
  DM_TRY(scan_ints(flag.as<int>(), (int)nc, frank.as<long long>(), wsum3, st, &d_nb));
  DM_HIP(hipMemcpyAsync(&nb, d_nb, 8, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  g->nb = (int)nb;
  DM_TRY(g->dalloc(&g->blk_a, nb));
  DM_TRY(g->dalloc(&g->blk_b, nb));
  DM_TRY(g->dalloc(&g->cstart, nb + 1));
  hipLaunchKernelGGL(k_nr_blocks, dim3(grid_of(nc)), dim3(256), 0, st, k1.as<uint32_t>(), flag.as<int>(),
                     frank.as<long long>(), nc, g->blk_a, g->blk_b, g->cstart);
  const int ncs = (int)nc;
  DM_HIP(hipMemcpyAsync(g->cstart + nb, &ncs, 4, hipMemcpyHostToDevice, st));
  // row index for the symmetric SpMV
  std::vector<int> ha(nb), hb(nb);
  DM_HIP(hipMemcpyAsync(ha.data(), g->blk_a, nb * 4, hipMemcpyDeviceToHost, st));
  DM_HIP(hipMemcpyAsync(hb.data(), g->blk_b, nb * 4, hipMemcpyDeviceToHost, st));
  DM_HIP(hipStreamSynchronize(st));
  long long nrow = 0;
  std::vector<long long> roff(nb);
  for (long long k = 0; k < nb; ++k) { roff[k] = nrow; nrow += ha[k] == hb[k] ? 1 : 2; }
  g->nrow = nrow;
  DevBuf droff, rk0, rk1, rv0;
  DM_TRY(dev_alloc(droff, nb * 8));
  DM_TRY(dev_alloc(rk0, nrow * 4));
  DM_TRY(dev_alloc(rk1, nrow * 4));
  DM_TRY(dev_alloc(rv0, nrow * 4));
  DM_TRY(g->dalloc(&g->row_ent, nrow));
  DM_TRY(g->dalloc(&g->row_ec, nrow));
  DM_TRY(g->dalloc(&g->row_start, nn + 1));
  DM_HIP(hipMemcpyAsync(droff.p, roff.data(), nb * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_nr_row_emit, dim3(grid_of(nb)), dim3(256), 0, st, g->blk_a, g->blk_b, (int)nb,
                     droff.as<long long>(), rk0.as<uint32_t>(), rv0.as<uint32_t>());
  DM_TRY(sort_pairs(rk0.as<uint32_t>(), rk1.as<uint32_t>(), rv0.as<uint32_t>(), g->row_ent, nrow, 2 * kNodeBits, st));
  hipLaunchKernelGGL(k_nr_row_start, dim3(grid_of(nrow + 1)), dim3(256), 0, st, rk1.as<uint32_t>(), nrow,
                     g->row_start, (int)nn);
  hipLaunchKernelGGL(k_nr_row_ec, dim3(grid_of(nrow)), dim3(256), 0, st, rk1.as<uint32_t>(), g->row_ent, nrow,
                     g->row_ec);
  DM_HIP(hipGetLastError());
  // state and per-iteration buffers
  DM_TRY(g->dalloc(&g->T, nn * 16));
  DM_TRY(g->dalloc(&g->dq, nn));
  hipLaunchKernelGGL(k_set_identity, dim3(grid_of(nn)), dim3(256), 0, st, g->T, (int)nn);
  hipLaunchKernelGGL(k_nr_node_dq, dim3(grid_of(nn)), dim3(256), 0, st, g->T, (int)nn, g->dq);
  DM_TRY(g->dalloc(&g->pv, (size_t)nv * 3));
  DM_TRY(g->dalloc(&g->pn, (size_t)nv * 3));
  DM_TRY(g->dalloc(&g->vdata, (size_t)nv * 8));
  DM_TRY(g->dalloc(&g->edata, (size_t)nn * kNK * 8));
  DM_TRY(g->dalloc(&g->H, (size_t)nb * 36));
  DM_TRY(g->dalloc(&g->rhs, (size_t)nn * 6));
  DM_TRY(g->dalloc(&g->Minv, (size_t)nn * 36));
  DM_TRY(g->dalloc(&g->Hd, (size_t)nn * 36));
  DM_TRY(g->dalloc(&g->stats, 4));
  DM_TRY(g->dalloc(&g->stats_part, 3 * kStatsWg));
  PcgState &S = g->pcg;
  S.nwg = (int)grid_of(nn, kSpmvNodes);
  for (double **v : {&S.x, &S.r, &S.u, &S.w, &S.p, &S.s, &S.q, &S.t, &S.m0, &S.m1}) DM_TRY(g->dalloc(v, nn * 6));
  DM_TRY(g->dalloc(&S.part, 8 * (size_t)S.nwg));
  DM_TRY(g->dalloc(&S.flag, 2));
  DM_TRY(g->dalloc(&g->d_frame, 1));
  DM_HIP(hipHostMalloc((void **)&g->h_frame, sizeof(NrFrame), hipHostMallocDefault));
  DM_HIP(hipStreamSynchronize(st));
  return DM_OK;
}

// The linearisation terms' arguments of graph g (vertex part, then edges).
static NrTerms terms_of(const dm_graph *g) {
  NrTerms a{};
  a.verts = g->verts; a.normals = g->normals;
  a.node_of_vertex = g->node_of_vertex; a.vnn = g->vnn; a.vnn_w = g->vnn_w; a.dq = g->dq;
  a.nv = g->nv;
  a.tv = g->tv; a.tn = g->tn; a.depth = g->depth; a.sdepth = g->sdepth;
  a.mv = g->mesh_v; a.mn = g->mesh_n;
  a.Fp = g->d_frame;
  a.vdata = g->vdata;
  a.T = g->T; a.node_vid = g->node_vid; a.nbr = g->nbr; a.nbr_w = g->nbr_w; a.nn = g->nn;
  a.edata = g->edata;
  a.vwg = grid_of(g->nv);
  a.grid = a.vwg + grid_of((long long)g->nn * kNK);
  return a;
}

// The kernels of one solve call: target maps, then gn Gauss-Newton
// iterations of (blend, data, ARAP, assembly, PCG, update).  Everything
// that changes between calls (pose, gates, depth) is read from device
// memory, so the sequence is captured once and replayed.
// mode 0: depth target (optionally gated by a source depth); 1: mesh target
static void nr_enqueue(dm_graph *g, int rows, int cols, int gn, int pcg_its, int mode) {
  hipStream_t st = g->stream;
  const size_t P = (size_t)rows * cols;
  const int nn = g->nn, ne = nn * kNK;
  const PcgState S = g->pcg;
  const unsigned pg = (unsigned)S.nwg;
  const NrFrame *F = g->d_frame;
  const NrTerms terms = terms_of(g);
  if (mode == 0) {
    hipLaunchKernelGGL(k_nr_target_vert, dim3(grid_of(P)), dim3(256), 0, st, g->depth, F, g->tv);
    hipLaunchKernelGGL(k_nr_target_norm, dim3(grid_of(P)), dim3(256), 0, st, g->tv, F, g->tn);
  }
  for (int it = 0; it < gn; ++it) {
    if (mode == 0) hipLaunchKernelGGL(k_nr_terms<false>, dim3(terms.grid), dim3(256), 0, st, terms);
    else hipLaunchKernelGGL(k_nr_terms<true>, dim3(terms.grid), dim3(256), 0, st, terms);
    hipLaunchKernelGGL(k_nr_assemble, dim3(grid_of(g->nb, 256 / kAsmLanes)), dim3(256), 0, st, g->blk_a, g->blk_b,
                       g->cstart, g->codes, g->cw, g->nb, g->vdata, g->nbr, g->edata, F, g->H, g->rhs, g->Hd);
    hipLaunchKernelGGL(k_nr_invdiag, dim3(grid_of(nn)), dim3(256), 0, st, g->Hd, nn, g->Minv, g->rhs, S.x, S.r,
                       S.u);
    if (it + 1 == gn) {
      hipLaunchKernelGGL(k_nr_stats, dim3(kStatsWg), dim3(256), 0, st, g->vdata, g->nv, g->edata, ne, F,
                         g->stats_part);
      hipLaunchKernelGGL(k_nr_stats_sum, dim3(1), dim3(64), 0, st, g->stats_part, g->stats);
    }
    hipLaunchKernelGGL(k_pcg_start, dim3(pg), dim3(64 * kSpmvNodes), 0, st, S, g->H, g->row_start, g->row_ec,
                       g->Minv, nn);
    for (int k = 0; k < pcg_its; ++k)
      hipLaunchKernelGGL(k_pcg_iter, dim3(pg), dim3(64 * kSpmvNodes), 0, st, S, g->H, g->row_start, g->row_ec,
                         g->Minv, nn, k, F);
    hipLaunchKernelGGL(k_nr_update, dim3(grid_of(nn)), dim3(256), 0, st, S.x, g->T, nn, g->dq);
  }
}

static int nr_reserve(dm_graph *g, size_t P, int pcg_its) {
  if (P > g->pixels_cap) {
    DM_TRY(g->dalloc(&g->depth, P));
    DM_TRY(g->dalloc(&g->sdepth, P));
    DM_TRY(g->dalloc(&g->tv, P));
    DM_TRY(g->dalloc(&g->tn, P));
    g->pixels_cap = P;
    g->key_gn = -1;  // buffers moved: capture again
    ++g->buf_gen;
  }
  if (pcg_its + 1 > g->pcg_cap) {
    DM_TRY(g->dalloc(&g->pcg.rz, pcg_its + 1));
    DM_TRY(g->dalloc(&g->pcg.al, pcg_its + 1));
    g->pcg_cap = pcg_its + 1;
    g->key_gn = -1;
    ++g->buf_gen;
  }
  return DM_OK;
}

// The solve's frame record (pinned host copy; the caller uploads it).
static void fill_frame(dm_graph *g, int mode, const dm_sensor *sensor, const double pose[16],
                       const dm_nonrigid_props *pr, bool use_src) {
  NrFrame &F = *g->h_frame;
  F = NrFrame{};
  if (mode == 0) {
    F.rows = (int)sensor->rows; F.cols = (int)sensor->cols;
    F.fx = sensor->fx; F.fy = sensor->fy; F.cx = sensor->cx; F.cy = sensor->cy;
  }
  for (int k = 0; k < 16; ++k) F.pose[k] = pose[k];
  F.min_depth = pr->min_depth; F.max_depth = pr->max_depth; F.dist_max = pr->dist_max;
  F.normal_cos = pr->normal_cos; F.lambda = pr->lambda_reg; F.damping = pr->damping; F.pcg_tol = pr->pcg_tol;
  F.use_src = use_src;
}

// Copy a solve's statistics (3 energies, 2 PCG flags) into the graph's pinned
// buffer on stream st; read_stats takes them once st is done.
static int stats_copy(dm_graph *g, const dm_nonrigid_props *pr, hipStream_t st) {
  if (!g->h_stats) DM_HIP(hipHostMalloc((void **)&g->h_stats, 8 * sizeof(double), hipHostMallocDefault));
  for (int k = 0; k < 8; ++k) g->h_stats[k] = 0.0;
  if (pr->gn_iterations > 0) {
    DM_HIP(hipMemcpyAsync(g->h_stats, g->stats, 24, hipMemcpyDeviceToHost, st));
    DM_HIP(hipMemcpyAsync(g->h_stats + 3, g->pcg.flag, 8, hipMemcpyDeviceToHost, st));
  }
  return DM_OK;
}
static void read_stats(const dm_graph *g, double stats[4]) {
  int flags[2];
  memcpy(flags, g->h_stats + 3, sizeof flags);
  stats[0] = g->h_stats[0]; stats[1] = g->h_stats[1]; stats[2] = g->h_stats[2]; stats[3] = flags[1];
}

// One solve call: mode 0 = depth target (src may gate it), 1 = mesh target.
// Enqueue one solve on the graph's stream (uploads, the captured graph, the
// statistics' copy into pinned memory); nr_finish waits for it.
static int nr_launch(dm_graph *g, int mode, const float *depth, const float *src, const float *mv, const float *mn,
                     const dm_sensor *sensor, const double pose[16], const dm_nonrigid_props *pr) {
  if (!g || !pose || !pr || (mode == 0 && (!sensor || sensor->rows == 0 || sensor->cols == 0)) ||
      pr->gn_iterations < 0 || pr->pcg_iterations < 0 || pr->pcg_iterations > 1000) {
    set_error("dm_nonrigid_solve: invalid argument");
    return DM_E_INVALID;
  }
  hipStream_t st = g->stream;
  const size_t P = mode == 0 ? (size_t)sensor->rows * sensor->cols : 0;
  DM_TRY(nr_reserve(g, P, pr->pcg_iterations));
  if (mode == 1 && !g->mesh_v) {
    DM_TRY(g->dalloc(&g->mesh_v, (size_t)g->nv * 3));
    DM_TRY(g->dalloc(&g->mesh_n, (size_t)g->nv * 3));
  }
  DM_HIP(hipStreamSynchronize(st));  // the pinned frame record may still be read by the last call
  fill_frame(g, mode, sensor, pose, pr, src != nullptr);
  const NrFrame &F = *g->h_frame;
  DM_HIP(hipMemcpyAsync(g->d_frame, g->h_frame, sizeof(NrFrame), hipMemcpyHostToDevice, st));
  if (mode == 0) {
    DM_HIP(hipMemcpyAsync(g->depth, depth, P * 4, hipMemcpyHostToDevice, st));
    if (src) DM_HIP(hipMemcpyAsync(g->sdepth, src, P * 4, hipMemcpyHostToDevice, st));
  } else {
    DM_HIP(hipMemcpyAsync(g->mesh_v, mv, (size_t)g->nv * 12, hipMemcpyHostToDevice, st));
    DM_HIP(hipMemcpyAsync(g->mesh_n, mn, (size_t)g->nv * 12, hipMemcpyHostToDevice, st));
  }
  if (!g->exec || g->key_gn != pr->gn_iterations || g->key_pcg != pr->pcg_iterations || g->key_rows != F.rows ||
      g->key_cols != F.cols || g->key_mode != mode) {
    if (g->exec) { (void)hipGraphExecDestroy(g->exec); g->exec = nullptr; }
    hipGraph_t graph = nullptr;
    DM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    nr_enqueue(g, F.rows, F.cols, pr->gn_iterations, pr->pcg_iterations, mode);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
    ce = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ce != hipSuccess) { g->exec = nullptr; return hip_fail(ce, "hipGraphInstantiate"); }
    g->key_gn = pr->gn_iterations; g->key_pcg = pr->pcg_iterations;
    g->key_rows = F.rows; g->key_cols = F.cols; g->key_mode = mode;
  }
  DM_HIP(hipGraphLaunch(g->exec, st));
  return stats_copy(g, pr, st);
}

static int nr_finish(dm_graph *g, double stats[4]) {
  DM_HIP(hipStreamSynchronize(g->stream));
  if (stats) read_stats(g, stats);
  return DM_OK;
}

static int nr_solve(dm_graph *g, int mode, const float *depth, const float *src, const float *mv, const float *mn,
                    const dm_sensor *sensor, const double pose[16], const dm_nonrigid_props *pr, double stats[4]) {
  DM_TRY(nr_launch(g, mode, depth, src, mv, mn, sensor, pose, pr));
  return nr_finish(g, stats);
}

// ---------------------------------------------------------------- batches
// Several graphs' solves in shared launches: every kernel of the solve runs
// once per step for all graphs (graph = blockIdx.y), reading that graph's
// arguments from a table (kb_* kernels; the bodies are the single-graph
// kernels').  The tables and one captured graph of the whole solve belong
// to the batch; the frame buffers and statistics stay the graphs' own.
struct dm_nonrigid_batch {
  std::vector<dm_graph *> gs;
  std::vector<int> gens;          // the graphs' buf_gen at capture
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  int key_gn = -1, key_pcg = -1, key_rows = -1, key_cols = -1;
  char *tables = nullptr;         // device: every kernel's argument records
  size_t table_cap = 0;
  ~dm_nonrigid_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (tables) (void)hipFree(tables);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Host images of the argument tables of one batch, at 256 B aligned offsets.
struct BatchTables {
  std::vector<char> host;
  template <class T> size_t add(const std::vector<T> &v) {
    const size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + v.size() * sizeof(T));
    memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};
template <class T> static unsigned max_grid(const std::vector<T> &v) {
  unsigned m = 0;
  for (const T &e : v) m = std::max(m, e.grid);
  return m;
}

// Capture one solve of every graph of the batch on its stream (depth
// target, no source gate; the frames are rows x cols).
static int batch_capture(dm_nonrigid_batch *b, int rows, int cols, int gn, int pcg_its) {
  const int K = (int)b->gs.size();
  const size_t P = (size_t)rows * cols;
  std::vector<b_nr_target_vert> tvt(K);
  std::vector<b_nr_target_norm> tnm(K);
  std::vector<NrTerms> trm(K);
  std::vector<b_nr_assemble> as(K);
  std::vector<b_nr_invdiag> inv(K);
  std::vector<b_nr_stats> sta(K);
  std::vector<b_nr_stats_sum> ssum(K);
  std::vector<b_pcg_start> pst(K);
  std::vector<b_pcg_iter> pit(K);
  std::vector<b_nr_update> nup(K);
  for (int i = 0; i < K; ++i) {
    dm_graph *g = b->gs[i];
    const int nn = g->nn, ne = nn * kNK;
    const PcgState S = g->pcg;
    const NrFrame *F = g->d_frame;
    tvt[i] = {g->depth, F, g->tv, grid_of(P)};
    tnm[i] = {g->tv, F, g->tn, grid_of(P)};
    trm[i] = terms_of(g);
    as[i] = {g->blk_a, g->blk_b, g->cstart, g->codes, g->cw, g->nb, g->vdata, g->nbr, g->edata, F, g->H, g->rhs,
             g->Hd, grid_of(g->nb, 256 / kAsmLanes)};
    inv[i] = {g->Hd, nn, g->Minv, g->rhs, S.x, S.r, S.u, grid_of(nn)};
    sta[i] = {g->vdata, g->nv, g->edata, ne, F, g->stats_part, (unsigned)kStatsWg};
    ssum[i] = {g->stats_part, g->stats, 1u};
    pst[i] = {S, g->H, g->row_start, g->row_ec, g->Minv, nn, (unsigned)S.nwg};
    pit[i] = {S, g->H, g->row_start, g->row_ec, g->Minv, nn, F, (unsigned)S.nwg};
    nup[i] = {S.x, g->T, nn, g->dq, grid_of(nn)};
  }
  BatchTables T;
  const size_t o_tvt = T.add(tvt), o_tnm = T.add(tnm), o_trm = T.add(trm),
               o_as = T.add(as), o_inv = T.add(inv), o_sta = T.add(sta), o_ssum = T.add(ssum), o_pst = T.add(pst),
               o_pit = T.add(pit), o_nup = T.add(nup);
  if (T.host.size() > b->table_cap) {
    if (b->tables) DM_HIP(hipFree(b->tables));
    b->tables = nullptr;
    DM_HIP(hipMalloc((void **)&b->tables, T.host.size()));
    b->table_cap = T.host.size();
  }
  DM_HIP(hipMemcpy(b->tables, T.host.data(), T.host.size(), hipMemcpyHostToDevice));
  char *D = b->tables;
  hipStream_t st = b->stream;
  const unsigned Ky = (unsigned)K;
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  hipGraph_t graph = nullptr;
  DM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(kb_nr_target_vert, dim3(max_grid(tvt), Ky), dim3(256), 0, st,
                     (const b_nr_target_vert *)(D + o_tvt));
  hipLaunchKernelGGL(kb_nr_target_norm, dim3(max_grid(tnm), Ky), dim3(256), 0, st,
                     (const b_nr_target_norm *)(D + o_tnm));
  for (int it = 0; it < gn; ++it) {
    hipLaunchKernelGGL(kb_nr_terms, dim3(max_grid(trm), Ky), dim3(256), 0, st, (const NrTerms *)(D + o_trm));
    hipLaunchKernelGGL(kb_nr_assemble, dim3(max_grid(as), Ky), dim3(256), 0, st,
                       (const b_nr_assemble *)(D + o_as));
    hipLaunchKernelGGL(kb_nr_invdiag, dim3(max_grid(inv), Ky), dim3(256), 0, st, (const b_nr_invdiag *)(D + o_inv));
    if (it + 1 == gn) {
      hipLaunchKernelGGL(kb_nr_stats, dim3(max_grid(sta), Ky), dim3(256), 0, st, (const b_nr_stats *)(D + o_sta));
      hipLaunchKernelGGL(kb_nr_stats_sum, dim3(1, Ky), dim3(64), 0, st, (const b_nr_stats_sum *)(D + o_ssum));
    }
    hipLaunchKernelGGL(kb_pcg_start, dim3(max_grid(pst), Ky), dim3(64 * kSpmvNodes), 0, st,
                       (const b_pcg_start *)(D + o_pst));
    for (int k = 0; k < pcg_its; ++k)
      hipLaunchKernelGGL(kb_pcg_iter, dim3(max_grid(pit), Ky), dim3(64 * kSpmvNodes), 0, st,
                         (const b_pcg_iter *)(D + o_pit), k);
    hipLaunchKernelGGL(kb_nr_update, dim3(max_grid(nup), Ky), dim3(256), 0, st, (const b_nr_update *)(D + o_nup));
  }
  hipError_t ce = hipStreamEndCapture(st, &graph);
  if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture");
  ce = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ce != hipSuccess) { b->exec = nullptr; return hip_fail(ce, "hipGraphInstantiate"); }
  b->key_gn = gn; b->key_pcg = pcg_its; b->key_rows = rows; b->key_cols = cols;
  b->gens.resize(K);
  for (int i = 0; i < K; ++i) b->gens[i] = b->gs[i]->buf_gen;
  return DM_OK;
}

extern "C" {

int dm_graph_create(const float *vertices, const float *normals, int64_t nv, const dm_graph_props *props,
                    dm_graph **out) {
  if (!vertices || !normals || !props || !out || nv <= 0 || nv >= (1LL << 27)) {
    set_error("dm_graph_create: invalid argument (1 <= vertices < 2^27)");
    return DM_E_INVALID;
  }
  if (!(props->node_radius > 0.f) || !(props->dgw > 0.f) || props->max_nodes < kNK + 1 ||
      props->max_nodes > kMaxNodes) {
    set_error("dm_graph_create: node_radius, dgw > 0 and 5 <= max_nodes <= 32768 required");
    return DM_E_INVALID;
  }
  dm_graph *g = new dm_graph();
  g->nv = nv;
  g->props = *props;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete g;
    return hip_fail(e, "hipStreamCreate");
  }
  int rc = graph_build(g, vertices, normals);
  if (rc != DM_OK) {
    g->release();
    delete g;
    return rc;
  }
  *out = g;
  return DM_OK;
}

int dm_graph_destroy(dm_graph *g) {
  if (!g) return DM_OK;
  g->release();
  delete g;
  return DM_OK;
}

int dm_graph_num_nodes(dm_graph *g, int32_t *n) {
  if (!g || !n) { set_error("null argument"); return DM_E_INVALID; }
  *n = g->nn;
  return DM_OK;
}

int dm_graph_get_nodes(dm_graph *g, int32_t *vertex_ids, int32_t *nbr_ids, float *nbr_weights) {
  if (!g) { set_error("null argument"); return DM_E_INVALID; }
  if (vertex_ids) DM_HIP(hipMemcpyAsync(vertex_ids, g->node_vid, (size_t)g->nn * 4, hipMemcpyDeviceToHost, g->stream));
  if (nbr_ids) DM_HIP(hipMemcpyAsync(nbr_ids, g->nbr, (size_t)g->nn * kNK * 4, hipMemcpyDeviceToHost, g->stream));
  if (nbr_weights)
    DM_HIP(hipMemcpyAsync(nbr_weights, g->nbr_w, (size_t)g->nn * kNK * 4, hipMemcpyDeviceToHost, g->stream));
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_graph_get_vertex_nn(dm_graph *g, int32_t *ids, float *weights) {
  if (!g) { set_error("null argument"); return DM_E_INVALID; }
  if (ids) DM_HIP(hipMemcpyAsync(ids, g->vnn, (size_t)g->nv * kNK * 4, hipMemcpyDeviceToHost, g->stream));
  if (weights) DM_HIP(hipMemcpyAsync(weights, g->vnn_w, (size_t)g->nv * kNK * 4, hipMemcpyDeviceToHost, g->stream));
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_graph_num_blocks(dm_graph *g, int32_t *nb) {
  if (!g || !nb) { set_error("null argument"); return DM_E_INVALID; }
  *nb = g->nb;
  return DM_OK;
}

int dm_graph_set_transforms(dm_graph *g, const double *T) {
  if (!g || !T) { set_error("null argument"); return DM_E_INVALID; }
  DM_HIP(hipMemcpyAsync(g->T, T, (size_t)g->nn * 128, hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_nr_node_dq, dim3(grid_of(g->nn)), dim3(256), 0, g->stream, g->T, g->nn, g->dq);
  DM_HIP(hipGetLastError());
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_graph_get_transforms(dm_graph *g, double *T) {
  if (!g || !T) { set_error("null argument"); return DM_E_INVALID; }
  DM_HIP(hipMemcpyAsync(T, g->T, (size_t)g->nn * 128, hipMemcpyDeviceToHost, g->stream));
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_blend(dm_graph *g, float *out_vertices, float *out_normals) {
  if (!g || !out_vertices) { set_error("null argument"); return DM_E_INVALID; }
  hipLaunchKernelGGL(k_nr_blend, dim3(grid_of(g->nv)), dim3(256), 0, g->stream, g->verts, g->normals, g->nv,
                     g->node_of_vertex, g->vnn, g->vnn_w, g->dq, g->pv, g->pn);
  DM_HIP(hipGetLastError());
  DM_HIP(hipMemcpyAsync(out_vertices, g->pv, (size_t)g->nv * 12, hipMemcpyDeviceToHost, g->stream));
  if (out_normals) DM_HIP(hipMemcpyAsync(out_normals, g->pn, (size_t)g->nv * 12, hipMemcpyDeviceToHost, g->stream));
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_nonrigid_reserve(dm_graph *g, int64_t observations, int32_t pcg_iterations) {
  if (!g || pcg_iterations < 0 || pcg_iterations > 1000 || observations < 0 || observations > (1LL << 28)) {
    set_error("dm_nonrigid_reserve: invalid argument");
    return DM_E_INVALID;
  }
  DM_TRY(nr_reserve(g, (size_t)observations, pcg_iterations));
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

int dm_nonrigid_solve(dm_graph *g, const float *depth, const dm_sensor *sensor, const double pose[16],
                      const dm_nonrigid_props *pr, double stats[4]) {
  if (!depth) { set_error("dm_nonrigid_solve: invalid argument"); return DM_E_INVALID; }
  return nr_solve(g, 0, depth, nullptr, nullptr, nullptr, sensor, pose, pr, stats);
}

int dm_nonrigid_solve_many(dm_graph *const *graphs, int32_t n, const float *const *depths, const dm_sensor *sensor,
                           const double *poses, const dm_nonrigid_props *pr, double *stats) {
  if (!graphs || n < 1 || !depths || !poses) { set_error("dm_nonrigid_solve_many: invalid argument"); return DM_E_INVALID; }
  for (int i = 0; i < n; ++i) {
    if (!graphs[i] || !depths[i]) { set_error("dm_nonrigid_solve_many: null graph or frame"); return DM_E_INVALID; }
    for (int j = 0; j < i; ++j)
      if (graphs[j] == graphs[i]) { set_error("dm_nonrigid_solve_many: a graph appears twice"); return DM_E_INVALID; }
  }
  // every solve enqueued on its graph's stream before any is waited for, so
  // that the solves' short kernels overlap
  int rc = DM_OK, launched = 0;
  for (; launched < n && rc == DM_OK; ++launched)
    rc = nr_launch(graphs[launched], 0, depths[launched], nullptr, nullptr, nullptr, sensor, poses + 16 * launched, pr);
  for (int i = 0; i < launched; ++i) {
    const int r = nr_finish(graphs[i], stats ? stats + 4 * i : nullptr);
    if (rc == DM_OK) rc = r;
  }
  return rc;
}

int dm_nonrigid_batch_create(dm_graph *const *graphs, int32_t n, dm_nonrigid_batch **out) {
  if (!graphs || n < 1 || n > 65535 || !out) { set_error("dm_nonrigid_batch_create: invalid argument"); return DM_E_INVALID; }
  for (int i = 0; i < n; ++i) {
    if (!graphs[i]) { set_error("dm_nonrigid_batch_create: null graph"); return DM_E_INVALID; }
    for (int j = 0; j < i; ++j)
      if (graphs[j] == graphs[i]) { set_error("dm_nonrigid_batch_create: a graph appears twice"); return DM_E_INVALID; }
  }
  auto *b = new dm_nonrigid_batch();
  b->gs.assign(graphs, graphs + n);
  const hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete b; return hip_fail(e, "hipStreamCreateWithFlags"); }
  *out = b;
  return DM_OK;
}

int dm_nonrigid_batch_solve(dm_nonrigid_batch *b, const float *const *depths, const dm_sensor *sensor,
                            const double *poses, const dm_nonrigid_props *pr, double *stats) {
  if (!b || !depths || !sensor || !poses || !pr || sensor->rows == 0 || sensor->cols == 0 ||
      pr->gn_iterations < 0 || pr->pcg_iterations < 0 || pr->pcg_iterations > 1000) {
    set_error("dm_nonrigid_batch_solve: invalid argument");
    return DM_E_INVALID;
  }
  const int K = (int)b->gs.size();
  for (int i = 0; i < K; ++i)
    if (!depths[i]) { set_error("dm_nonrigid_batch_solve: null frame"); return DM_E_INVALID; }
  const size_t P = (size_t)sensor->rows * sensor->cols;
  DM_HIP(hipStreamSynchronize(b->stream));  // the pinned frame records may still be read by the last call
  for (dm_graph *g : b->gs) {
    DM_HIP(hipStreamSynchronize(g->stream));  // single-graph work on the graph's own stream
    DM_TRY(nr_reserve(g, P, pr->pcg_iterations));
  }
  for (int i = 0; i < K; ++i) {
    dm_graph *g = b->gs[i];
    fill_frame(g, 0, sensor, poses + 16 * (size_t)i, pr, false);
    DM_HIP(hipMemcpyAsync(g->d_frame, g->h_frame, sizeof(NrFrame), hipMemcpyHostToDevice, b->stream));
    DM_HIP(hipMemcpyAsync(g->depth, depths[i], P * 4, hipMemcpyHostToDevice, b->stream));
  }
  bool moved = !b->exec || b->key_gn != pr->gn_iterations || b->key_pcg != pr->pcg_iterations ||
               b->key_rows != (int)sensor->rows || b->key_cols != (int)sensor->cols || (int)b->gens.size() != K;
  for (int i = 0; i < K && !moved; ++i) moved = b->gens[i] != b->gs[i]->buf_gen;
  if (moved) {
    DM_HIP(hipStreamSynchronize(b->stream));
    DM_TRY(batch_capture(b, (int)sensor->rows, (int)sensor->cols, pr->gn_iterations, pr->pcg_iterations));
  }
  DM_HIP(hipGraphLaunch(b->exec, b->stream));
  for (dm_graph *g : b->gs) DM_TRY(stats_copy(g, pr, b->stream));
  DM_HIP(hipStreamSynchronize(b->stream));
  if (stats)
    for (int i = 0; i < K; ++i) read_stats(b->gs[i], stats + 4 * (size_t)i);
  return DM_OK;
}

int dm_nonrigid_batch_destroy(dm_nonrigid_batch *b) {
  delete b;
  return DM_OK;
}

int dm_nonrigid_solve_pair(dm_graph *g, const float *target_depth, const float *source_depth,
                           const dm_sensor *sensor, const double pose[16], const dm_nonrigid_props *pr,
                           double stats[4]) {
  if (!target_depth || !source_depth) { set_error("dm_nonrigid_solve_pair: invalid argument"); return DM_E_INVALID; }
  return nr_solve(g, 0, target_depth, source_depth, nullptr, nullptr, sensor, pose, pr, stats);
}

int dm_nonrigid_solve_mesh(dm_graph *g, const float *target_vertices, const float *target_normals,
                           const double pose[16], const dm_nonrigid_props *pr, double stats[4]) {
  if (!target_vertices || !target_normals) {
    set_error("dm_nonrigid_solve_mesh: invalid argument");
    return DM_E_INVALID;
  }
  return nr_solve(g, 1, nullptr, nullptr, target_vertices, target_normals, nullptr, pose, pr, stats);
}

int dm_graph_create_nodes(const float *vertices, const float *normals, int64_t nv, const int32_t *node_vertex_ids,
                          int32_t num_nodes, float dgw, dm_graph **out) {
  if (!vertices || !normals || !node_vertex_ids || !out || nv <= 0 || nv >= (1LL << 27) || !(dgw > 0.f) ||
      num_nodes < kNK + 1 || num_nodes > kMaxNodes || num_nodes > nv) {
    set_error("dm_graph_create_nodes: invalid argument (5 <= nodes <= min(32768, vertices), dgw > 0)");
    return DM_E_INVALID;
  }
  dm_graph *g = new dm_graph();
  g->nv = nv;
  g->props = dm_graph_props{0.f, dgw, num_nodes};
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete g;
    return hip_fail(e, "hipStreamCreate");
  }
  int rc = graph_build_nodes(g, vertices, normals, node_vertex_ids, num_nodes);
  if (rc != DM_OK) {
    g->release();
    delete g;
    return rc;
  }
  *out = g;
  return DM_OK;
}

int dm_graph_update_nodes(dm_graph *g) {
  if (!g) { set_error("null argument"); return DM_E_INVALID; }
  DM_TRY(graph_links(g));
  hipLaunchKernelGGL(k_nr_contrib_weights, dim3(grid_of(g->ncontrib)), dim3(256), 0, g->stream, g->codes,
                     g->ncontrib, g->node_of_vertex, g->vnn_wn, g->cw);
  DM_HIP(hipGetLastError());
  DM_HIP(hipStreamSynchronize(g->stream));
  return DM_OK;
}

}  // extern "C"
