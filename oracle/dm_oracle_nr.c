/*
 * CPU oracle, non-rigid part: deformation graph, dual-quaternion blending and
 * the non-rigid Gauss-Newton step.  TEST INFRASTRUCTURE ONLY (see
 * dm_oracle.c): used by the tests, the smoke check and the benchmark's CPU
 * baseline as the checker; nothing in the product links it.
 *
 * Built as FP32 transcription (-DOR_REAL=float: integer outputs -- node set,
 * neighbour ids -- are defined by this build) and as FP64 oracle.
 *
 * Reference (paths relative to the reference root):
 *   include/geometry/defGraph/defGraph.cu   graph nodes, distances, sort, weights
 *   include/blender/dqBlender/dqBlender.cu  blendVertex :162-186
 *   include/math/dualQuat.h                 transformPosition :293-302
 *   include/math/Quaternion.h               rotate :278-282
 *   include/solver/nonRigidICP/nonRigidICP.cu getRegResiduals :494-515
 *
 * K (kept): distances clamped below at 1e-4 (defGraph.cu:436); weights
 * exp(-d^2 / dgw^2) (defGraph.cu:131,410); a vertex that is a node takes
 * that node's dual quaternion (dqBlender.cu:178-182); translation
 * t = 2 (v_d w_r - v_r w_d + v_r x v_d) and rotation
 * v + 2 q_v x (q_v x v + w v) (dualQuat.h:293-302, Quaternion.h:278-282);
 * regulariser residual w_ij (T_i g_j - T_j g_j) with
 * w_ij = max(w_i[k], w_j[k]) for the k-th neighbour j of node i
 * (nonRigidICP.cu:494-515).
 * D (diverge, with reason):
 *  - nodes: in every occupied cell of a grid of edge node_radius, the vertex
 *    closest to the cell centre (ties to the lower index), numbered in
 *    vertex order (the reference takes the first NODE_NUM vertices or random
 *    ones, defGraph.cu:673-702); deterministic and spatially uniform;
 *  - neighbours: K = 4 nearest by squared distance, ties to the lower index
 *    (the reference's insertion sort keeps equal distances in input order,
 *    which is the same rule);
 *  - blending: weighted sum of the neighbours' dual quaternions with Kavan's
 *    antipodality sign (dot with the first), normalised once at the end
 *    (Kavan et al. 2007); the reference normalises after every addition
 *    (dqBlender.cu:176), which over-weights later neighbours;
 *  - normals are rotated only (the reference adds the translation,
 *    dualQuat.h:315-324);
 *  - Gauss-Newton: point-to-plane data term with analytic Jacobians of the
 *    linear-blend approximation dp'/dxi_k = w_k [p' x ., I] (the reference
 *    uses an unsigned finite-difference norm, nonRigidICP.cu:131-236), ARAP
 *    term weighted lambda, block-sparse 6x6 normal equations, block-Jacobi
 *    PCG, increments composed per node T_i <- exp(xi_i) T_i (the reference
 *    discards its solution, nonRigidICP.cu:1168-1243).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OR_REAL
#define OR_REAL float
#endif
typedef OR_REAL real;
#define OR_EXPORT __attribute__((visibility("default")))
#define NR_K 4

void or_se3_exp(const double xi[6], double T[16]); /* dm_oracle.c */
int or_ldlt6(const double *Ain, const double *b, double *x);

typedef struct { real w, x, y, z; } quat;
typedef struct { quat r, d; } dquat;

typedef struct {
  int64_t nv;
  int nn;
  float *verts, *normals;     /* canonical, nv * 3 */
  int32_t *node_vid;          /* nn */
  int32_t *node_of_vertex;    /* nv, -1 = none */
  int32_t *nbr;               /* nn * K node neighbours */
  real *nbr_w;                /* nn * K */
  int32_t *vnn;               /* nv * K vertex -> nodes */
  real *vnn_w;                /* nv * K raw weights */
  double *T;                  /* nn * 16 node transforms (canonical -> current) */
  real dgw;
} or_graph;

/* ---------------------------------------------------------------- graph */
static real clamp_dist(real d2) {
  real d = (real)sqrt((double)d2); /* correctly rounded in either precision */
  return d < (real)1e-4f ? (real)1e-4f : d;
}
static real weight_of(real d, real dgw) { return (real)exp(-(double)((d * d) / (dgw * dgw))); }

static real sqdist(const float *a, const float *b) {
  real dx = (real)a[0] - (real)b[0], dy = (real)a[1] - (real)b[1], dz = (real)a[2] - (real)b[2];
  return dx * dx + dy * dy + dz * dz;
}

/* K nearest of query among the nodes (excluding node `self`, -1 = none);
 * distance ascending, ties to the lower node index */
static void knn_nodes(const or_graph *g, const float *q, int self, int32_t *ids, real *d2s) {
  for (int k = 0; k < NR_K; ++k) { ids[k] = -1; d2s[k] = (real)INFINITY; }
  for (int j = 0; j < g->nn; ++j) {
    if (j == self) continue;
    real d2 = sqdist(q, g->verts + 3 * (size_t)g->node_vid[j]);
    if (!(d2 < d2s[NR_K - 1])) continue;
    int k = NR_K - 1;
    while (k > 0 && d2 < d2s[k - 1]) { d2s[k] = d2s[k - 1]; ids[k] = ids[k - 1]; --k; }
    d2s[k] = d2; ids[k] = j;
  }
}

static int cell_coord(float x, real r) { return (int)floor((double)((real)x / r)); }

/* squared distance of a vertex to the centre of its cell, (c + 1/2) r */
static real cell_d2(const float *p, const int c[3], real r) {
  real d2 = 0;
  for (int k = 0; k < 3; ++k) {
    real e = (real)p[k] - ((real)c[k] + (real)0.5f) * r;
    d2 = d2 + e * e;
  }
  return d2;
}

typedef struct { int cx, cy, cz; real d2; int64_t v; } cell_vertex;
static int cmp_cell_vertex(const void *a, const void *b) {
  const cell_vertex *p = (const cell_vertex *)a, *q = (const cell_vertex *)b;
  if (p->cx != q->cx) return p->cx < q->cx ? -1 : 1;
  if (p->cy != q->cy) return p->cy < q->cy ? -1 : 1;
  if (p->cz != q->cz) return p->cz < q->cz ? -1 : 1;
  if (p->d2 != q->d2) return p->d2 < q->d2 ? -1 : 1;
  return p->v < q->v ? -1 : (p->v > q->v);
}

static void graph_links(or_graph *g);

/*
 * defGraph::init + defGraphUpdateNodes (defGraph.cu:15-69, :483-545).
 * status: 0 ok, -1 fewer than K+1 nodes, -2 more than max_nodes.
 */
OR_EXPORT or_graph *or_graph_create(const float *verts, const float *normals, int64_t nv, double node_radius,
                                    double dgw, int max_nodes, int *status) {
  or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
  g->nv = nv;
  g->dgw = (real)dgw;
  g->verts = (float *)malloc(nv * 3 * sizeof(float));
  g->normals = (float *)malloc(nv * 3 * sizeof(float));
  memcpy(g->verts, verts, nv * 3 * sizeof(float));
  memcpy(g->normals, normals, nv * 3 * sizeof(float));
  /* cells: per occupied cell the vertex closest to the cell centre (ties to
   * the lower index), by sorting (cell, d2, index) */
  const real r = (real)node_radius;
  cell_vertex *c = (cell_vertex *)malloc(nv * sizeof(cell_vertex));
  for (int64_t v = 0; v < nv; ++v) {
    int cc[3] = {cell_coord(verts[3 * v], r), cell_coord(verts[3 * v + 1], r), cell_coord(verts[3 * v + 2], r)};
    c[v].cx = cc[0]; c[v].cy = cc[1]; c[v].cz = cc[2];
    c[v].d2 = cell_d2(verts + 3 * v, cc, r);
    c[v].v = v;
  }
  qsort(c, nv, sizeof(cell_vertex), cmp_cell_vertex);
  uint8_t *isnode = (uint8_t *)calloc(nv, 1);
  int nn = 0;
  for (int64_t i = 0; i < nv; ++i)
    if (i == 0 || c[i].cx != c[i - 1].cx || c[i].cy != c[i - 1].cy || c[i].cz != c[i - 1].cz) {
      isnode[c[i].v] = 1;
      ++nn;
    }
  free(c);
  *status = 0;
  if (nn < NR_K + 1) *status = -1;
  if (nn > max_nodes) *status = -2;
  g->nn = nn;
  g->node_vid = (int32_t *)malloc((nn ? nn : 1) * sizeof(int32_t));
  g->node_of_vertex = (int32_t *)malloc(nv * sizeof(int32_t));
  int k = 0;
  for (int64_t v = 0; v < nv; ++v) {
    g->node_of_vertex[v] = isnode[v] ? k : -1;
    if (isnode[v]) g->node_vid[k++] = (int32_t)v;
  }
  free(isnode);
  if (*status != 0) {
    g->T = (double *)calloc((size_t)(nn ? nn : 1) * 16, sizeof(double));
    return g;
  }
  graph_links(g);
  return g;
}

/* defGraph::defGraphUpdateNodes (defGraph.cu:483-545) and the vertex
 * neighbours of updateActiveNeighbourNodes (:243-302): the K nearest nodes
 * of every node and of every vertex with weights exp(-d^2/dgw^2).  The
 * first call also sets every node transform to the identity. */
static void graph_links(or_graph *g) {
  const int nn = g->nn;
  const int64_t nv = g->nv;
  if (!g->nbr) {
    g->nbr = (int32_t *)malloc((size_t)(nn ? nn : 1) * NR_K * sizeof(int32_t));
    g->nbr_w = (real *)malloc((size_t)(nn ? nn : 1) * NR_K * sizeof(real));
    g->vnn = (int32_t *)malloc((size_t)nv * NR_K * sizeof(int32_t));
    g->vnn_w = (real *)malloc((size_t)nv * NR_K * sizeof(real));
    g->T = (double *)calloc((size_t)(nn ? nn : 1) * 16, sizeof(double));
    for (int i = 0; i < nn; ++i) { g->T[16 * i] = g->T[16 * i + 5] = g->T[16 * i + 10] = g->T[16 * i + 15] = 1.0; }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nn; ++i) {
    real d2[NR_K];
    knn_nodes(g, g->verts + 3 * (size_t)g->node_vid[i], i, g->nbr + NR_K * i, d2);
    for (int q = 0; q < NR_K; ++q) g->nbr_w[NR_K * i + q] = weight_of(clamp_dist(d2[q]), g->dgw);
  }
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < nv; ++v) {
    real d2[NR_K];
    knn_nodes(g, g->verts + 3 * v, -1, g->vnn + NR_K * v, d2);
    for (int q = 0; q < NR_K; ++q) g->vnn_w[NR_K * v + q] = weight_of(clamp_dist(d2[q]), g->dgw);
  }
}

/*
 * defGraph::init with the nodes given (defGraph.cu:15-69 with
 * sampleDownMeshHost, :673-702): node k is vertex node_ids[k] (mode 1 of the
 * reference passes the first NODE_NUM vertices, mode 0 random ones).
 * status: 0 ok, -1 fewer than K+1 nodes, -3 an id out of range or repeated.
 */
OR_EXPORT or_graph *or_graph_create_nodes(const float *verts, const float *normals, int64_t nv,
                                          const int32_t *node_ids, int nn, double dgw, int *status) {
  or_graph *g = (or_graph *)calloc(1, sizeof(or_graph));
  g->nv = nv;
  g->dgw = (real)dgw;
  g->verts = (float *)malloc(nv * 3 * sizeof(float));
  g->normals = (float *)malloc(nv * 3 * sizeof(float));
  memcpy(g->verts, verts, nv * 3 * sizeof(float));
  memcpy(g->normals, normals, nv * 3 * sizeof(float));
  g->nn = nn;
  g->node_vid = (int32_t *)malloc((nn > 0 ? nn : 1) * sizeof(int32_t));
  g->node_of_vertex = (int32_t *)malloc(nv * sizeof(int32_t));
  for (int64_t v = 0; v < nv; ++v) g->node_of_vertex[v] = -1;
  *status = nn < NR_K + 1 ? -1 : 0;
  for (int k = 0; k < nn && *status == 0; ++k) {
    const int32_t id = node_ids[k];
    if (id < 0 || id >= nv || g->node_of_vertex[id] >= 0) { *status = -3; break; }
    g->node_of_vertex[id] = k;
    g->node_vid[k] = id;
  }
  if (*status != 0) {
    g->nn = 0;
    g->T = (double *)calloc(16, sizeof(double));
    return g;
  }
  graph_links(g);
  return g;
}

/* defGraph::defGraphUpdateNodes: neighbours and weights recomputed from the
 * canonical positions; node transforms kept. */
OR_EXPORT void or_graph_update_nodes(or_graph *g) { graph_links(g); }

OR_EXPORT void or_graph_destroy(or_graph *g) {
  if (!g) return;
  free(g->verts); free(g->normals); free(g->node_vid); free(g->node_of_vertex); free(g->nbr); free(g->nbr_w);
  free(g->vnn); free(g->vnn_w); free(g->T); free(g);
}

OR_EXPORT int or_graph_num_nodes(const or_graph *g) { return g->nn; }

OR_EXPORT void or_graph_nodes(const or_graph *g, int32_t *vid, int32_t *nbr, double *nbr_w) {
  for (int i = 0; i < g->nn; ++i) {
    if (vid) vid[i] = g->node_vid[i];
    for (int q = 0; q < NR_K; ++q) {
      if (nbr) nbr[NR_K * i + q] = g->nbr[NR_K * i + q];
      if (nbr_w) nbr_w[NR_K * i + q] = (double)g->nbr_w[NR_K * i + q];
    }
  }
}

OR_EXPORT void or_graph_vertex_nn(const or_graph *g, int32_t *ids, double *w) {
  for (int64_t v = 0; v < g->nv * NR_K; ++v) {
    if (ids) ids[v] = g->vnn[v];
    if (w) w[v] = (double)g->vnn_w[v];
  }
}

OR_EXPORT void or_graph_set_transforms(or_graph *g, const double *T) { memcpy(g->T, T, (size_t)g->nn * 16 * sizeof(double)); }
OR_EXPORT void or_graph_get_transforms(const or_graph *g, double *T) { memcpy(T, g->T, (size_t)g->nn * 16 * sizeof(double)); }

/* ---------------------------------------------------------------- DQ */
/* unit quaternion of a rotation matrix (row-major 3x3 inside a 4x4), w >= 0,
 * in FP64 (Shepperd's method) */
static void quat_of_rot(const double *T, double q[4]) {
  const double m00 = T[0], m01 = T[1], m02 = T[2], m10 = T[4], m11 = T[5], m12 = T[6], m20 = T[8], m21 = T[9],
               m22 = T[10];
  const double tr = m00 + m11 + m22;
  double w, x, y, z;
  if (tr > 0) {
    double s = sqrt(tr + 1.0) * 2.0;
    w = 0.25 * s; x = (m21 - m12) / s; y = (m02 - m20) / s; z = (m10 - m01) / s;
  } else if (m00 > m11 && m00 > m22) {
    double s = sqrt(1.0 + m00 - m11 - m22) * 2.0;
    w = (m21 - m12) / s; x = 0.25 * s; y = (m01 + m10) / s; z = (m02 + m20) / s;
  } else if (m11 > m22) {
    double s = sqrt(1.0 + m11 - m00 - m22) * 2.0;
    w = (m02 - m20) / s; x = (m01 + m10) / s; y = 0.25 * s; z = (m12 + m21) / s;
  } else {
    double s = sqrt(1.0 + m22 - m00 - m11) * 2.0;
    w = (m10 - m01) / s; x = (m02 + m20) / s; y = (m12 + m21) / s; z = 0.25 * s;
  }
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  double n = sqrt(w * w + x * x + y * y + z * z);
  q[0] = w / n; q[1] = x / n; q[2] = y / n; q[3] = z / n;
}

/* unit dual quaternion of a rigid transform: real = q, dual = 0.5 (0, t) q,
 * in FP64 then rounded */
static dquat dq_of_T(const double *T) {
  double q[4];
  quat_of_rot(T, q);
  const double tx = T[3], ty = T[7], tz = T[11];
  /* (0, t) * q = (-t.qv, q_w t + t x qv) */
  double dw = -(tx * q[1] + ty * q[2] + tz * q[3]);
  double dx = q[0] * tx + (ty * q[3] - tz * q[2]);
  double dy = q[0] * ty + (tz * q[1] - tx * q[3]);
  double dz = q[0] * tz + (tx * q[2] - ty * q[1]);
  dquat o;
  o.r.w = (real)q[0]; o.r.x = (real)q[1]; o.r.y = (real)q[2]; o.r.z = (real)q[3];
  o.d.w = (real)(0.5 * dw); o.d.x = (real)(0.5 * dx); o.d.y = (real)(0.5 * dy); o.d.z = (real)(0.5 * dz);
  return o;
}

OR_EXPORT void or_dq_of_transform(const double T[16], double out[8]) {
  dquat d = dq_of_T(T);
  out[0] = d.r.w; out[1] = d.r.x; out[2] = d.r.y; out[3] = d.r.z;
  out[4] = d.d.w; out[5] = d.d.x; out[6] = d.d.y; out[7] = d.d.z;
}

/* Quaternion::rotate (Quaternion.h:278-282): v + 2 qv x (qv x v + w v) */
static void q_rotate(const quat *q, const real v[3], real o[3]) {
  const real qx = q->x * (real)2, qy = q->y * (real)2, qz = q->z * (real)2;
  const real cx = q->y * v[2] - q->z * v[1], cy = q->z * v[0] - q->x * v[2], cz = q->x * v[1] - q->y * v[0];
  const real ax = cx + v[0] * q->w, ay = cy + v[1] * q->w, az = cz + v[2] * q->w;
  o[0] = v[0] + (qy * az - qz * ay);
  o[1] = v[1] + (qz * ax - qx * az);
  o[2] = v[2] + (qx * ay - qy * ax);
}

/* blended, normalised dual quaternion of vertex v under the node dqs */
static dquat blend_dq(const or_graph *g, const dquat *ndq, int64_t v) {
  if (g->node_of_vertex[v] >= 0) return ndq[g->node_of_vertex[v]];
  dquat b;
  memset(&b, 0, sizeof b);
  const quat *r0 = &ndq[g->vnn[NR_K * v]].r;
  for (int k = 0; k < NR_K; ++k) {
    const dquat *q = &ndq[g->vnn[NR_K * v + k]];
    real w = g->vnn_w[NR_K * v + k];
    real dt = r0->w * q->r.w + r0->x * q->r.x + r0->y * q->r.y + r0->z * q->r.z;
    if (dt < (real)0) w = -w;
    b.r.w = b.r.w + w * q->r.w; b.r.x = b.r.x + w * q->r.x; b.r.y = b.r.y + w * q->r.y; b.r.z = b.r.z + w * q->r.z;
    b.d.w = b.d.w + w * q->d.w; b.d.x = b.d.x + w * q->d.x; b.d.y = b.d.y + w * q->d.y; b.d.z = b.d.z + w * q->d.z;
  }
  real n2 = b.r.w * b.r.w + b.r.x * b.r.x + b.r.y * b.r.y + b.r.z * b.r.z;
  real n = sizeof(real) == sizeof(float) ? (real)sqrtf((float)n2) : (real)sqrt((double)n2);
  b.r.w = b.r.w / n; b.r.x = b.r.x / n; b.r.y = b.r.y / n; b.r.z = b.r.z / n;
  b.d.w = b.d.w / n; b.d.x = b.d.x / n; b.d.y = b.d.y / n; b.d.z = b.d.z / n;
  return b;
}

/* dualQuat::transformPosition (dualQuat.h:293-302) for a unit dq; normals
 * rotated only (D) */
static void dq_apply(const dquat *b, const float *p, const float *n, real po[3], real no[3]) {
  const real pv[3] = {(real)p[0], (real)p[1], (real)p[2]};
  real rp[3];
  q_rotate(&b->r, pv, rp);
  const real vrx = b->r.x, vry = b->r.y, vrz = b->r.z, vdx = b->d.x, vdy = b->d.y, vdz = b->d.z;
  const real tx = ((vdx * b->r.w - vrx * b->d.w) + (vry * vdz - vrz * vdy)) * (real)2;
  const real ty = ((vdy * b->r.w - vry * b->d.w) + (vrz * vdx - vrx * vdz)) * (real)2;
  const real tz = ((vdz * b->r.w - vrz * b->d.w) + (vrx * vdy - vry * vdx)) * (real)2;
  po[0] = rp[0] + tx; po[1] = rp[1] + ty; po[2] = rp[2] + tz;
  if (n) {
    const real nv[3] = {(real)n[0], (real)n[1], (real)n[2]};
    q_rotate(&b->r, nv, no);
  }
}

static dquat *node_dqs(const or_graph *g) {
  dquat *d = (dquat *)malloc((size_t)g->nn * sizeof(dquat));
  for (int i = 0; i < g->nn; ++i) d[i] = dq_of_T(g->T + 16 * (size_t)i);
  return d;
}

/* dqBlender::blendMesh (dqBlender.cu:187-218): warped vertices and normals */
OR_EXPORT void or_blend(const or_graph *g, float *out_v, float *out_n) {
  dquat *nd = node_dqs(g);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < g->nv; ++v) {
    dquat b = blend_dq(g, nd, v);
    real po[3], no[3];
    dq_apply(&b, g->verts + 3 * v, g->normals + 3 * v, po, no);
    for (int k = 0; k < 3; ++k) {
      out_v[3 * v + k] = (float)po[k];
      if (out_n) out_n[3 * v + k] = (float)no[k];
    }
  }
  free(nd);
}

/* ---------------------------------------------------------------- GN */
typedef struct {
  int gn_iterations;     /* Gauss-Newton iterations per call */
  int pcg_iterations;    /* PCG iteration cap */
  double pcg_tol;        /* stop when r.z <= tol^2 * r0.z0 */
  double lambda_reg;     /* ARAP weight */
  double damping;        /* added to every diagonal entry */
  double dist_max;       /* association distance gate (m) */
  double normal_cos;     /* |n_src . n_tgt| gate */
  double min_depth, max_depth;
} or_nr_props;

/* block-sparse symmetric matrix: for node a, the list of b >= a */
typedef struct {
  int nn;
  int *start, *cols; /* CSR over upper blocks (a, b >= a), cols sorted */
  double *blk;       /* 36 per block, row-major */
  double *rhs;       /* 6 per node */
} bsm;

static int bsm_find(const bsm *m, int a, int b) {
  for (int k = m->start[a]; k < m->start[a + 1]; ++k)
    if (m->cols[k] == b) return k;
  return -1;
}

static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return (x > y) - (x < y); }

/* pattern: data pairs from the vertices' neighbour sets + ARAP edges +
 * every diagonal */
typedef struct { int *cnt, *caps; int **lst; } pat;
static void pat_add(pat *P, int a, int b) {
  if (a > b) { int t = a; a = b; b = t; }
  for (int k = 0; k < P->cnt[a]; ++k) if (P->lst[a][k] == b) return;
  if (P->cnt[a] == P->caps[a]) { P->caps[a] *= 2; P->lst[a] = (int *)realloc(P->lst[a], P->caps[a] * sizeof(int)); }
  P->lst[a][P->cnt[a]++] = b;
}

static void bsm_build(bsm *m, const or_graph *g) {
  int nn = g->nn;
  pat P;
  P.cnt = (int *)calloc(nn, sizeof(int));
  P.caps = (int *)malloc(nn * sizeof(int));
  P.lst = (int **)malloc(nn * sizeof(int *));
  for (int i = 0; i < nn; ++i) { P.caps[i] = 32; P.lst[i] = (int *)malloc(32 * sizeof(int)); }
  for (int i = 0; i < nn; ++i) pat_add(&P, i, i);
  for (int64_t v = 0; v < g->nv; ++v) {
    if (g->node_of_vertex[v] >= 0) continue;
    for (int p = 0; p < NR_K; ++p)
      for (int q = p + 1; q < NR_K; ++q) pat_add(&P, g->vnn[NR_K * v + p], g->vnn[NR_K * v + q]);
  }
  for (int i = 0; i < nn; ++i)
    for (int k = 0; k < NR_K; ++k) pat_add(&P, i, g->nbr[NR_K * i + k]);
  m->nn = nn;
  m->start = (int *)malloc((nn + 1) * sizeof(int));
  m->start[0] = 0;
  for (int i = 0; i < nn; ++i) m->start[i + 1] = m->start[i] + P.cnt[i];
  m->cols = (int *)malloc((m->start[nn] ? m->start[nn] : 1) * sizeof(int));
  for (int i = 0; i < nn; ++i) {
    qsort(P.lst[i], P.cnt[i], sizeof(int), cmp_int);
    memcpy(m->cols + m->start[i], P.lst[i], P.cnt[i] * sizeof(int));
    free(P.lst[i]);
  }
  free(P.lst); free(P.cnt); free(P.caps);
  m->blk = (double *)calloc((size_t)m->start[nn] * 36, sizeof(double));
  m->rhs = (double *)calloc((size_t)nn * 6, sizeof(double));
}

static void bsm_free(bsm *m) { free(m->start); free(m->cols); free(m->blk); free(m->rhs); }

/* H[a][b] += s * u v^T (a <= b handled by the caller's ordering) */
static void blk_add(bsm *m, int a, int b, double s, const double *u, const double *v) {
  if (a > b) { int t = a; a = b; b = t; const double *w = u; u = v; v = w; }
  double *B = m->blk + 36 * (size_t)bsm_find(m, a, b);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) B[6 * r + c] += s * u[r] * v[c];
}

/* y = H x over the symmetric block matrix */
static void bsm_mul(const bsm *m, const double *x, double *y) {
  memset(y, 0, (size_t)m->nn * 6 * sizeof(double));
  for (int a = 0; a < m->nn; ++a)
    for (int k = m->start[a]; k < m->start[a + 1]; ++k) {
      int b = m->cols[k];
      const double *B = m->blk + 36 * (size_t)k;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          y[6 * a + r] += B[6 * r + c] * x[6 * b + c];
          if (b != a) y[6 * b + c] += B[6 * r + c] * x[6 * a + r];
        }
    }
}

/* inverse of the 6x6 diagonal blocks (FP64, via LDL^T columns) */
static void diag_inv(const bsm *m, double *Minv) {
  for (int a = 0; a < m->nn; ++a) {
    const double *B = m->blk + 36 * (size_t)bsm_find(m, a, a);
    for (int c = 0; c < 6; ++c) {
      double e[6] = {0, 0, 0, 0, 0, 0}, col[6];
      e[c] = 1.0;
      if (or_ldlt6(B, e, col) != 0) memset(col, 0, sizeof col);
      for (int r = 0; r < 6; ++r) Minv[36 * (size_t)a + 6 * r + c] = col[r];
    }
  }
}

/* block-Jacobi preconditioned CG; returns the iterations run */
static int pcg(const bsm *m, const double *b, double *x, int iters, double tol) {
  const int n = 6 * m->nn;
  double *r = (double *)malloc(n * 8), *z = (double *)malloc(n * 8), *p = (double *)malloc(n * 8),
         *q = (double *)malloc(n * 8), *Mi = (double *)malloc((size_t)m->nn * 36 * 8);
  diag_inv(m, Mi);
  memset(x, 0, n * 8);
  memcpy(r, b, n * 8);
  for (int a = 0; a < m->nn; ++a)
    for (int i = 0; i < 6; ++i) {
      double s = 0;
      for (int j = 0; j < 6; ++j) s += Mi[36 * a + 6 * i + j] * r[6 * a + j];
      z[6 * a + i] = s;
    }
  memcpy(p, z, n * 8);
  double rz = 0;
  for (int i = 0; i < n; ++i) rz += r[i] * z[i];
  const double rz0 = rz;
  int it = 0;
  while (it < iters && rz > tol * tol * rz0 && rz > 0) {
    bsm_mul(m, p, q);
    double pq = 0;
    for (int i = 0; i < n; ++i) pq += p[i] * q[i];
    const double alpha = rz / pq;
    for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    for (int a = 0; a < m->nn; ++a)
      for (int i = 0; i < 6; ++i) {
        double s = 0;
        for (int j = 0; j < 6; ++j) s += Mi[36 * a + 6 * i + j] * r[6 * a + j];
        z[6 * a + i] = s;
      }
    double rzn = 0;
    for (int i = 0; i < n; ++i) rzn += r[i] * z[i];
    const double beta = rzn / rz;
    rz = rzn;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  free(r); free(z); free(p); free(q); free(Mi);
  return it;
}

/* target vertex / normal at pixel (u, v) in camera coordinates (FP64, the
 * rigid path's live-frame map rule: cross(V[i+2]-V[i-2], V[i+W]-V[i-W]) with
 * out-of-range neighbours replaced by the pixel itself) */
static int target_vn(const float *depth, int rows, int cols, double fx, double fy, double cx, double cy, int i,
                     double q[3], double n[3]) {
  const int size = rows * cols;
  double V[5][3];
  const int idx[5] = {i, i - 2 < 0 ? i : i - 2, i + 2 >= size ? i : i + 2, i - cols <= 0 ? i : i - cols,
                      i + cols >= size ? i : i + cols};
  for (int k = 0; k < 5; ++k) {
    const float d = depth[idx[k]];
    if (!(d > 0.0f)) return 0;
    const int vv = idx[k] / cols, uu = idx[k] - vv * cols;
    const double dd = (double)d;
    V[k][0] = (double)(float)((((double)uu - cx) * dd) / fx);
    V[k][1] = (double)(float)((((double)vv - cy) * dd) / fy);
    V[k][2] = dd;
  }
  q[0] = V[0][0]; q[1] = V[0][1]; q[2] = V[0][2];
  const double ax = V[2][0] - V[1][0], ay = V[2][1] - V[1][1], az = V[2][2] - V[1][2];
  const double bx = V[4][0] - V[3][0], by = V[4][1] - V[3][1], bz = V[4][2] - V[3][2];
  const double c0 = ay * bz - az * by, c1 = az * bx - ax * bz, c2 = ax * by - ay * bx;
  const double nn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  if (!(nn > 0)) return 0;
  n[0] = (double)(float)(c0 / nn); n[1] = (double)(float)(c1 / nn); n[2] = (double)(float)(c2 / nn);
  return 1;
}

typedef struct {
  const float *depth;
  int rows, cols;
  double fx, fy, cx, cy;
  const double *pose;
  /* optional: the source (predicted) depth of the same view; a pixel takes
   * part only where it lies in [min_depth, max_depth] (the reference's
   * range test on both images, nonRigidICP.cu:291-292) */
  const float *src_depth;
  /* optional mesh target in camera coordinates, vertex v corresponding to
   * canonical vertex v (nonRigidICP.cu:354-376); NULL = the depth target */
  const float *mesh_v, *mesh_n;
} nr_frame;

/* Correspondence of the warped vertex p (world) with normal nw: target point
 * q and normal nt in camera coordinates, pc / ncam the warped vertex and
 * normal in camera coordinates.  0 = no correspondence. */
static int nr_associate(const nr_frame *F, const or_nr_props *pr, int64_t v, const double p[3], const double nw[3],
                        double pc[3], double ncam[3], double q[3], double nt[3]) {
  const double *pose = F->pose;
  for (int r = 0; r < 3; ++r) {
    pc[r] = pose[r] * (p[0] - pose[3]) + pose[4 + r] * (p[1] - pose[7]) + pose[8 + r] * (p[2] - pose[11]);
    ncam[r] = pose[r] * nw[0] + pose[4 + r] * nw[1] + pose[8 + r] * nw[2];
  }
  if (F->mesh_v) {
    for (int r = 0; r < 3; ++r) { q[r] = F->mesh_v[3 * v + r]; nt[r] = F->mesh_n[3 * v + r]; }
    const double n2 = nt[0] * nt[0] + nt[1] * nt[1] + nt[2] * nt[2];
    if (!(n2 > 0) || !(q[0] == q[0] && q[1] == q[1] && q[2] == q[2])) return 0;
  } else {
    if (!(pc[2] > 0)) return 0;
    const double uf = F->fx * pc[0] / pc[2] + F->cx, vf = F->fy * pc[1] / pc[2] + F->cy;
    if (!(uf > -1.0 && vf > -1.0 && uf < (double)F->cols && vf < (double)F->rows)) return 0;
    const int u = (int)floor(uf + 0.5), vv = (int)floor(vf + 0.5);
    if (u < 0 || vv < 0 || u >= F->cols || vv >= F->rows) return 0;
    const float d = F->depth[vv * F->cols + u];
    if (!(d >= (float)pr->min_depth && d <= (float)pr->max_depth)) return 0;
    if (F->src_depth) {
      const float ds = F->src_depth[vv * F->cols + u];
      if (!(ds >= (float)pr->min_depth && ds <= (float)pr->max_depth)) return 0;
    }
    if (!target_vn(F->depth, F->rows, F->cols, F->fx, F->fy, F->cx, F->cy, vv * F->cols + u, q, nt)) return 0;
  }
  const double dx = pc[0] - q[0], dy = pc[1] - q[1], dz = pc[2] - q[2];
  if (dx * dx + dy * dy + dz * dz > pr->dist_max * pr->dist_max) return 0;
  if (fabs(ncam[0] * nt[0] + ncam[1] * nt[1] + ncam[2] * nt[2]) < pr->normal_cos) return 0;
  return 1;
}

/* Gauss-Newton linearisation at the current node transforms: fills m (data
 * and ARAP blocks, rhs = -J^T r, damping on the diagonal) */
static void nr_assemble(const or_graph *g, const nr_frame *F, const or_nr_props *pr, const float *pv,
                        const float *pn, bsm *m, double *e_data_out, double *e_reg_out, int64_t *nres_out) {
  const int nn = g->nn;
  const double *pose = F->pose;
  memset(m->blk, 0, (size_t)m->start[nn] * 36 * sizeof(double));
  memset(m->rhs, 0, (size_t)nn * 6 * sizeof(double));
  double e_data = 0, e_reg = 0;
  int64_t nres = 0;
  for (int64_t v = 0; v < g->nv; ++v) {
    /* association in FP64 from the FP32 warped vertex */
    const double p[3] = {pv[3 * v], pv[3 * v + 1], pv[3 * v + 2]};
    const double nw[3] = {pn[3 * v], pn[3 * v + 1], pn[3 * v + 2]};
    double pc[3], ncam[3], q[3], nt[3];
    if (!nr_associate(F, pr, v, p, nw, pc, ncam, q, nt)) continue;
    /* world-frame target normal and point, residual n . (p' - q) */
    double nW[3], qW[3];
    for (int r = 0; r < 3; ++r) {
      nW[r] = pose[4 * r] * nt[0] + pose[4 * r + 1] * nt[1] + pose[4 * r + 2] * nt[2];
      qW[r] = pose[4 * r] * q[0] + pose[4 * r + 1] * q[1] + pose[4 * r + 2] * q[2] + pose[4 * r + 3];
    }
    const double res = nW[0] * (p[0] - qW[0]) + nW[1] * (p[1] - qW[1]) + nW[2] * (p[2] - qW[2]);
    const double gv[6] = {p[1] * nW[2] - p[2] * nW[1], p[2] * nW[0] - p[0] * nW[2], p[0] * nW[1] - p[1] * nW[0],
                          nW[0], nW[1], nW[2]};
    e_data += res * res;
    ++nres;
    int ids[NR_K];
    double w[NR_K];
    int nk;
    if (g->node_of_vertex[v] >= 0) {
      nk = 1; ids[0] = g->node_of_vertex[v]; w[0] = 1.0;
    } else {
      nk = NR_K;
      double sw = 0;
      for (int k = 0; k < NR_K; ++k) { ids[k] = g->vnn[NR_K * v + k]; w[k] = (double)g->vnn_w[NR_K * v + k]; sw += w[k]; }
      for (int k = 0; k < NR_K; ++k) w[k] /= sw;
    }
    for (int a = 0; a < nk; ++a) {
      for (int r = 0; r < 6; ++r) m->rhs[6 * ids[a] + r] -= w[a] * gv[r] * res;
      for (int b = a; b < nk; ++b) blk_add(m, ids[a], ids[b], w[a] * w[b], gv, gv);
    }
  }
  /* ARAP: r_c = w (a_c - b_c) with a = T_i g_j, b = T_j g_j;
   * d a / d xi_i = [-[a]x | I], d b / d xi_j = [-[b]x | I] */
  const double lam = pr->lambda_reg;
  for (int i = 0; i < nn; ++i) {
    const double *Ti = g->T + 16 * (size_t)i;
    for (int k = 0; k < NR_K; ++k) {
      const int j = g->nbr[NR_K * i + k];
      const double *Tj = g->T + 16 * (size_t)j;
      const float *gj = g->verts + 3 * (size_t)g->node_vid[j];
      double wij = (double)g->nbr_w[NR_K * i + k];
      const double wj = (double)g->nbr_w[NR_K * j + k];
      if (wj > wij) wij = wj;
      double A[3], B[3];
      for (int r = 0; r < 3; ++r) {
        A[r] = Ti[4 * r] * gj[0] + Ti[4 * r + 1] * gj[1] + Ti[4 * r + 2] * gj[2] + Ti[4 * r + 3];
        B[r] = Tj[4 * r] * gj[0] + Tj[4 * r + 1] * gj[1] + Tj[4 * r + 2] * gj[2] + Tj[4 * r + 3];
      }
      for (int c = 0; c < 3; ++c) {
        double Ji[6] = {0, 0, 0, 0, 0, 0}, Jj[6] = {0, 0, 0, 0, 0, 0};
        /* row c of -[a]x: c=0: (0, a2, -a1); c=1: (-a2, 0, a0); c=2: (a1, -a0, 0) */
        if (c == 0) { Ji[1] = A[2]; Ji[2] = -A[1]; Jj[1] = B[2]; Jj[2] = -B[1]; }
        if (c == 1) { Ji[0] = -A[2]; Ji[2] = A[0]; Jj[0] = -B[2]; Jj[2] = B[0]; }
        if (c == 2) { Ji[0] = A[1]; Ji[1] = -A[0]; Jj[0] = B[1]; Jj[1] = -B[0]; }
        Ji[3 + c] = 1.0; Jj[3 + c] = 1.0;
        for (int q2 = 0; q2 < 6; ++q2) { Ji[q2] *= wij; Jj[q2] *= -wij; }
        const double rc = wij * (A[c] - B[c]);
        e_reg += lam * rc * rc;
        for (int r = 0; r < 6; ++r) {
          m->rhs[6 * i + r] -= lam * Ji[r] * rc;
          m->rhs[6 * j + r] -= lam * Jj[r] * rc;
        }
        blk_add(m, i, i, lam, Ji, Ji);
        blk_add(m, j, j, lam, Jj, Jj);
        blk_add(m, i, j, lam, Ji, Jj);
      }
    }
  }
  for (int a = 0; a < nn; ++a) {
    double *B = m->blk + 36 * (size_t)bsm_find(m, a, a);
    for (int r = 0; r < 6; ++r) B[7 * r] += pr->damping;
  }
  *e_data_out = e_data; *e_reg_out = e_reg; *nres_out = nres;
}

static void nr_update(or_graph *g, const double *x) {
  for (int i = 0; i < g->nn; ++i) {
    double E[16], Tn[16];
    or_se3_exp(x + 6 * (size_t)i, E);
    const double *T = g->T + 16 * (size_t)i;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c)
        Tn[4 * r + c] = E[4 * r] * T[c] + E[4 * r + 1] * T[4 + c] + E[4 * r + 2] * T[8 + c] + E[4 * r + 3] * T[12 + c];
    memcpy(g->T + 16 * (size_t)i, Tn, sizeof Tn);
  }
}

/*
 * nonRigidICP::solve (nonRigidICP.cu:557-704), the build's formulation:
 * gn_iterations Gauss-Newton steps of
 *   E = sum_v (n_t . (p'_v - q_v))^2 + lambda sum_(i, k) |w_ik (T_i g_j - T_j g_j)|^2
 * over the node transforms T_i <- exp(xi_i) T_i, the target being one depth
 * frame seen from `pose` (camera-to-world).  stats (may be NULL):
 * [0] data residuals, [1] data energy, [2] ARAP energy (at the last
 * linearisation), [3] PCG iterations of the last step.
 */
static int nonrigid_solve(or_graph *g, const nr_frame *Fp, const or_nr_props *pr, double *stats) {
  const nr_frame F = *Fp;
  bsm m;
  bsm_build(&m, g);
  float *pv = (float *)malloc(g->nv * 3 * sizeof(float)), *pn = (float *)malloc(g->nv * 3 * sizeof(float));
  double *x = (double *)malloc((size_t)g->nn * 6 * sizeof(double));
  double e_data = 0, e_reg = 0;
  int64_t nres = 0;
  int pcg_it = 0;
  for (int it = 0; it < pr->gn_iterations; ++it) {
    or_blend(g, pv, pn);
    nr_assemble(g, &F, pr, pv, pn, &m, &e_data, &e_reg, &nres);
    pcg_it = pcg(&m, m.rhs, x, pr->pcg_iterations, pr->pcg_tol);
    nr_update(g, x);
  }
  if (stats) { stats[0] = (double)nres; stats[1] = e_data; stats[2] = e_reg; stats[3] = pcg_it; }
  free(pv); free(pn); free(x);
  bsm_free(&m);
  return 0;
}

OR_EXPORT int or_nonrigid_solve(or_graph *g, const float *depth, int rows, int cols, double fx, double fy, double cx,
                                double cy, const double pose[16], const or_nr_props *pr, double *stats) {
  const nr_frame F = {depth, rows, cols, fx, fy, cx, cy, pose, NULL, NULL, NULL};
  return nonrigid_solve(g, &F, pr, stats);
}

/* nonRigidICP::solve(warpField, blender, targetImage, sourceImage, pose,
 * sensor) (nonRigidICP.cu:557-704): the depth target gated by the source
 * image's depth range at the associated pixel (nonRigidICP.cu:291-292). */
OR_EXPORT int or_nonrigid_solve_pair(or_graph *g, const float *depth, const float *src_depth, int rows, int cols,
                                     double fx, double fy, double cx, double cy, const double pose[16],
                                     const or_nr_props *pr, double *stats) {
  const nr_frame F = {depth, rows, cols, fx, fy, cx, cy, pose, src_depth, NULL, NULL};
  return nonrigid_solve(g, &F, pr, stats);
}

/* nonRigidICP::solve(warpField, blender, targetMesh, pose)
 * (nonRigidICP.cu:706-852): target vertex v (camera coordinates, with its
 * normal) corresponds to canonical vertex v (nonRigidICP.cu:370-376). */
OR_EXPORT int or_nonrigid_solve_mesh(or_graph *g, const float *mesh_v, const float *mesh_n, const double pose[16],
                                     const or_nr_props *pr, double *stats) {
  const nr_frame F = {NULL, 0, 0, 0, 0, 0, 0, pose, NULL, mesh_v, mesh_n};
  return nonrigid_solve(g, &F, pr, stats);
}

/* The linearisation at the current transforms, densified (tests, small
 * graphs): H (6N x 6N, damping included), rhs (6N), the PCG solution x and
 * the number of upper blocks. */
OR_EXPORT int or_nonrigid_system(const or_graph *g, const float *depth, int rows, int cols, double fx, double fy,
                                 double cx, double cy, const double pose[16], const or_nr_props *pr, double *H,
                                 double *rhs, double *x, double *energy) {
  const nr_frame F = {depth, rows, cols, fx, fy, cx, cy, pose, NULL, NULL, NULL};
  const int nn = g->nn, n = 6 * nn;
  bsm m;
  bsm_build(&m, g);
  float *pv = (float *)malloc(g->nv * 3 * sizeof(float)), *pn = (float *)malloc(g->nv * 3 * sizeof(float));
  or_blend(g, pv, pn);
  double e_data, e_reg;
  int64_t nres;
  nr_assemble(g, &F, pr, pv, pn, &m, &e_data, &e_reg, &nres);
  if (H) {
    memset(H, 0, (size_t)n * n * sizeof(double));
    for (int a = 0; a < nn; ++a)
      for (int k = m.start[a]; k < m.start[a + 1]; ++k) {
        const int b = m.cols[k];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) {
            H[(size_t)(6 * a + r) * n + 6 * b + c] = m.blk[36 * (size_t)k + 6 * r + c];
            H[(size_t)(6 * b + c) * n + 6 * a + r] = m.blk[36 * (size_t)k + 6 * r + c];
          }
      }
  }
  if (rhs) memcpy(rhs, m.rhs, (size_t)n * sizeof(double));
  if (x) pcg(&m, m.rhs, x, pr->pcg_iterations, pr->pcg_tol);
  if (energy) { energy[0] = e_data; energy[1] = e_reg; energy[2] = (double)nres; energy[3] = m.start[nn]; }
  free(pv); free(pn);
  const int nb = m.start[nn];
  bsm_free(&m);
  return nb;
}
