/*
 * dynamap_nonrigid.h -- C ABI of the non-rigid path: deformation graph
 * (geometry::defGraph, include/geometry/defGraph/defGraph.h:45-87),
 * dual-quaternion blending (blender::dqBlender,
 * include/blender/dqBlender/dqBlender.h:19-83) and the non-rigid
 * Gauss-Newton solver (solver::nonRigidICP,
 * include/solver/nonRigidICP/nonRigidICP.h:59-183).  Paths are relative to
 * the reference repository root.
 *
 * Poses and node transforms are row-major 4x4 doubles; a node transform maps
 * canonical coordinates to the current frame (world).
 */
#ifndef DYNAMAP_NONRIGID_H
#define DYNAMAP_NONRIGID_H
#include "dynamap.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dm_graph dm_graph;

/* Graph construction parameters.  The reference fixes NODE_NUM = 273,
 * KNN = 4 and dgw = 0.015 at compile time (include/DynaMap.h:20-25); here
 * the node count follows from the sampling radius. */
typedef struct dm_graph_props {
  float node_radius; /* grid cell edge of the node sampling, metres */
  float dgw;         /* weight scale: w = exp(-d^2 / dgw^2) (defGraph.cu:131,410) */
  int32_t max_nodes; /* capacity, 5 .. 32768; more sampled nodes is an error */
} dm_graph_props;

/* nonRigidICP solver parameters (the reference's solver::Properties,
 * gnSolver.h:47-55, plus the terms of the build's formulation). */
typedef struct dm_nonrigid_props {
  int32_t gn_iterations;  /* Gauss-Newton iterations per call */
  int32_t pcg_iterations; /* PCG iteration cap per Gauss-Newton iteration */
  double pcg_tol;         /* PCG stops when r.z <= tol^2 r0.z0 */
  double lambda_reg;      /* ARAP weight */
  double damping;         /* added to every diagonal entry of the normal equations */
  double dist_max;        /* data association distance gate, metres */
  double normal_cos;      /* data association gate on |n_src . n_tgt| */
  double min_depth, max_depth;
} dm_nonrigid_props;

/* defGraph::init + defGraphUpdateNodes (defGraph.cu:15-69, :483-545): nodes
 * sampled from the canonical mesh (per grid cell of edge node_radius, the
 * vertex closest to the cell centre), 4 nearest nodes of every node and of
 * every vertex with weights exp(-d^2/dgw^2), node transforms = identity.
 * vertices, normals: nv x 3 floats (host). */
DM_API int dm_graph_create(const float *vertices, const float *normals, int64_t nv, const dm_graph_props *props,
                           dm_graph **out);
/* defGraph::init(MeshSTD&, int ObservationNum, int mode) with the node
 * choice made by the caller (defGraph.cu:15-69, sampleDownMeshHost
 * :673-702: mode 1 takes the first NODE_NUM vertices, mode 0 random ones):
 * node k is vertex node_vertex_ids[k] (distinct, 5 <= num_nodes <= 32768);
 * neighbours and weights as dm_graph_create with weight scale dgw. */
DM_API int dm_graph_create_nodes(const float *vertices, const float *normals, int64_t nv,
                                 const int32_t *node_vertex_ids, int32_t num_nodes, float dgw, dm_graph **out);
/* defGraph::defGraphUpdateNodes (defGraph.cu:483-545) and the vertex part of
 * updateActiveNeighbourNodes (:243-302): recompute the 4 nearest nodes of
 * every node and of every vertex and their weights from the canonical
 * positions; node transforms are kept. */
DM_API int dm_graph_update_nodes(dm_graph *g);
/* defGraph::Free (defGraph.cu:86-101) */
DM_API int dm_graph_destroy(dm_graph *g);
/* defGraph::nodeNum */
DM_API int dm_graph_num_nodes(dm_graph *g, int32_t *n);
/* defGraphNode::vertexID / nIds / nWeights (defGraph.h:35-43): per node its
 * vertex index, its 4 nearest nodes (nearest first) and their weights.  Any
 * output may be NULL. */
DM_API int dm_graph_get_nodes(dm_graph *g, int32_t *vertex_ids, int32_t *nbr_ids, float *nbr_weights);
/* defGraph::visibleNodeIds / visibleNWeights (defGraph.h:71-73): the 4
 * nearest nodes of every vertex and their weights (nv x 4 each). */
DM_API int dm_graph_get_vertex_nn(dm_graph *g, int32_t *ids, float *weights);
/* Number of 6x6 blocks of the normal equations' upper triangle. */
DM_API int dm_graph_num_blocks(dm_graph *g, int32_t *nb);
/* defGraph::updateNodesDQ (defGraph.cu:704-730): node transforms (n x 16). */
DM_API int dm_graph_set_transforms(dm_graph *g, const double *T);
DM_API int dm_graph_get_transforms(dm_graph *g, double *T);
/* dqBlender::blendMesh (dqBlender.cu:187-218): the canonical mesh warped by
 * the blended node dual quaternions; out_normals may be NULL. */
DM_API int dm_blend(dm_graph *g, float *out_vertices, float *out_normals);
/* nonRigidICP::solve (nonRigidICP.cu:557-704): gn_iterations Gauss-Newton
 * steps of the point-to-plane data term against one depth frame (host,
 * rows*cols metres) seen from pose (camera-to-world) plus the ARAP term,
 * updating the node transforms.  stats (may be NULL): [0] data residuals,
 * [1] data energy and [2] ARAP energy at the last linearisation, [3] PCG
 * iterations of the last step. */
DM_API int dm_nonrigid_solve(dm_graph *g, const float *depth, const dm_sensor *sensor, const double pose[16],
                             const dm_nonrigid_props *props, double stats[4]);
/* dm_nonrigid_solve for n distinct graphs at once (no reference
 * counterpart; for callers that track several deforming objects): graph i
 * solves against depths[i] seen from poses[16 i .. 16 i + 15], all with the
 * same sensor and props.  Every solve is enqueued on its graph's own stream
 * before any is waited for, so that the solves' short kernels overlap; each
 * result equals that of dm_nonrigid_solve on the same graph and frame.
 * stats (may be NULL): 4 per graph, as dm_nonrigid_solve. */
DM_API int dm_nonrigid_solve_many(dm_graph *const *graphs, int32_t n, const float *const *depths,
                                  const dm_sensor *sensor, const double *poses, const dm_nonrigid_props *props,
                                  double *stats);
/* A fixed set of n distinct graphs solved in shared launches (no reference
 * counterpart; for callers that track several deforming objects at frame
 * rate): each kernel of the solve runs once per step for all n graphs
 * (graph = blockIdx.y, its arguments in a device table), and the whole
 * solve is one captured HIP graph replayed per call.  The batch keeps
 * pointers to the graphs: destroy it before any of them.  A graph may be in
 * several batches and solved on its own between batch calls. */
typedef struct dm_nonrigid_batch dm_nonrigid_batch;
DM_API int dm_nonrigid_batch_create(dm_graph *const *graphs, int32_t n, dm_nonrigid_batch **out);
/* One dm_nonrigid_solve per graph of the batch: graph i against depths[i]
 * seen from poses[16 i .. 16 i + 15], same sensor and props; each result
 * equals that of dm_nonrigid_solve on the same graph and frame.  stats (may
 * be NULL): 4 per graph, as dm_nonrigid_solve. */
DM_API int dm_nonrigid_batch_solve(dm_nonrigid_batch *b, const float *const *depths, const dm_sensor *sensor,
                                   const double *poses, const dm_nonrigid_props *props, double *stats);
DM_API int dm_nonrigid_batch_destroy(dm_nonrigid_batch *b);
/* nonRigidICP::init(warpField, Prop, problemSize) (nonRigidICP.cu:18-40):
 * sizes the solver's frame buffers for `observations` pixels (the
 * reference's problemSize, rows * cols) and pcg_iterations, so that later
 * solves up to that size allocate no frame buffers. */
DM_API int dm_nonrigid_reserve(dm_graph *g, int64_t observations, int32_t pcg_iterations);
/* nonRigidICP::solve(warpField, blender, targetImage, sourceImage, pose,
 * sensor) (nonRigidICP.cu:557-704): as dm_nonrigid_solve against
 * target_depth, where a pixel takes part only if source_depth (the predicted
 * image of the same view, rows*cols metres) also lies in [min_depth,
 * max_depth] there (the range test of both images, nonRigidICP.cu:291-292). */
DM_API int dm_nonrigid_solve_pair(dm_graph *g, const float *target_depth, const float *source_depth,
                                  const dm_sensor *sensor, const double pose[16], const dm_nonrigid_props *props,
                                  double stats[4]);
/* nonRigidICP::solve(warpField, blender, targetMesh, pose)
 * (nonRigidICP.cu:706-852): the target is a mesh with the canonical mesh's
 * vertex order (target vertex v corresponds to canonical vertex v,
 * nonRigidICP.cu:370-376), nv x 3 positions and unit normals in the
 * coordinates of the camera at `pose`; same gates (dist_max, normal_cos),
 * terms and stats as dm_nonrigid_solve.  Zero or NaN entries drop a vertex. */
DM_API int dm_nonrigid_solve_mesh(dm_graph *g, const float *target_vertices, const float *target_normals,
                                  const double pose[16], const dm_nonrigid_props *props, double stats[4]);

#ifdef __cplusplus
}
#endif
#endif /* DYNAMAP_NONRIGID_H */
