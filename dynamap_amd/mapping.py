"""Host-side classes with the reference's method names over the C ABI.

TsdfVolume  - tsdf::tsdfVolume / tsdf::HashTable (include/geometry/tsdf/tsdf.h:37-81,
              include/dataStructure/hashTable/hashTable.h:32-141)
Session     - builder::buildModelLS + solver::gnSolver (include/builder/builder.cu:70-139,
              include/solver/gnSolver/gnSolver.cu:317-438)
Server      - the multi-session serving loop (no reference counterpart)

Arrays are numpy; poses are 4x4 row-major camera-to-world.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .configs import sensor as _sensor, tsdf_props as _tsdf, icp_props as _icp


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _u8(a):
    return None if a is None else np.ascontiguousarray(a, np.uint8)


class TsdfVolume:
    """tsdfVolume.  `props` is a dict of tsdfProperties fields
    (tsdf.h:19-35); `max_rows/max_cols` bound the frames it accepts."""

    def __init__(self, props, max_rows, max_cols):
        self._lib = N.lib()
        self.props = dict(props)
        h = C.c_void_p()
        N.check(self._lib.dm_volume_create(C.byref(_tsdf(props)), max_rows, max_cols, C.byref(h)))
        self._h = h
        self._owned = True

    @classmethod
    def _borrow(cls, handle, props):
        v = cls.__new__(cls)
        v._lib, v.props, v._h, v._owned = N.lib(), dict(props), C.c_void_p(handle), False
        return v

    def close(self):
        if getattr(self, "_h", None) and self._owned:
            self._lib.dm_volume_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # tsdfVolume::IntegrateScan (tsdf.cu:454-469)
    def IntegrateScan(self, depth, rgb, sensor, pose):
        depth, rgb = _f32(depth), _u8(rgb)
        pf = np.ascontiguousarray(pose, np.float32)
        N.check(self._lib.dm_volume_integrate(self._h, N.ptr(depth), N.ptr(rgb), C.byref(_sensor(sensor)),
                                              N.ptr(pf)))

    def AllocateFromDepth(self, depth, sensor, pose):
        depth = _f32(depth)
        pf = np.ascontiguousarray(pose, np.float32)
        N.check(self._lib.dm_volume_allocate(self._h, N.ptr(depth), C.byref(_sensor(sensor)), N.ptr(pf)))

    def Raycast(self, pose, sensor):
        """depth, vertex, normal, rgb of the model seen from `pose`."""
        r, c = sensor["rows"], sensor["cols"]
        d = np.zeros((r, c), np.float32)
        v = np.zeros((r, c, 3), np.float32)
        n = np.zeros((r, c, 3), np.float32)
        rgb = np.zeros((r, c, 3), np.uint8)
        pf = np.ascontiguousarray(pose, np.float32)
        N.check(self._lib.dm_volume_raycast(self._h, N.ptr(pf), C.byref(_sensor(sensor)), N.ptr(d), N.ptr(v),
                                            N.ptr(n), N.ptr(rgb)))
        return d, v, n, rgb

    # tsdfVolume::GenerateDepth / GenerateRgb (tsdf.cu:527-550)
    def GenerateDepth(self, pose, sensor):
        return self.Raycast(pose, sensor)[0]

    def GenerateRgb(self, pose, sensor):
        return self.Raycast(pose, sensor)[3]

    # HashTable::GetNumAllocatedBlocks (hashTable.cu:165-167)
    def GetNumAllocatedBlocks(self):
        n = C.c_int32()
        N.check(self._lib.dm_volume_num_allocated(self._h, C.byref(n)))
        return n.value

    def GetNumEntries(self):
        n = C.c_int64()
        N.check(self._lib.dm_volume_num_entries(self._h, C.byref(n)))
        return n.value

    def GetHashEntries(self):
        """(E, 4) int32: position x, y, z and pointer (-1 = free) per entry."""
        out = np.zeros((self.GetNumEntries(), 4), np.int32)
        N.check(self._lib.dm_volume_get_entries(self._h, N.ptr(out)))
        return out

    def GetHashEntry(self, i):
        e = self.GetHashEntries()[i]
        return (int(e[0]), int(e[1]), int(e[2])), int(e[3])

    def blocks(self, with_voxels=True):
        """Allocated blocks in pool order: xyz (n,3), sdf (n,512), rgb (n,512,3), weight (n,512)."""
        n = self.GetNumAllocatedBlocks()
        xyz = np.zeros((max(n, 1), 3), np.int32)
        sdf = np.zeros((max(n, 1), 512), np.float32) if with_voxels else None
        rgb = np.zeros((max(n, 1), 512, 3), np.uint8) if with_voxels else None
        w = np.zeros((max(n, 1), 512), np.uint8) if with_voxels else None
        got = C.c_int32()
        N.check(self._lib.dm_volume_get_blocks(self._h, n, N.ptr(xyz), N.ptr(sdf), N.ptr(rgb), N.ptr(w),
                                               C.byref(got)))
        if not with_voxels:
            return xyz[:n]
        return xyz[:n], sdf[:n], rgb[:n], w[:n]

    def snapshot(self):
        """The reconstruction state as bytes (hash table, blocks, flags, pose, model maps)."""
        n = C.c_uint64()
        N.check(self._lib.dm_volume_snapshot(self._h, None, 0, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        N.check(self._lib.dm_volume_snapshot(self._h, N.ptr(buf), n.value, C.byref(n)))
        return buf.tobytes()

    def restore(self, data):
        buf = np.frombuffer(data, np.uint8)
        N.check(self._lib.dm_volume_restore(self._h, N.ptr(np.ascontiguousarray(buf)), len(buf)))

    # tsdfVolume::ExtractMesh (tsdf.cu:551-560): 200 000 triangles at most
    def ExtractMesh(self, lower_corner=None, upper_corner=None, max_triangles=200000):
        return MeshExtractor(max_triangles).ExtractMesh(self, lower_corner, upper_corner)

    # HashTable::DeleteBlock (hashTable.cu:112-134) for each block of xyz (n, 3)
    def DeleteBlocks(self, xyz):
        k = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        n = C.c_int64()
        N.check(self._lib.dm_volume_delete_blocks(self._h, N.ptr(k), len(k), C.byref(n)))
        return n.value

    def DeleteBlock(self, x, y, z):
        return self.DeleteBlocks([[x, y, z]]) == 1

    def Recycle(self, center, radius):
        """Delete every allocated block whose centre lies farther than radius
        (metres) from center; returns the number deleted."""
        c = np.ascontiguousarray(center, np.float64).reshape(3)
        n = C.c_int64()
        N.check(self._lib.dm_volume_recycle(self._h, N.ptr(c), float(radius), C.byref(n)))
        return n.value

    def counters(self):
        out = np.zeros(5, np.int64)
        N.check(self._lib.dm_volume_counters(self._h, N.ptr(out)))
        return dict(allocated=int(out[0]), dropped_bucket=int(out[1]), dropped_pool=int(out[2]),
                    dropped_range=int(out[3]), dropped_keyset=int(out[4]))


class Mesh:
    """geometry::Mesh (include/geometry/mesh/mesh.h:25-66) as host arrays.

    triangles (T, 3, 3) vertex positions, colors (T, 3, 3) rgb in [0, 1],
    edge_ids (T, 3) uint64 grid-edge id of each vertex (equal id = same
    vertex).  num_triangles_ is the total the extraction found; at most
    max_triangles_ of them are held."""

    def __init__(self, triangles, colors, edge_ids, num_triangles, max_triangles):
        self.triangles = triangles
        self.colors = colors
        self.edge_ids = edge_ids
        self.num_triangles_ = int(num_triangles)
        self.max_triangles_ = int(max_triangles)

    def __len__(self):
        return len(self.triangles)

    def weld(self):
        """Indexed mesh: (vertices (V,3), colors (V,3), faces (T,3) int64),
        vertices numbered in order of first appearance."""
        flat = self.edge_ids.reshape(-1)
        uniq, first, inv = np.unique(flat, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(len(order))
        faces = rank[inv].reshape(-1, 3)
        src = first[order]
        return (self.triangles.reshape(-1, 3)[src], self.colors.reshape(-1, 3)[src], faces)

    def SaveToFile(self, filename):
        """Mesh::SaveToFile (mesh.cu:38-72): "v x y z r g b" per triangle
        vertex, then "f 3i+1 3i+2 3i+3"; returns num_triangles_."""
        n = len(self.triangles)
        with open(filename, "w") as f:
            P = self.triangles.reshape(-1, 3)
            Cc = self.colors.reshape(-1, 3)
            f.writelines(f"v {p[0]:g} {p[1]:g} {p[2]:g} {c[0]:g} {c[1]:g} {c[2]:g}\n" for p, c in zip(P, Cc))
            f.writelines(f"f {i} {i + 1} {i + 2}\n" for i in range(1, 3 * n + 1, 3))
        return self.num_triangles_

    def SavePly(self, filename):
        """Binary PLY of the welded mesh (shared vertices, uchar colours)."""
        from .io import save_ply
        V, Cv, F = self.weld()
        save_ply(filename, V, F, Cv)
        return len(F)

    def SaveWelded(self, filename):
        """OBJ with shared vertices (one "v" line per distinct edge id)."""
        V, Cv, F = self.weld()
        with open(filename, "w") as f:
            f.writelines(f"v {p[0]:g} {p[1]:g} {p[2]:g} {c[0]:g} {c[1]:g} {c[2]:g}\n" for p, c in zip(V, Cv))
            f.writelines(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in F)
        return len(F)


class MeshExtractor:
    """geometry::MeshExtractor (include/geometry/mesh/meshExtractor/meshExtractor.h:20-59):
    marching cubes over the allocated blocks of a volume on the GPU."""

    def __init__(self, max_triangles=200000):
        self._lib = N.lib()
        self.max_triangles = int(max_triangles)
        self._mesh = None

    # MeshExtractor::Init (meshExtractor.cu:11-15); the voxel size is the volume's
    def Init(self, max_triangles, voxel_size=None):
        self.max_triangles = int(max_triangles)

    def ExtractMesh(self, volume, lower_corner=None, upper_corner=None):
        """MeshExtractor::ExtractMesh (meshExtractor.cu:391-402).  Without a
        box every allocated block is meshed."""
        box = lower_corner is not None
        lo = np.ascontiguousarray(lower_corner if box else [0, 0, 0], np.float64)
        hi = np.ascontiguousarray(upper_corner if box else [0, 0, 0], np.float64)
        n = C.c_int64()
        h = volume._h
        # one pass into buffers of the extractor's capacity (untouched pages
        # of the zeroed arrays cost no memory)
        cap = max(int(self.max_triangles), 1)
        pos = np.zeros((cap, 3, 3), np.float32)
        col = np.zeros((cap, 3, 3), np.float32)
        eid = np.zeros((cap, 3), np.uint64)
        N.check(self._lib.dm_mesh_extract(h, N.ptr(lo) if box else None, N.ptr(hi) if box else None,
                                          int(self.max_triangles), N.ptr(pos), N.ptr(col), N.ptr(eid), C.byref(n)))
        m = min(n.value, self.max_triangles)
        self._mesh = Mesh(pos[:m], col[:m], eid[:m], n.value, self.max_triangles)
        return self._mesh

    def GetMesh(self):
        return self._mesh


# ---------------------------------------------------------------- non-rigid
NONRIGID_DEFAULTS = dict(gn_iterations=5, pcg_iterations=40, pcg_tol=1e-6, lambda_reg=10.0, damping=1e-3,
                         dist_max=0.05, normal_cos=0.5, min_depth=0.1, max_depth=5.0)


class DefGraph:
    """geometry::defGraph (include/geometry/defGraph/defGraph.h:45-87) on the GPU:
    nodes sampled from a canonical mesh, 4 nearest nodes per node and per
    vertex with weights exp(-d^2/dgw^2), node transforms (canonical ->
    current).  `init` = defGraph::init + defGraphUpdateNodes."""

    def __init__(self, vertices, normals, node_radius=0.025, dgw=0.05, max_nodes=4096, node_ids=None):
        """Nodes sampled per grid cell of edge node_radius, or, with
        node_ids, node k = vertex node_ids[k] (defGraph::init's own choice of
        the first or random vertices, defGraph.cu:673-702)."""
        self._lib = N.lib()
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        h = C.c_void_p()
        if node_ids is None:
            pr = N.dm_graph_props(node_radius=node_radius, dgw=dgw, max_nodes=max_nodes)
            N.check(self._lib.dm_graph_create(N.ptr(self.vertices), N.ptr(self.normals), len(self.vertices),
                                              C.byref(pr), C.byref(h)))
        else:
            ids = np.ascontiguousarray(node_ids, np.int32)
            N.check(self._lib.dm_graph_create_nodes(N.ptr(self.vertices), N.ptr(self.normals), len(self.vertices),
                                                    N.ptr(ids), len(ids), float(dgw), C.byref(h)))
        self._h = h
        n = C.c_int32()
        N.check(self._lib.dm_graph_num_nodes(h, C.byref(n)))
        self.nodeNum = n.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dm_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nodes(self):
        """(vertex ids (N,), neighbour ids (N,4), neighbour weights (N,4))."""
        vid = np.zeros(self.nodeNum, np.int32)
        nbr = np.zeros((self.nodeNum, 4), np.int32)
        w = np.zeros((self.nodeNum, 4), np.float32)
        N.check(self._lib.dm_graph_get_nodes(self._h, N.ptr(vid), N.ptr(nbr), N.ptr(w)))
        return vid, nbr, w

    def vertex_neighbours(self):
        ids = np.zeros((len(self.vertices), 4), np.int32)
        w = np.zeros((len(self.vertices), 4), np.float32)
        N.check(self._lib.dm_graph_get_vertex_nn(self._h, N.ptr(ids), N.ptr(w)))
        return ids, w

    def num_blocks(self):
        n = C.c_int32()
        N.check(self._lib.dm_graph_num_blocks(self._h, C.byref(n)))
        return n.value

    # defGraph::defGraphUpdateNodes (defGraph.cu:483-545)
    def defGraphUpdateNodes(self):
        N.check(self._lib.dm_graph_update_nodes(self._h))

    # defGraph::updateNodesDQ (defGraph.cu:704-730)
    def set_transforms(self, T):
        T = np.ascontiguousarray(T, np.float64).reshape(self.nodeNum, 16)
        N.check(self._lib.dm_graph_set_transforms(self._h, N.ptr(T)))

    def transforms(self):
        T = np.zeros((self.nodeNum, 4, 4), np.float64)
        N.check(self._lib.dm_graph_get_transforms(self._h, N.ptr(T)))
        return T


class DqBlender:
    """blender::dqBlender (include/blender/dqBlender/dqBlender.h:19-83)."""

    def __init__(self, graph=None):
        self.graph = graph

    def init(self, graph):
        self.graph = graph

    def blendMesh(self, graph=None):
        """(warped vertices (V,3), rotated normals (V,3)) of the graph's canonical mesh."""
        g = graph or self.graph
        v = np.zeros_like(g.vertices)
        n = np.zeros_like(g.normals)
        N.check(g._lib.dm_blend(g._h, N.ptr(v), N.ptr(n)))
        return v, n


class NonRigidICP:
    """solver::nonRigidICP (include/solver/nonRigidICP/nonRigidICP.h:59-183):
    Gauss-Newton over the node transforms, point-to-plane data term against a
    depth frame + ARAP regulariser, block-Jacobi PCG on the GPU."""

    def __init__(self, **props):
        self.props = dict(NONRIGID_DEFAULTS)
        self.props.update(props)

    def init(self, graph=None, **props):
        self.props.update(props)

    def reserve(self, graph, observations):
        """nonRigidICP::init's sizing (nonRigidICP.cu:18-40): frame buffers for
        `observations` pixels (rows * cols), so later solves of that size
        allocate no frame buffers."""
        N.check(graph._lib.dm_nonrigid_reserve(graph._h, int(observations), int(self.props["pcg_iterations"])))

    def _props(self, override):
        p = dict(self.props)
        p.update(override)
        return N.dm_nonrigid_props(**p)

    @staticmethod
    def _stats(st):
        return dict(residuals=int(st[0]), data_energy=float(st[1]), arap_energy=float(st[2]),
                    pcg_iterations=int(st[3]))

    def solve(self, graph, depth, sensor, pose, source_depth=None, **override):
        """Against one depth frame seen from pose; with source_depth (the
        predicted image of the same view) a pixel takes part only where both
        depths lie in range (nonRigidICP.cu:291-292)."""
        pr = self._props(override)
        depth = _f32(depth)
        pose = np.ascontiguousarray(pose, np.float64)
        st = np.zeros(4, np.float64)
        if source_depth is None:
            N.check(graph._lib.dm_nonrigid_solve(graph._h, N.ptr(depth), C.byref(_sensor(sensor)), N.ptr(pose),
                                                C.byref(pr), N.ptr(st)))
        else:
            src = _f32(source_depth)
            N.check(graph._lib.dm_nonrigid_solve_pair(graph._h, N.ptr(depth), N.ptr(src), C.byref(_sensor(sensor)),
                                                     N.ptr(pose), C.byref(pr), N.ptr(st)))
        return self._stats(st)

    def solve_many(self, graphs, depths, sensor, poses, **override):
        """One solve per graph (distinct graphs), each against its own depth
        frame and pose, enqueued together so that their kernels overlap;
        returns the stats of each."""
        pr = self._props(override)
        n = len(graphs)
        if n != len(depths) or n != len(poses):
            raise ValueError("one depth frame and one pose per graph")
        ds = [_f32(d) for d in depths]
        dptr = (C.c_void_p * n)(*[d.ctypes.data for d in ds])
        gptr = (C.c_void_p * n)(*[g._h.value for g in graphs])
        P = np.ascontiguousarray(np.stack([np.asarray(p, np.float64).reshape(16) for p in poses]))
        st = np.zeros((n, 4), np.float64)
        N.check(graphs[0]._lib.dm_nonrigid_solve_many(gptr, n, dptr, C.byref(_sensor(sensor)), N.ptr(P),
                                                     C.byref(pr), N.ptr(st)))
        return [self._stats(s) for s in st]

    def batch(self, graphs):
        """A NonRigidBatch of these distinct graphs, solved with this
        solver's properties."""
        return NonRigidBatch(self, graphs)

    def solve_mesh(self, graph, target_vertices, target_normals, pose, **override):
        """nonRigidICP::solve(warpField, blender, targetMesh, pose)
        (nonRigidICP.cu:706-852): target vertex v corresponds to canonical
        vertex v; positions and normals in the camera frame at pose."""
        pr = self._props(override)
        mv = np.ascontiguousarray(target_vertices, np.float32).reshape(-1, 3)
        mn = np.ascontiguousarray(target_normals, np.float32).reshape(-1, 3)
        if len(mv) != len(graph.vertices) or len(mn) != len(mv):
            raise ValueError("the target mesh needs one vertex and normal per canonical vertex")
        pose = np.ascontiguousarray(pose, np.float64)
        st = np.zeros(4, np.float64)
        N.check(graph._lib.dm_nonrigid_solve_mesh(graph._h, N.ptr(mv), N.ptr(mn), N.ptr(pose), C.byref(pr),
                                                 N.ptr(st)))
        return self._stats(st)


class NonRigidBatch:
    """A fixed set of distinct graphs whose solves share launches: every
    kernel of the solve runs once per step for all of them (graph =
    blockIdx.y), replayed from one captured HIP graph.  Holds references to
    its graphs; close it before closing them."""

    def __init__(self, solver, graphs):
        self.solver = solver
        self.graphs = list(graphs)
        self._lib = self.graphs[0]._lib
        n = len(self.graphs)
        gptr = (C.c_void_p * n)(*[g._h.value for g in self.graphs])
        h = C.c_void_p()
        N.check(self._lib.dm_nonrigid_batch_create(gptr, n, C.byref(h)))
        self._h = h

    def solve(self, depths, sensor, poses, **override):
        """One solve per graph, graph i against depths[i] seen from poses[i];
        returns the stats of each (as NonRigidICP.solve)."""
        pr = self.solver._props(override)
        n = len(self.graphs)
        if n != len(depths) or n != len(poses):
            raise ValueError("one depth frame and one pose per graph")
        ds = [_f32(d) for d in depths]
        dptr = (C.c_void_p * n)(*[d.ctypes.data for d in ds])
        P = np.ascontiguousarray(np.stack([np.asarray(p, np.float64).reshape(16) for p in poses]))
        st = np.zeros((n, 4), np.float64)
        N.check(self._lib.dm_nonrigid_batch_solve(self._h, dptr, C.byref(_sensor(sensor)), N.ptr(P), C.byref(pr),
                                                  N.ptr(st)))
        return [self.solver._stats(s) for s in st]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dm_nonrigid_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bilateral_filter(depth, radius=3, sigma_depth=0.03, sigma_space=2.0):
    """pyramid::bilateralFilter (pyramid.cu:200-282) on the GPU."""
    d = _f32(depth)
    out = np.zeros_like(d)
    N.check(N.lib().dm_depth_bilateral(N.ptr(d), N.ptr(out), d.shape[0], d.shape[1], int(radius),
                                       float(sigma_depth), float(sigma_space)))
    return out


def depth_pyramid(depth, sigma_depth=0.03, levels=3):
    """Gaussian depth pyramid (pyramid::downPyramidDepth's role) on the GPU:
    list of levels 0..levels-1."""
    d = _f32(depth)
    r, c = d.shape
    sizes = [(r >> l, c >> l) for l in range(levels)]
    out = np.zeros(sum(a * b for a, b in sizes), np.float32)
    N.check(N.lib().dm_depth_pyramid(N.ptr(d), r, c, float(sigma_depth), int(levels), N.ptr(out)))
    res, o = [], 0
    for a, b in sizes:
        res.append(out[o:o + a * b].reshape(a, b))
        o += a * b
    return res


def mc_tri_table():
    """The build's marching-cubes triangle table: 256 packed words."""
    out = np.zeros(256, np.uint64)
    N.check(N.lib().dm_mc_tri_table(N.ptr(out)))
    return out


class Session:
    """One reconstruction: builder::buildModelLS per frame (track, integrate, raycast)."""

    def __init__(self, tsdf, icp, sensor, init_pose):
        self._lib = N.lib()
        self.sensor = dict(sensor)
        self.tsdf = dict(tsdf)
        h = C.c_void_p()
        ip = np.ascontiguousarray(init_pose, np.float64)
        N.check(self._lib.dm_session_create(C.byref(_tsdf(tsdf)), C.byref(_icp(icp)), C.byref(_sensor(sensor)),
                                            N.ptr(ip), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dm_session_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def volume(self):
        return TsdfVolume._borrow(self._lib.dm_session_volume(self._h), self.tsdf)

    def set_depth_filter(self, radius=3, sigma_depth=0.03, sigma_space=2.0):
        """Bilateral-filter each frame before tracking (radius 0 = off)."""
        N.check(self._lib.dm_session_set_depth_filter(self._h, int(radius), float(sigma_depth), float(sigma_space)))

    def process_frame(self, depth, rgb=None, solve_on_host=False):
        depth, rgb = _f32(depth), _u8(rgb)
        pose = np.zeros((4, 4), np.float64)
        N.check(self._lib.dm_session_process_frame(self._h, N.ptr(depth), N.ptr(rgb), int(solve_on_host),
                                                   N.ptr(pose)))
        return pose

    def pose(self):
        p = np.zeros((4, 4), np.float64)
        N.check(self._lib.dm_session_get_pose(self._h, N.ptr(p)))
        return p

    def icp_increment(self):
        """gnSolver::solve's summed twist (omega, t) of the last tracking."""
        xi = np.zeros(6, np.float64)
        N.check(self._lib.dm_session_icp_increment(self._h, N.ptr(xi)))
        return xi

    def writePoseToTxt(self, filename, timestamp, pose=None):
        """builder::writePoseToTxt (builder.cu:243-262): append one TUM line
        "t tx ty tz qx qy qz qw" for the current (or the given) pose."""
        from .io import write_trajectory
        write_trajectory(filename, [timestamp], [self.pose() if pose is None else pose], append=True)

    def run_dataset(self, dataset, nframes=None, trajectory=None, solve_on_host=False):
        """builder::buildModelLS over the frames of a TUM RGB-D sequence
        (io.TumDataset, the reference's dataSet::invokeNextFrame,
        utils_Dataset.cpp:223-237), appending each frame's pose to the TUM
        trajectory file `trajectory` when given.  Returns the poses (n,4,4)."""
        n = len(dataset) if nframes is None else min(int(nframes), len(dataset))
        poses = np.zeros((n, 4, 4), np.float64)
        for k in range(n):
            depth, rgb, _, stamp = dataset.frame(k)
            poses[k] = self.process_frame(depth, rgb, solve_on_host)
            if trajectory is not None:
                self.writePoseToTxt(trajectory, stamp, poses[k])
        return poses

    def source_maps(self, depth, rgb, level):
        r, c = self.sensor["rows"] >> level, self.sensor["cols"] >> level
        v = np.zeros((r, c, 3), np.float32)
        n = np.zeros((r, c, 3), np.float32)
        it = np.zeros((r, c), np.float32)
        N.check(self._lib.dm_session_source_maps(self._h, N.ptr(_f32(depth)), N.ptr(_u8(rgb)), level, N.ptr(v),
                                                 N.ptr(n), N.ptr(it)))
        return v, n, it

    def set_model_maps(self, level, vertex, normal, intensity=None):
        N.check(self._lib.dm_session_set_model_maps(self._h, level, N.ptr(_f32(vertex)), N.ptr(_f32(normal)),
                                                    N.ptr(None if intensity is None else _f32(intensity))))

    def set_icp_reduce(self, variant):
        """ICP linearisation kernel: 'shuffle' (shuffle-tree reduction, the
        default) or 'mfma' (FP32 matrix-core F^T F)."""
        if variant not in ("shuffle", "mfma"):
            raise ValueError("variant is 'shuffle' or 'mfma'")
        N.check(self._lib.dm_session_set_icp_reduce(self._h, 1 if variant == "mfma" else 0))

    def icp_system(self, level, Tinc):
        A = np.zeros((6, 6)); b = np.zeros(6); s = np.zeros(3)
        T = np.ascontiguousarray(Tinc, np.float64)
        N.check(self._lib.dm_session_icp_system(self._h, level, N.ptr(T), N.ptr(A), N.ptr(b), N.ptr(s)))
        return A, b, s

    def icp_track(self, solve_on_host=False):
        T = np.zeros((4, 4))
        N.check(self._lib.dm_session_icp_track(self._h, int(solve_on_host), N.ptr(T)))
        return T


class Server:
    """Continuous-batching serving loop over many sessions."""

    def __init__(self, max_sessions, max_rows, max_cols, hbm_budget_bytes=1 << 40, use_graph=True):
        self._lib = N.lib()
        cfg = N.dm_server_config()
        cfg.max_sessions, cfg.max_rows, cfg.max_cols = max_sessions, max_rows, max_cols
        cfg.hbm_budget_bytes, cfg.use_graph = hbm_budget_bytes, int(use_graph)
        h = C.c_void_p()
        N.check(self._lib.dm_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.traces = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dm_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_trace(self, tsdf, icp, sensor, init_pose, depth, rgb=None):
        depth, rgb = _f32(depth), _u8(rgb)
        tid = C.c_int32()
        ip = np.ascontiguousarray(init_pose, np.float64)
        N.check(self._lib.dm_server_add_trace(self._h, C.byref(_tsdf(tsdf)), C.byref(_icp(icp)),
                                              C.byref(_sensor(sensor)), N.ptr(ip), depth.shape[0], N.ptr(depth),
                                              N.ptr(rgb), C.byref(tid)))
        self.traces.append(depth.shape[0])
        return tid.value

    def trace_bytes(self, tid):
        b = C.c_uint64()
        N.check(self._lib.dm_server_trace_bytes(self._h, tid, C.byref(b)))
        return b.value

    def run(self, mode, first_frame, nframes, arrivals=None, want_poses=False):
        nt = len(self.traces)
        comp = np.zeros((nt, nframes), np.float64)
        poses = np.zeros((nt, nframes, 4, 4), np.float64) if want_poses else None
        arr = None if arrivals is None else np.ascontiguousarray(arrivals, np.float64)
        wall = C.c_double()
        nb = C.c_int64()
        N.check(self._lib.dm_server_run(self._h, mode, first_frame, nframes, N.ptr(arr), N.ptr(comp),
                                        N.ptr(poses), C.byref(wall), C.byref(nb)))
        return dict(completions=comp, poses=poses, wall=wall.value, batches=nb.value)

    def reset(self):
        N.check(self._lib.dm_server_reset(self._h))

    STAGES = ("source_maps", "icp_l0", "icp_l0_reduce_only", "alloc_all", "integrate", "raycast", "alloc_rays")

    def autotune(self, frame=0, reps=3):
        """Pick the raycast tile shape, integration grid and ICP reduction
        (shuffle tree or matrix cores) by timing them on the sessions' state
        at `frame`; the sessions are reset afterwards."""
        choice = np.zeros(3, np.int32)
        ms = np.zeros(13)
        N.check(self._lib.dm_server_autotune(self._h, int(frame), int(reps), N.ptr(choice), N.ptr(ms)))
        N.check(self._lib.dm_server_reset(self._h))
        return dict(ray_tile=f"{int(choice[0])}x{256 // int(choice[0])}", integrate_workgroups=int(choice[1]),
                    icp_reduce="mfma" if int(choice[2]) else "shuffle",
                    raycast_ms={f"{w}x{256 // w}": float(t) for w, t in zip((8, 16, 32, 64), ms[:4])},
                    integrate_ms={str(n): float(t) for n, t in zip((1024, 2048, 4096), ms[4:7])},
                    icp_reduce_ms={"shuffle": [float(t) for t in ms[7:10]], "mfma": [float(t) for t in ms[10:13]]})

    def set_shapes(self, ray_tile_width=16, integrate_workgroups=2048, icp_reduce="shuffle"):
        """Set the serving graph's launch shapes (the autotuner's choices)."""
        if icp_reduce not in ("shuffle", "mfma"):
            raise ValueError("icp_reduce is 'shuffle' or 'mfma'")
        shapes = np.array([ray_tile_width, integrate_workgroups, 1 if icp_reduce == "mfma" else 0], np.int32)
        N.check(self._lib.dm_server_set_shapes(self._h, N.ptr(shapes)))

    def profile(self, frame, reps=5):
        ms = np.zeros(7)
        work = np.zeros(7, np.int64)
        N.check(self._lib.dm_server_profile(self._h, frame, reps, N.ptr(ms), N.ptr(work)))
        return dict(zip(self.STAGES, ms.tolist())), dict(
            pixels=int(work[0]), visible_blocks=int(work[1]), block_visits=int(work[2]), new_blocks=int(work[3]),
            sessions=int(work[4]), allocated_blocks=int(work[5]), unique_blocks=int(work[6]))
