"""ctypes bindings of the CPU oracle (oracle/_build/liboracle_f32.so and
liboracle_f64.so) and a frame pipeline composed from its functions.

Test infrastructure only: the product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# DM_ORACLE_BUILD: another build of the oracle (a copy with a planted mistake)
BUILD = os.environ.get("DM_ORACLE_BUILD", os.path.join(ORACLE_DIR, "_build"))
_vp = C.c_void_p
_d = C.c_double
_i = C.c_int

_PROTOS = {
    "or_real_size": (_i, []),
    "or_hash": (_i, [_i, _i, _i, _i, _i]),
    "or_pack_key": (C.c_uint64, [_i, _i, _i]),
    "or_volume_create": (_vp, [_d, _i, _i, _i, _i, _d, _d, _d]),
    "or_volume_destroy": (None, [_vp]),
    "or_volume_num_alloc": (_i, [_vp]),
    "or_volume_dropped": (None, [_vp, _vp]),
    "or_ray_blocks": (_i, [_vp, _i, _d, _i, _i, _d, _d, _d, _d, _d, _d, _vp, _i]),
    "or_allocate": (None, [_vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "or_integrate": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "or_integrate_scan": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp]),
    "or_get_block": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "or_get_table": (None, [_vp, _vp, _vp]),
    "or_dump_blocks": (None, [_vp, _vp, _vp, _vp, _vp]),
    "or_delete_blocks": (C.c_int64, [_vp, _vp, C.c_int64]),
    "or_recycle": (C.c_int64, [_vp, _vp, _d]),
    "or_raycast": (None, [_vp, _i, _i, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    "or_vertex_map": (None, [_vp, _i, _i, _d, _d, _d, _d, _vp]),
    "or_normal_map": (None, [_vp, _i, _i, _vp]),
    "or_subsample": (None, [_vp, _i, _i, _i, _vp]),
    "or_icp_system": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp, _vp, _vp, _vp, _d, _d, _d, _d,
                             _vp, _vp, _vp]),
    "or_ldlt6": (_i, [_vp, _vp, _vp]),
    "or_se3_exp": (None, [_vp, _vp]),
    "or_icp_step": (_i, [_vp, _vp, _i, _d, _vp]),
    "or_icp_track": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp, _d, _d, _d, _d, _d, _d,
                            _d, _vp]),
}

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(bits):
    """bits = 32 (FP32 transcription) or 64 (FP64 oracle)."""
    if bits not in _libs:
        path = os.path.join(BUILD, f"liboracle_f{bits}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for n, (r, a) in _PROTOS.items():
            f = getattr(L, n, None)
            if f is None:
                continue
            f.restype, f.argtypes = r, a
        assert L.or_real_size() == bits // 8
        _libs[bits] = L
    return _libs[bits]


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _intr(s):
    return s["rows"], s["cols"], float(s["fx"]), float(s["fy"]), float(s["cx"]), float(s["cy"])


class OracleVolume:
    def __init__(self, props, bits=32):
        self.L = load(bits)
        self.bits = bits
        self.props = props
        p = props
        self.h = self.L.or_volume_create(p["voxel_size"], p["num_buckets"], p["bucket_size"], p["num_blocks"],
                                         p["max_sdf_weight"], p["truncation_distance"], p["max_sensor_depth"],
                                         p["min_sensor_depth"])

    def __del__(self):
        try:
            self.L.or_volume_destroy(self.h)
        except Exception:
            pass

    def allocate(self, depth, sensor, pose):
        d = np.ascontiguousarray(depth, np.float32)
        T = np.ascontiguousarray(pose, np.float64)
        self.L.or_allocate(self.h, P(d), *_intr(sensor), P(T))

    def integrate_scan(self, depth, rgb, sensor, pose):
        d = np.ascontiguousarray(depth, np.float32)
        c = None if rgb is None else np.ascontiguousarray(rgb, np.uint8)
        T = np.ascontiguousarray(pose, np.float64)
        self.L.or_integrate_scan(self.h, P(d), P(c), *_intr(sensor), P(T))

    def num_alloc(self):
        return self.L.or_volume_num_alloc(self.h)

    def dropped(self):
        out = np.zeros(3, np.int64)
        self.L.or_volume_dropped(self.h, P(out))
        return dict(bucket=int(out[0]), pool=int(out[1]), range=int(out[2]))

    def table(self):
        B, S = self.props["num_buckets"], self.props["bucket_size"]
        keys = np.zeros(B * S, np.uint64)
        counts = np.zeros(B, np.int32)
        self.L.or_get_table(self.h, P(keys), P(counts))
        return keys, counts

    def delete_blocks(self, xyz):
        """HashTable::DeleteBlock for each block of xyz (n, 3); returns the number deleted."""
        k = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        return int(self.L.or_delete_blocks(self.h, P(k), len(k)))

    def recycle(self, center, radius):
        c = np.ascontiguousarray(center, np.float64)
        return int(self.L.or_recycle(self.h, P(c), float(radius)))

    def blocks(self):
        n = self.num_alloc()
        xyz = np.zeros((max(n, 1), 3), np.int32)
        sdf = np.zeros((max(n, 1), 512), np.float64)
        rgb = np.zeros((max(n, 1), 512, 3), np.uint8)
        w = np.zeros((max(n, 1), 512), np.uint8)
        self.L.or_dump_blocks(self.h, P(xyz), P(sdf), P(rgb), P(w))
        return xyz[:n], sdf[:n], rgb[:n], w[:n]

    def raycast(self, pose, sensor, want_rgb=True):
        r, c = sensor["rows"], sensor["cols"]
        d = np.zeros((r, c), np.float32)
        v = np.zeros((r, c, 3), np.float32)
        n = np.zeros((r, c, 3), np.float32)
        rgb = np.zeros((r, c, 3), np.uint8) if want_rgb else None
        T = np.ascontiguousarray(pose, np.float64)
        self.L.or_raycast(self.h, *_intr(sensor), P(T), P(d), P(v), P(n), P(rgb))
        return d, v, n, rgb


def pack_key(x, y, z):
    M = (1 << 21) - 1
    return np.uint64((1 << 63) | ((int(x) & M) << 42) | ((int(y) & M) << 21) | (int(z) & M))


def intensity(rgb):
    r, g, b = (rgb[..., k].astype(np.float32) for k in range(3))
    return ((np.float32(0.2126) * r + np.float32(0.7152) * g) + np.float32(0.0722) * b) * np.float32(1.0 / 255.0)


def level_sensor(sensor, l):
    s = float(1 << l)
    return dict(rows=sensor["rows"] >> l, cols=sensor["cols"] >> l, fx=sensor["fx"] / s, fy=sensor["fy"] / s,
                cx=sensor["cx"] / s, cy=sensor["cy"] / s)


def source_maps(depth, rgb, sensor, bits=32):
    """Per level: vertex (r,c,3), normal (r,c,3), intensity (r,c), as the
    reference's getPointCloudXYZ / getNormalsfromVertices on a nearest
    subsampled pyramid."""
    L = load(bits)
    out = []
    d = np.ascontiguousarray(depth, np.float32)
    inten = intensity(rgb) if rgb is not None else np.zeros(d.shape, np.float32)
    for l in range(3):
        s = level_sensor(sensor, l)
        if l > 0:
            nd = np.zeros((s["rows"], s["cols"]), np.float32)
            L.or_subsample(P(d), d.shape[0], d.shape[1], 1, P(nd))
            d = nd
            inten = np.ascontiguousarray(inten[::2, ::2][: s["rows"], : s["cols"]])
        v = np.zeros((s["rows"], s["cols"], 3), np.float32)
        n = np.zeros_like(v)
        L.or_vertex_map(P(d), *_intr(s), P(v))
        L.or_normal_map(P(v), s["rows"], s["cols"], P(n))
        out.append((v, n, np.ascontiguousarray(inten, np.float32)))
    return out


def model_maps(vertex, normal, inten):
    """Levels 1, 2 of the model maps by nearest subsampling of level 0."""
    out = [(vertex, normal, inten)]
    for l in (1, 2):
        r, c = vertex.shape[0] >> l, vertex.shape[1] >> l
        step = 1 << l
        out.append((np.ascontiguousarray(vertex[::step, ::step][:r, :c]),
                    np.ascontiguousarray(normal[::step, ::step][:r, :c]),
                    np.ascontiguousarray(inten[::step, ::step][:r, :c])))
    return out


def icp_system(src, tgt, sensor, level, Tinc, icp, bits=32):
    L = load(bits)
    s = level_sensor(sensor, level)
    (sv, sn, si), (tv, tn, ti) = src[level], tgt[level]
    A = np.zeros(36); b = np.zeros(6); sums = np.zeros(3)
    T = np.ascontiguousarray(Tinc, np.float64)
    L.or_icp_system(P(sv), P(sn), P(si), *_intr(s), P(tv), P(tn), P(ti), P(T), icp["min_depth"],
                    icp["max_depth"], icp["dist_max"], icp["photo_weight"], P(A), P(b), P(sums))
    return A.reshape(6, 6), b, sums


def icp_track(src, tgt, sensor, icp, bits=64):
    L = load(bits)
    arr = lambda xs: (C.c_void_p * 3)(*[x.ctypes.data for x in xs])
    sv = arr([x[0] for x in src]); sn = arr([x[1] for x in src]); si = arr([x[2] for x in src])
    tv = arr([x[0] for x in tgt]); tn = arr([x[1] for x in tgt]); ti = arr([x[2] for x in tgt])
    iters = np.array(icp["lvl_iterations"], np.int32)
    T = np.zeros(16)
    L.or_icp_track(sv, sn, si, tv, tn, ti, *_intr(sensor), P(iters), icp["min_depth"], icp["max_depth"],
                   icp["dist_max"], icp["photo_weight"], icp["regularization"], icp["max_increment"],
                   float(np.float32(icp.get("min_increment", 0.0))), P(T))
    return T.reshape(4, 4)


class OraclePipeline:
    """builder::buildModelLS restated on the oracle: track the frame against
    the model raycast (frame-to-model), integrate at the tracked pose, raycast
    the model at that pose."""

    def __init__(self, tsdf, icp, sensor, init_pose, bits=64):
        self.vol = OracleVolume(tsdf, bits)
        self.icp, self.sensor, self.bits = icp, sensor, bits
        self.T = np.array(init_pose, np.float64)
        self.model = None

    def process_frame(self, depth, rgb):
        if self.model is not None:
            src = source_maps(depth, rgb, self.sensor, self.bits)
            Tinc = icp_track(src, self.model, self.sensor, self.icp, self.bits)
            self.T = self.T @ Tinc
        Tf = self.T.astype(np.float32).astype(np.float64)
        self.vol.integrate_scan(depth, rgb, self.sensor, Tf if self.bits == 32 else self.T)
        d, v, n, c = self.vol.raycast(Tf if self.bits == 32 else self.T, self.sensor)
        self.model = model_maps(v, n, intensity(c) if c is not None else np.zeros(d.shape, np.float32))
        return self.T.copy()


def mc_extract(vol, box=None):
    """Oracle marching cubes: (positions (T,3,3), colours (T,3,3), edge ids (T,3))."""
    L = vol.L
    L.or_mc_extract.restype = C.c_int64
    L.or_mc_extract.argtypes = [_vp, _i, _vp, _vp, _vp, _vp, _vp, C.c_int64]
    lo = np.asarray(box[0] if box else [0, 0, 0], np.float64)
    hi = np.asarray(box[1] if box else [0, 0, 0], np.float64)
    n = L.or_mc_extract(vol.h, 1 if box else 0, P(lo), P(hi), None, None, None, 0)
    pos = np.zeros((max(n, 1), 3, 3), np.float32)
    col = np.zeros((max(n, 1), 3, 3), np.float32)
    eid = np.zeros((max(n, 1), 3), np.uint64)
    L.or_mc_extract(vol.h, 1 if box else 0, P(lo), P(hi), P(pos), P(col), P(eid), n)
    return pos[:n], col[:n], eid[:n]


def mc_table(bits=32):
    out = np.zeros(256, np.uint64)
    load(bits).or_mc_table(P(out))
    return out


def weld(eid):
    """Indexed mesh from edge ids: vertices numbered by first appearance."""
    flat = eid.reshape(-1)
    uniq, first, inv = np.unique(flat, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[inv].reshape(eid.shape), len(uniq)


def _set_blocks(self, xyz, sdf, w, rgb=None):
    """Write whole blocks in the given order (test hook): xyz (n,3) int32,
    sdf (n,512), w (n,512) uint8, rgb (n,512,3) uint8 or None."""
    L = self.L
    L.or_set_blocks.restype = _i
    L.or_set_blocks.argtypes = [_vp, _i, _vp, _vp, _vp, _vp]
    xyz = np.ascontiguousarray(xyz, np.int32)
    sdf = np.ascontiguousarray(sdf, np.float64)
    w = np.ascontiguousarray(w, np.uint8)
    rgb = None if rgb is None else np.ascontiguousarray(rgb, np.uint8)
    return L.or_set_blocks(self.h, len(xyz), P(xyz), P(sdf), P(w), P(rgb))


OracleVolume.set_blocks = _set_blocks


def _interp(self, pts):
    """Trilinear samples of the volume at world points (n, 3) (test hook of
    GetInterpolatedVoxel): (sdf (n,), colour (n, 3), observed (n,) bool)."""
    L = self.L
    L.or_interp.restype = None
    L.or_interp.argtypes = [_vp, _i, _vp, _vp, _vp, _vp]
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    sdf, col, ok = np.zeros(n), np.zeros((n, 3)), np.zeros(n, np.int32)
    L.or_interp(self.h, n, P(pts), P(sdf), P(col), P(ok))
    return sdf, col, ok.astype(bool)


OracleVolume.interp = _interp


# ---------------------------------------------------------------- non-rigid
class or_nr_props(C.Structure):
    _fields_ = [("gn_iterations", C.c_int), ("pcg_iterations", C.c_int), ("pcg_tol", C.c_double),
                ("lambda_reg", C.c_double), ("damping", C.c_double), ("dist_max", C.c_double),
                ("normal_cos", C.c_double), ("min_depth", C.c_double), ("max_depth", C.c_double)]


NR_DEFAULTS = dict(gn_iterations=5, pcg_iterations=40, pcg_tol=1e-6, lambda_reg=10.0, damping=1e-3,
                   dist_max=0.05, normal_cos=0.5, min_depth=0.1, max_depth=5.0)


def nr_props(**kw):
    p = dict(NR_DEFAULTS)
    p.update(kw)
    return or_nr_props(**p)


class OracleGraph:
    """Deformation graph + blender + non-rigid solver of the CPU oracle."""

    def __init__(self, verts, normals, node_radius=0.025, dgw=0.05, max_nodes=4096, bits=64, node_ids=None):
        L = self.L = load(bits)
        L.or_graph_create.restype = _vp
        L.or_graph_create.argtypes = [_vp, _vp, C.c_int64, _d, _d, _i, _vp]
        L.or_graph_create_nodes.restype = _vp
        L.or_graph_create_nodes.argtypes = [_vp, _vp, C.c_int64, _vp, _i, _d, _vp]
        self.verts = np.ascontiguousarray(verts, np.float32)
        self.normals = np.ascontiguousarray(normals, np.float32)
        st = np.zeros(1, np.int32)
        if node_ids is None:
            self.h = L.or_graph_create(P(self.verts), P(self.normals), len(self.verts), node_radius, dgw,
                                       max_nodes, P(st))
        else:
            ids = np.ascontiguousarray(node_ids, np.int32)
            self.h = L.or_graph_create_nodes(P(self.verts), P(self.normals), len(self.verts), P(ids), len(ids),
                                             dgw, P(st))
        self.status = int(st[0])
        L.or_graph_num_nodes.restype = _i
        L.or_graph_num_nodes.argtypes = [_vp]
        self.nn = L.or_graph_num_nodes(self.h)
        for n, a in (("or_graph_nodes", [_vp, _vp, _vp, _vp]), ("or_graph_vertex_nn", [_vp, _vp, _vp]),
                     ("or_graph_set_transforms", [_vp, _vp]), ("or_graph_get_transforms", [_vp, _vp]),
                     ("or_blend", [_vp, _vp, _vp]), ("or_graph_destroy", [_vp]),
                     ("or_dq_of_transform", [_vp, _vp]), ("or_graph_update_nodes", [_vp])):
            f = getattr(L, n)
            f.restype, f.argtypes = None, a
        L.or_nonrigid_solve.restype = _i
        L.or_nonrigid_solve.argtypes = [_vp, _vp, _i, _i, _d, _d, _d, _d, _vp, C.POINTER(or_nr_props), _vp]
        L.or_nonrigid_solve_pair.restype = _i
        L.or_nonrigid_solve_pair.argtypes = [_vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _vp, C.POINTER(or_nr_props),
                                             _vp]
        L.or_nonrigid_solve_mesh.restype = _i
        L.or_nonrigid_solve_mesh.argtypes = [_vp, _vp, _vp, _vp, C.POINTER(or_nr_props), _vp]
        L.or_nonrigid_system.restype = _i
        L.or_nonrigid_system.argtypes = [_vp, _vp, _i, _i, _d, _d, _d, _d, _vp, C.POINTER(or_nr_props), _vp, _vp,
                                         _vp, _vp]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_graph_destroy(self.h)
            self.h = None

    def nodes(self):
        vid = np.zeros(self.nn, np.int32)
        nbr = np.zeros((self.nn, 4), np.int32)
        w = np.zeros((self.nn, 4), np.float64)
        self.L.or_graph_nodes(self.h, P(vid), P(nbr), P(w))
        return vid, nbr, w

    def vertex_nn(self):
        ids = np.zeros((len(self.verts), 4), np.int32)
        w = np.zeros((len(self.verts), 4), np.float64)
        self.L.or_graph_vertex_nn(self.h, P(ids), P(w))
        return ids, w

    def set_transforms(self, T):
        T = np.ascontiguousarray(T, np.float64).reshape(self.nn, 16)
        self.L.or_graph_set_transforms(self.h, P(T))

    def transforms(self):
        T = np.zeros((self.nn, 4, 4))
        self.L.or_graph_get_transforms(self.h, P(T))
        return T

    def blend(self):
        v = np.zeros_like(self.verts)
        n = np.zeros_like(self.normals)
        self.L.or_blend(self.h, P(v), P(n))
        return v, n

    def solve(self, depth, sensor, pose, **kw):
        depth = np.ascontiguousarray(depth, np.float32)
        pose = np.ascontiguousarray(pose, np.float64)
        st = np.zeros(4)
        pr = nr_props(**kw)
        self.L.or_nonrigid_solve(self.h, P(depth), *_intr(sensor), P(pose), C.byref(pr), P(st))
        return st

    def solve_pair(self, depth, src_depth, sensor, pose, **kw):
        depth = np.ascontiguousarray(depth, np.float32)
        src = np.ascontiguousarray(src_depth, np.float32)
        pose = np.ascontiguousarray(pose, np.float64)
        st = np.zeros(4)
        pr = nr_props(**kw)
        self.L.or_nonrigid_solve_pair(self.h, P(depth), P(src), *_intr(sensor), P(pose), C.byref(pr), P(st))
        return st

    def solve_mesh(self, mesh_v, mesh_n, pose, **kw):
        mv = np.ascontiguousarray(mesh_v, np.float32)
        mn = np.ascontiguousarray(mesh_n, np.float32)
        pose = np.ascontiguousarray(pose, np.float64)
        st = np.zeros(4)
        pr = nr_props(**kw)
        self.L.or_nonrigid_solve_mesh(self.h, P(mv), P(mn), P(pose), C.byref(pr), P(st))
        return st

    def update_nodes(self):
        self.L.or_graph_update_nodes(self.h)

    def system(self, depth, sensor, pose, **kw):
        n = 6 * self.nn
        H = np.zeros((n, n))
        b = np.zeros(n)
        x = np.zeros(n)
        e = np.zeros(4)
        depth = np.ascontiguousarray(depth, np.float32)
        pose = np.ascontiguousarray(pose, np.float64)
        pr = nr_props(**kw)
        self.L.or_nonrigid_system(self.h, P(depth), *_intr(sensor), P(pose), C.byref(pr), P(H), P(b), P(x), P(e))
        return H, b, x, e


def dq_of_transform(T, bits=64):
    out = np.zeros(8)
    L = load(bits)
    L.or_dq_of_transform.restype, L.or_dq_of_transform.argtypes = None, [_vp, _vp]
    L.or_dq_of_transform(P(np.ascontiguousarray(T, np.float64)), P(out))
    return out


# ---------------------------------------------------------------- preprocessing
def bilateral(depth, radius, sigma_depth, sigma_space, bits=64):
    L = load(bits)
    L.or_bilateral.restype = None
    L.or_bilateral.argtypes = [_vp, _vp, _i, _i, _i, _d, _d]
    d = np.ascontiguousarray(depth, np.float32)
    out = np.zeros_like(d)
    L.or_bilateral(P(d), P(out), d.shape[0], d.shape[1], radius, sigma_depth, sigma_space)
    return out


def pyr_down(depth, sigma_depth, bits=64):
    L = load(bits)
    L.or_pyr_down.restype = None
    L.or_pyr_down.argtypes = [_vp, _i, _i, _d, _vp]
    d = np.ascontiguousarray(depth, np.float32)
    out = np.zeros((d.shape[0] >> 1, d.shape[1] >> 1), np.float32)
    L.or_pyr_down(P(d), d.shape[0], d.shape[1], sigma_depth, P(out))
    return out
