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

    def counters(self):
        out = np.zeros(5, np.int64)
        N.check(self._lib.dm_volume_counters(self._h, N.ptr(out)))
        return dict(allocated=int(out[0]), dropped_bucket=int(out[1]), dropped_pool=int(out[2]),
                    dropped_range=int(out[3]), dropped_keyset=int(out[4]))


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

    STAGES = ("source_maps", "icp_reduce_l0", "icp_solve", "alloc_all", "integrate", "raycast", "alloc_rays")

    def profile(self, frame, reps=5):
        ms = np.zeros(7)
        work = np.zeros(6, np.int64)
        N.check(self._lib.dm_server_profile(self._h, frame, reps, N.ptr(ms), N.ptr(work)))
        return dict(zip(self.STAGES, ms.tolist())), dict(
            pixels=int(work[0]), visible_blocks=int(work[1]), block_visits=int(work[2]), sessions=int(work[4]),
            allocated_blocks=int(work[5]))
