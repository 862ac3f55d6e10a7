"""ctypes bindings of libdynamap_hip.so (include/dynamap.h, include/dynamap_session.h)
and of the synthetic-data library libdm_synth.so.

The HIP library is required: loading fails loudly if it has not been built
(`python -m dynamap_amd.build` or `__graft_entry__.build()`); there is no
CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libdynamap_hip.so")
SYNTH_LIB_PATH = os.path.join(LIB_DIR, "libdm_synth.so")

DM_OK, DM_E_INVALID, DM_E_HIP, DM_E_CAPACITY, DM_E_STATE = 0, -1, -2, -3, -4
DM_RUN_BATCHED, DM_RUN_PER_FRAME_HOST, DM_RUN_PER_FRAME_DEV = 0, 1, 2


class DmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dynamap error {code}: {msg}")
        self.code = code


class dm_sensor(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("depth_scaler", C.c_float)]


class dm_tsdf_props(C.Structure):
    _fields_ = [("voxel_size", C.c_float), ("num_buckets", C.c_int32), ("bucket_size", C.c_int32),
                ("num_blocks", C.c_int32), ("block_size", C.c_int32), ("max_sdf_weight", C.c_int32),
                ("truncation_distance", C.c_float), ("max_sensor_depth", C.c_float),
                ("min_sensor_depth", C.c_float)]


class dm_icp_props(C.Structure):
    _fields_ = [("lvl_iterations", C.c_int32 * 3), ("min_depth", C.c_float), ("max_depth", C.c_float),
                ("dist_max", C.c_float), ("photo_weight", C.c_float), ("regularization", C.c_float),
                ("max_increment", C.c_float), ("min_increment", C.c_float)]


class dm_server_config(C.Structure):
    _fields_ = [("max_sessions", C.c_int32), ("max_rows", C.c_uint32), ("max_cols", C.c_uint32),
                ("hbm_budget_bytes", C.c_uint64), ("use_graph", C.c_int32)]


_vp = C.c_void_p
_P = C.POINTER

# name -> (restype, argtypes); every function declared in include/*.h
class dm_graph_props(C.Structure):
    _fields_ = [("node_radius", C.c_float), ("dgw", C.c_float), ("max_nodes", C.c_int32)]


class dm_nonrigid_props(C.Structure):
    _fields_ = [("gn_iterations", C.c_int32), ("pcg_iterations", C.c_int32), ("pcg_tol", C.c_double),
                ("lambda_reg", C.c_double), ("damping", C.c_double), ("dist_max", C.c_double),
                ("normal_cos", C.c_double), ("min_depth", C.c_double), ("max_depth", C.c_double)]


PROTOTYPES = {
    "dm_last_error": (C.c_char_p, []),
    "dm_device_count": (C.c_int, [_P(C.c_int32)]),
    "dm_volume_create": (C.c_int, [_P(dm_tsdf_props), C.c_uint32, C.c_uint32, _P(_vp)]),
    "dm_volume_destroy": (C.c_int, [_vp]),
    "dm_volume_integrate": (C.c_int, [_vp, _vp, _vp, _P(dm_sensor), _vp]),
    "dm_volume_allocate": (C.c_int, [_vp, _vp, _P(dm_sensor), _vp]),
    "dm_volume_raycast": (C.c_int, [_vp, _vp, _P(dm_sensor), _vp, _vp, _vp, _vp]),
    "dm_volume_num_allocated": (C.c_int, [_vp, _P(C.c_int32)]),
    "dm_volume_num_entries": (C.c_int, [_vp, _P(C.c_int64)]),
    "dm_volume_get_entries": (C.c_int, [_vp, _vp]),
    "dm_volume_get_blocks": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp, _P(C.c_int32)]),
    "dm_volume_counters": (C.c_int, [_vp, _vp]),
    "dm_volume_delete_blocks": (C.c_int, [_vp, _vp, C.c_int64, _P(C.c_int64)]),
    "dm_volume_recycle": (C.c_int, [_vp, _vp, C.c_double, _P(C.c_int64)]),
    "dm_volume_snapshot": (C.c_int, [_vp, _vp, C.c_uint64, _P(C.c_uint64)]),
    "dm_volume_restore": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dm_mc_tri_table": (C.c_int, [_vp]),
    "dm_depth_bilateral": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_int32, C.c_float, C.c_float]),
    "dm_depth_pyramid": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_int32, _vp]),
    "dm_graph_create": (C.c_int, [_vp, _vp, C.c_int64, _P(dm_graph_props), _P(_vp)]),
    "dm_graph_destroy": (C.c_int, [_vp]),
    "dm_graph_num_nodes": (C.c_int, [_vp, _P(C.c_int32)]),
    "dm_graph_get_nodes": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dm_graph_get_vertex_nn": (C.c_int, [_vp, _vp, _vp]),
    "dm_graph_num_blocks": (C.c_int, [_vp, _P(C.c_int32)]),
    "dm_graph_set_transforms": (C.c_int, [_vp, _vp]),
    "dm_graph_get_transforms": (C.c_int, [_vp, _vp]),
    "dm_blend": (C.c_int, [_vp, _vp, _vp]),
    "dm_nonrigid_solve": (C.c_int, [_vp, _vp, _P(dm_sensor), _vp, _P(dm_nonrigid_props), _vp]),
    "dm_nonrigid_reserve": (C.c_int, [_vp, C.c_int64, C.c_int32]),
    "dm_nonrigid_solve_pair": (C.c_int, [_vp, _vp, _vp, _P(dm_sensor), _vp, _P(dm_nonrigid_props), _vp]),
    "dm_nonrigid_solve_many": (C.c_int, [_vp, C.c_int32, _vp, _P(dm_sensor), _vp, _P(dm_nonrigid_props), _vp]),
    "dm_nonrigid_batch_create": (C.c_int, [_vp, C.c_int32, _P(_vp)]),
    "dm_nonrigid_batch_solve": (C.c_int, [_vp, _vp, _P(dm_sensor), _vp, _P(dm_nonrigid_props), _vp]),
    "dm_nonrigid_batch_destroy": (C.c_int, [_vp]),
    "dm_nonrigid_solve_mesh": (C.c_int, [_vp, _vp, _vp, _vp, _P(dm_nonrigid_props), _vp]),
    "dm_graph_create_nodes": (C.c_int, [_vp, _vp, C.c_int64, _vp, C.c_int32, C.c_float, _P(_vp)]),
    "dm_graph_update_nodes": (C.c_int, [_vp]),
    "dm_mesh_extract": (C.c_int, [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _P(C.c_int64)]),
    "dm_mesh_extract_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _P(C.c_int64)]),
    "dm_session_create": (C.c_int, [_P(dm_tsdf_props), _P(dm_icp_props), _P(dm_sensor), _vp, _P(_vp)]),
    "dm_session_destroy": (C.c_int, [_vp]),
    "dm_session_volume": (_vp, [_vp]),
    "dm_session_process_frame": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp]),
    "dm_session_process_frame_dev": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp]),
    "dm_session_get_pose": (C.c_int, [_vp, _vp]),
    "dm_session_set_depth_filter": (C.c_int, [_vp, C.c_int32, C.c_float, C.c_float]),
    "dm_session_source_maps": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp, _vp, _vp]),
    "dm_session_set_model_maps": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp]),
    "dm_session_icp_system": (C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp]),
    "dm_session_set_icp_reduce": (C.c_int, [_vp, C.c_int32]),
    "dm_session_icp_track": (C.c_int, [_vp, C.c_int32, _vp]),
    "dm_session_icp_increment": (C.c_int, [_vp, _vp]),
    "dm_server_create": (C.c_int, [_P(dm_server_config), _P(_vp)]),
    "dm_server_destroy": (C.c_int, [_vp]),
    "dm_server_add_trace": (C.c_int, [_vp, _P(dm_tsdf_props), _P(dm_icp_props), _P(dm_sensor), _vp, C.c_int32,
                                      _vp, _vp, _P(C.c_int32)]),
    "dm_server_trace_bytes": (C.c_int, [_vp, C.c_int32, _P(C.c_uint64)]),
    "dm_server_run": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp, _P(C.c_double),
                                _P(C.c_int64)]),
    "dm_server_reset": (C.c_int, [_vp]),
    "dm_server_autotune": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp, _vp]),
    "dm_server_set_shapes": (C.c_int, [_vp, _vp]),
    "dm_server_profile": (C.c_int, [_vp, C.c_int32, C.c_int32, _vp, _vp]),
}

_lib = None
_synth = None


def lib():
    """The HIP library (loaded once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError(f"{HIP_LIB_PATH} is missing: build it with `python -m dynamap_amd.build`")
        L = C.CDLL(HIP_LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != DM_OK:
        msg = lib().dm_last_error()
        raise DmError(code, msg.decode() if msg else "")
    return code


def synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB_PATH):
            raise RuntimeError(f"{SYNTH_LIB_PATH} is missing: build it with `python -m dynamap_amd.build`")
        S = C.CDLL(SYNTH_LIB_PATH)
        S.dms_scene_size.restype = C.c_int
        S.dms_scene_init.restype = C.c_int
        S.dms_scene_init.argtypes = [_vp, C.c_int, C.c_uint64]
        S.dms_render.restype = None
        S.dms_render.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_uint64, C.c_int, _vp, _vp]
        S.dms_trajectory.restype = None
        S.dms_trajectory.argtypes = [_vp, C.c_int, C.c_uint64, _vp]
        S.dms_cylinder_warp.restype = None
        S.dms_cylinder_warp.argtypes = [C.c_double, C.c_int64, _vp, _vp]
        S.dms_cylinder_mesh.restype = None
        S.dms_cylinder_mesh.argtypes = [C.c_int, C.c_int, _vp, _vp]
        S.dms_cylinder_render.restype = None
        S.dms_cylinder_render.argtypes = [C.c_double, _vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_uint64, C.c_int, _vp]
        S.dms_cylinder_trajectory.restype = None
        S.dms_cylinder_trajectory.argtypes = [C.c_int, C.c_uint64, _vp]
        _synth = S
    return _synth


def ptr(a):
    """Data pointer of a numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)
