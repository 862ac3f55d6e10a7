"""Workload definitions (sensor, volume and tracking parameters) and the
seeded synthetic data behind them.

Seeds are 20261202 + 1000 * workload + session.
  rigid-320x240-2cm : 320x240, fx=fy=262.5, 2 cm voxels, 2^16 buckets, trunc 6 cm,
                      depth 0.3..5 m, ICP 3 levels 10/5/4 (fine..coarse), depth only
  rigid-640x480-8mm : 640x480, fx=fy=525 (the reference's default sensor,
                      src/dynaMap.cpp:124-127), 8 mm voxels, 2^20 buckets, trunc 3 cm,
                      geometric + photometric ICP
  serve-16          : 8 x rigid-320x240-2cm + 8 x rigid-640x480-8mm, 300 frames each,
                      Poisson arrivals at 30 Hz per session
  large-640x480-4mm : 8 x 8 x 3 m multi-room (12 boxes, 20 spheres), 4 mm voxels,
                      2^22 buckets, 2^21 blocks, trunc 2 cm, geometric + photometric ICP
"""
import numpy as np

from . import _native as N

SEED_BASE = 20261202

SENSOR_320 = dict(rows=240, cols=320, fx=262.5, fy=262.5, cx=159.5, cy=119.5)
SENSOR_640 = dict(rows=480, cols=640, fx=525.0, fy=525.0, cx=319.5, cy=239.5)

# bucket_size 10 and max weight 64 are the reference defaults
# (src/dynaMap.cpp:202-205, config/tsdfParams.yaml:2-11)
TSDF_2CM = dict(voxel_size=0.02, num_buckets=1 << 16, bucket_size=10, num_blocks=1 << 15, block_size=8,
                max_sdf_weight=64, truncation_distance=0.06, max_sensor_depth=5.0, min_sensor_depth=0.3)
TSDF_8MM = dict(voxel_size=0.008, num_buckets=1 << 20, bucket_size=10, num_blocks=1 << 18, block_size=8,
                max_sdf_weight=64, truncation_distance=0.03, max_sensor_depth=5.0, min_sensor_depth=0.3)
TSDF_4MM = dict(voxel_size=0.004, num_buckets=1 << 22, bucket_size=10, num_blocks=1 << 21, block_size=8,
                max_sdf_weight=64, truncation_distance=0.02, max_sensor_depth=5.0, min_sensor_depth=0.3)
ICP_DEPTH = dict(lvl_iterations=(10, 5, 4), min_depth=0.3, max_depth=5.0, dist_max=0.1, photo_weight=0.0,
                 regularization=0.02, max_increment=0.3, min_increment=1e-3)
ICP_RGBD = dict(ICP_DEPTH, photo_weight=1e-3)


def sensor(d):
    s = N.dm_sensor()
    s.rows, s.cols = d["rows"], d["cols"]
    s.fx, s.fy, s.cx, s.cy = d["fx"], d["fy"], d["cx"], d["cy"]
    s.depth_scaler = 1.0
    return s


def tsdf_props(d):
    p = N.dm_tsdf_props()
    p.block_size = 8  # tsdfProperties::blockSize; 8 is the only supported edge
    for k, v in d.items():
        setattr(p, k, v)
    return p


def icp_props(d):
    p = N.dm_icp_props()
    for k, v in d.items():
        if k == "lvl_iterations":
            p.lvl_iterations[:] = list(v)
        else:
            setattr(p, k, v)
    return p


class Scene:
    """Seeded analytic scene (kind 1: 4x3x3 m room + 5 spheres; kind 5:
    8x8x3 m multi-room with 12 boxes and 20 spheres)."""

    def __init__(self, kind, seed):
        S = N.synth()
        self._buf = (np.zeros(S.dms_scene_size(), np.uint8))
        rc = S.dms_scene_init(N.ptr(self._buf), kind, seed)
        if rc != 0:
            raise RuntimeError(f"scene generation failed ({rc})")
        self.kind, self.seed = kind, seed

    def trajectory(self, nframes, seed):
        poses = np.zeros((nframes, 4, 4), np.float64)
        N.synth().dms_trajectory(N.ptr(self._buf), nframes, seed, N.ptr(poses))
        return poses

    def render(self, pose, sens, frame_seed, noise=True, rgb=True):
        rows, cols = sens["rows"], sens["cols"]
        depth = np.zeros((rows, cols), np.float32)
        col = np.zeros((rows, cols, 3), np.uint8) if rgb else None
        pose = np.ascontiguousarray(pose, np.float64)
        N.synth().dms_render(N.ptr(self._buf), N.ptr(pose), rows, cols, sens["fx"], sens["fy"], sens["cx"],
                             sens["cy"], frame_seed, 1 if noise else 0, N.ptr(depth), N.ptr(col))
        return depth, col


WORKLOADS = {
    "rigid-320x240-2cm": dict(sensor=SENSOR_320, tsdf=TSDF_2CM, icp=ICP_DEPTH, scene=1, rgb=False, wid=1),
    "rigid-640x480-8mm": dict(sensor=SENSOR_640, tsdf=TSDF_8MM, icp=ICP_RGBD, scene=1, rgb=True, wid=2),
    "large-640x480-4mm": dict(sensor=SENSOR_640, tsdf=TSDF_4MM, icp=ICP_RGBD, scene=5, rgb=True, wid=5),
}


def make_sequence(name, nframes, session=0, noise=True):
    """Frames of one session: (depth[n,r,c], rgb[n,r,c,3] or None, gt poses[n,4,4])."""
    w = WORKLOADS[name]
    seed = SEED_BASE + 1000 * w["wid"] + session
    sc = Scene(w["scene"], seed)
    poses = sc.trajectory(nframes, seed)
    s = w["sensor"]
    depth = np.zeros((nframes, s["rows"], s["cols"]), np.float32)
    rgb = np.zeros((nframes, s["rows"], s["cols"], 3), np.uint8)
    for f in range(nframes):
        d, c = sc.render(poses[f], s, seed * 7919 + f, noise=noise, rgb=True)
        depth[f] = d
        rgb[f] = c
    return depth, (rgb if w["rgb"] else None), poses
