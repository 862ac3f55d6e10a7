"""The single-session workload at its full length: 500 frames at 640x480 /
8 mm voxels with geometric + photometric ICP through the per-frame pipeline
(builder::buildModelLS, include/builder/builder.cu:70-139).  The first
frames agree with the FP64 oracle pipeline within 1 mm ATE; over the whole
sequence no increment is reset by the reference's guard
(gnSolver.cu:426-428: the pose would repeat), no key is dropped, and the trajectory stays within a
drift bound of the ground truth."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O

pytestmark = pytest.mark.gpu


def ate(A, B):
    return float(np.sqrt(np.mean(np.sum((A[:, :3, 3] - B[:, :3, 3]) ** 2, -1))))


def test_single_session_500_frames():
    name = "rigid-640x480-8mm"
    W = configs.WORKLOADS[name]
    n, n_oracle = 500, 16
    depth, rgb, gt = configs.make_sequence(name, n, session=1)
    s = dm.Session(W["tsdf"], W["icp"], W["sensor"], gt[0])
    G = np.stack([s.process_frame(depth[f], rgb[f]) for f in range(n)])
    cnt = s.volume.counters()
    # a guard reset leaves the pose of the previous frame (the camera moves
    # in every frame of the trajectory)
    fails = sum(np.array_equal(G[f], G[f - 1]) for f in range(1, n))
    err = np.linalg.norm(G[:, :3, 3] - gt[:, :3, 3], axis=1)
    print("blocks", cnt, "ICP guard resets", fails, "ATE vs ground truth (m)", ate(G, gt),
          "max translation error (m)", err.max(), flush=True)
    s.close()
    ora = O.OraclePipeline(W["tsdf"], W["icp"], W["sensor"], gt[0], bits=64)
    R = np.stack([ora.process_frame(depth[f], rgb[f]) for f in range(n_oracle)])
    print("ATE gpu-vs-fp64 over the first", n_oracle, "frames (m)", ate(G[:n_oracle], R), flush=True)
    assert np.all(np.isfinite(G))
    assert ate(G[:n_oracle], R) < 1e-3
    assert cnt["dropped_bucket"] == 0 and cnt["dropped_pool"] == 0 and cnt["dropped_keyset"] == 0
    assert fails == 0
    # drift bound: 5 mm over 500 frames of noisy depth (the per-frame motion
    # is up to 2 cm and 1 degree)
    assert err.max() < 5e-3
