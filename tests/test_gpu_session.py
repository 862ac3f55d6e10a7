"""End-to-end per-frame pipeline (builder::buildModelLS,
include/builder/builder.cu:70-139): GPU trajectory vs the FP64 oracle pipeline."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O

pytestmark = pytest.mark.gpu


def ate(A, B):
    return float(np.sqrt(np.mean(np.sum((A[:, :3, 3] - B[:, :3, 3]) ** 2, -1))))


@pytest.mark.parametrize("host", [False, True])
def test_trajectory_320_vs_fp64(host):
    name = "rigid-320x240-2cm"
    W = configs.WORKLOADS[name]
    n = 8
    depth, rgb, gt = configs.make_sequence(name, n)
    s = dm.Session(W["tsdf"], W["icp"], W["sensor"], gt[0])
    ora = O.OraclePipeline(W["tsdf"], W["icp"], W["sensor"], gt[0], bits=64)
    G = np.stack([s.process_frame(depth[f], rgb[f] if rgb is not None else None, solve_on_host=host)
                  for f in range(n)])
    R = np.stack([ora.process_frame(depth[f], rgb[f] if rgb is not None else None) for f in range(n)])
    print("ATE gpu-vs-fp64 (m)", ate(G, R), "gpu-vs-gt", ate(G, gt), "fp64-vs-gt", ate(R, gt))
    assert ate(G, R) < 1e-3
    assert ate(G, gt) < 5e-3
