"""GPU bilateral filter and Gaussian pyramid against the oracle (FP32
transcription; the exponentials of the GPU and host libraries may differ in
the last bit, so within 2e-6 relative), and a tracked session with the
filter on.  Reference: include/image/pyramid/pyramid.cu:200-282."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frame():
    depth, _, _ = configs.make_sequence("rigid-640x480-8mm", 1)
    return depth[0]


@pytest.mark.parametrize("radius", [0, 2, 5])
def test_bilateral_matches_oracle(frame, radius):
    g = dm.bilateral_filter(frame, radius, 0.03, 2.5)
    o = O.bilateral(frame, radius, 0.03, 2.5, bits=32)
    assert np.array_equal(g == 0, o == 0)
    assert np.abs(g - o).max() <= 2e-6 * np.abs(o).max()


def test_pyramid_matches_oracle(frame):
    levels = dm.depth_pyramid(frame, 0.03, 3)
    assert np.array_equal(levels[0], frame)
    prev = frame
    for l in (1, 2):
        o = O.pyr_down(prev, 0.03, bits=32)
        assert levels[l].shape == o.shape
        assert np.abs(levels[l] - o).max() <= 2e-6 * np.abs(o).max()
        prev = levels[l]


def test_session_with_filter_tracks():
    name = "rigid-640x480-8mm"
    w = configs.WORKLOADS[name]
    depth, rgb, gt = configs.make_sequence(name, 12)
    props = dict(w["tsdf"], num_buckets=1 << 18, num_blocks=1 << 16)
    s = dm.Session(props, w["icp"], w["sensor"], gt[0])
    s.set_depth_filter(3, 0.03, 2.0)
    poses = np.stack([s.process_frame(depth[f], rgb[f]) for f in range(12)])
    err = np.linalg.norm(poses[:, :3, 3] - gt[:, :3, 3], axis=1)
    print("filtered tracking error mm", 1000 * err)
    assert err.max() < 5e-3
    s.set_depth_filter(0, 0.0, 0.0)
    s.close()
