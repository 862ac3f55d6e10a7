"""GPU parity of the live-frame maps and the ICP tracker against the oracle.

Reference: include/image/sensor/rgbdImage.cu:22-49,126-167 (vertex and normal
maps), include/solver/gnSolver/gnSolver.cu:25-196 (normal equations),
:317-438 (multi-level solve).
"""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O

pytestmark = pytest.mark.gpu

W = configs.WORKLOADS["rigid-640x480-8mm"]
SENS = W["sensor"]
TSDF = dict(W["tsdf"], num_buckets=1 << 18, num_blocks=1 << 16)
ICP = W["icp"]


@pytest.fixture(scope="module")
def data():
    depth, rgb, poses = configs.make_sequence("rigid-640x480-8mm", 4)
    # model maps: FP32 oracle raycast of frame 0 integrated at its true pose
    ov = O.OracleVolume(TSDF, 32)
    ov.integrate_scan(depth[0], rgb[0], SENS, poses[0].astype(np.float32).astype(np.float64))
    d, v, n, c = ov.raycast(poses[0].astype(np.float32).astype(np.float64), SENS)
    model = O.model_maps(v, n, O.intensity(c))
    return depth, rgb, poses, model


@pytest.fixture(scope="module")
def sess(data):
    depth, rgb, poses, model = data
    s = dm.Session(TSDF, ICP, SENS, poses[0])
    for l in range(3):
        s.set_model_maps(l, *model[l])
    return s


def test_source_maps_bit_exact(sess, data):
    depth, rgb, poses, model = data
    ref = O.source_maps(depth[1], rgb[1], SENS, 32)
    for l in range(3):
        v, n, it = sess.source_maps(depth[1], rgb[1], l)
        np.testing.assert_array_equal(v, ref[l][0])
        np.testing.assert_array_equal(np.isnan(n), np.isnan(ref[l][1]))
        ok = ~np.isnan(n)
        np.testing.assert_array_equal(n[ok], ref[l][1][ok])
        np.testing.assert_array_equal(it, ref[l][2])


@pytest.mark.parametrize("variant", ["shuffle", "mfma"])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_icp_system(sess, data, level, variant):
    """The linearisation of both kernels (shuffle-tree reduction, FP32
    matrix-core F^T F) against the oracle's A, b and counts."""
    depth, rgb, poses, model = data
    src = O.source_maps(depth[1], rgb[1], SENS, 32)
    sess.source_maps(depth[1], rgb[1], 0)
    rng = np.random.default_rng(level)
    xi = rng.normal(size=6) * np.array([2e-3, 2e-3, 2e-3, 5e-3, 5e-3, 5e-3])
    L = O.load(64)
    T = np.zeros(16); L.or_se3_exp(O.P(xi), O.P(T)); T = T.reshape(4, 4)
    sess.set_icp_reduce(variant)
    try:
        A, b, s = sess.icp_system(level, T)
    finally:
        sess.set_icp_reduce("shuffle")
    Ao, bo, so = O.icp_system(src, model, SENS, level, T, ICP, 32)
    assert s[0] == so[0] and s[2] == so[2] and s[0] > 1000
    scale = np.abs(Ao).max()
    np.testing.assert_allclose(A, Ao, rtol=0, atol=2e-6 * scale)
    np.testing.assert_allclose(b, bo, rtol=0, atol=2e-6 * np.abs(bo).max() + 1e-9)


@pytest.mark.parametrize("host", [False, True])
def test_icp_track_vs_fp64(sess, data, host):
    depth, rgb, poses, model = data
    sess.source_maps(depth[2], rgb[2], 0)
    T = sess.icp_track(solve_on_host=host)
    src = O.source_maps(depth[2], rgb[2], SENS, 64)
    To = O.icp_track(src, model, SENS, ICP, 64)
    E = np.linalg.inv(To) @ T
    ang = np.arccos(np.clip((np.trace(E[:3, :3]) - 1) / 2, -1, 1))
    print("vs fp64: trans", np.linalg.norm(E[:3, 3]), "rot", ang)
    assert np.linalg.norm(E[:3, 3]) < 1e-6 and ang < 1e-6
    gt = np.linalg.inv(poses[0]) @ poses[2]
    Eg = np.linalg.inv(gt) @ T
    assert np.linalg.norm(Eg[:3, 3]) < 3e-3
