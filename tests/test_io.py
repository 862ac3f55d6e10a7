"""Host I/O: TUM RGB-D layout and trajectories (include/utils/
utils_Dataset.cpp:58-247, include/builder/builder.cu:222-262), ATE / RPE by
closed forms, OBJ / PLY round trips (include/geometry/mesh/mesh_IO.h)."""
import os

import numpy as np
from scipy.spatial.transform import Rotation

from dynamap_amd import configs
from dynamap_amd import io as dio


def random_poses(n, seed):
    rng = np.random.default_rng(seed)
    T = np.tile(np.eye(4), (n, 1, 1))
    T[:, :3, :3] = Rotation.from_rotvec(rng.normal(size=(n, 3))).as_matrix()
    T[:, :3, 3] = rng.normal(size=(n, 3))
    return T


def test_quaternions_match_scipy():
    for T in random_poses(50, 1):
        q = dio.rot_to_quat(T[:3, :3])
        qs = Rotation.from_matrix(T[:3, :3]).as_quat()
        qs = qs if qs[3] >= 0 else -qs
        assert np.abs(q - qs).max() < 1e-12
        assert np.abs(dio.quat_to_rot(*q) - T[:3, :3]).max() < 1e-12


def test_trajectory_round_trip(tmp_path):
    T = random_poses(20, 2)
    st = np.arange(20) * 0.033 + 1305031102.0
    p = tmp_path / "traj.txt"
    dio.write_trajectory(str(p), st, T)
    lines = p.read_text().splitlines()
    assert len(lines) == 20 and len(lines[0].split()) == 8
    s2, T2 = dio.read_trajectory(str(p))
    assert np.abs(s2 - st).max() < 6e-5
    assert np.abs(T2[:, :3, 3] - T[:, :3, 3]).max() < 6e-5
    assert np.abs(T2[:, :3, :3] - T[:, :3, :3]).max() < 1e-3


def test_ate_rpe_closed_forms():
    gt = random_poses(40, 3)
    M = random_poses(1, 4)[0]
    est = np.einsum("ij,njk->nik", M, gt)
    assert dio.ate(est, gt) < 1e-9                       # a rigid offset is aligned away
    assert abs(dio.ate(est, gt, align=False) - np.sqrt(np.mean(np.sum((est[:, :3, 3] - gt[:, :3, 3]) ** 2, 1)))) < 1e-12
    assert dio.rpe(est, gt)[0] < 1e-9 and dio.rpe(est, gt)[1] < 1e-6
    # known translational noise orthogonal to nothing: ATE without alignment = RMS noise
    rng = np.random.default_rng(5)
    noisy = gt.copy()
    d = rng.normal(size=(40, 3)) * 0.01
    noisy[:, :3, 3] += d
    assert abs(dio.ate(noisy, gt, align=False) - np.sqrt(np.mean(np.sum(d ** 2, 1)))) < 1e-12
    assert dio.ate(noisy, gt) <= dio.ate(noisy, gt, align=False) + 1e-12
    # a constant extra motion per step: RPE equals its size
    step = np.eye(4)
    step[:3, 3] = [0.003, 0.0, 0.004]
    drift = gt.copy()
    for k in range(1, 40):
        drift[k] = drift[k - 1] @ (np.linalg.inv(gt[k - 1]) @ gt[k]) @ step
    assert abs(dio.rpe(drift, gt)[0] - 0.005) < 1e-9


def test_tum_round_trip(tmp_path):
    depth, rgb, gt = configs.make_sequence("rigid-320x240-2cm", 4)
    rgb = np.stack([np.full(depth.shape[1:] + (3,), 17 * k, np.uint8) for k in range(4)]) if rgb is None else rgb
    root = str(tmp_path / "seq")
    st = dio.write_tum_sequence(root, depth, rgb, gt)
    for assoc in (True, False):
        if not assoc:
            os.remove(os.path.join(root, "associations.txt"))
        ds = dio.TumDataset(root)
        assert len(ds) == 4
        for k in range(4):
            d, c, T, t = ds.frame(k)
            assert abs(t - st[k]) < 1e-6
            assert d.dtype == np.float32 and d.shape == depth[k].shape
            assert np.abs(d - depth[k]).max() <= 0.5 / 5000 + 1e-6
            assert np.array_equal(c, rgb[k])
            assert np.abs(T - gt[k]).max() < 1e-8


def test_mesh_files_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    V = rng.normal(size=(50, 3)).astype(np.float32)
    F = rng.integers(0, 50, size=(80, 3))
    C = rng.random(size=(50, 3)).astype(np.float32)
    dio.save_ply(str(tmp_path / "m.ply"), V, F, C)
    V2, F2, C2 = dio.read_ply(str(tmp_path / "m.ply"))
    assert np.array_equal(V2, V) and np.array_equal(F2, F)
    assert np.abs(C2 / 255.0 - C).max() <= 0.5 / 255 + 1e-7
    dio.save_obj(str(tmp_path / "m.obj"), V, F, C)
    V3, F3, C3 = dio.read_obj(str(tmp_path / "m.obj"))
    assert np.abs(V3 - V).max() < 1e-5 and np.array_equal(F3, F) and np.abs(C3 - C).max() < 1e-5
    # normals of a flat square: the plane normal
    Vq = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    Fq = np.array([[0, 1, 2], [0, 2, 3]])
    assert np.allclose(dio.vertex_normals(Vq, Fq), [[0, 0, 1]] * 4)
