"""Pins the oracle's tracking pieces.

- 6x6 LDL^T (replacing Eigen's ldlt, include/solver/gnSolver/gnSolver.cu:439-442)
  against numpy.linalg.solve; the step rule with no damping likewise.
- SE(3) exponential of v2t(xi) (gnSolver.cu:443-455) against scipy.linalg.expm.
- Point-to-plane projective ICP (gnSolver.cu:25-196, 317-438) recovers a
  known relative pose on a noise-free three-plane scene, and the residuals
  vanish at the true pose.
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.spatial.transform import Rotation

import oracle_lib as O


@pytest.mark.parametrize("bits", [32, 64])
def test_ldlt_vs_numpy(bits):
    L = O.load(bits)
    rng = np.random.default_rng(3)
    for k in range(50):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 0.1 * np.eye(6)
        b = rng.normal(size=6)
        x = np.zeros(6)
        assert L.or_ldlt6(O.P(np.ascontiguousarray(A)), O.P(b), O.P(x)) == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
        xi = np.zeros(6)
        L.or_icp_step(O.P(np.ascontiguousarray(A)), O.P(b), 1, 0.0, O.P(xi))
        np.testing.assert_allclose(xi, np.linalg.solve(A, -b), rtol=1e-9, atol=1e-12)


def test_ldlt_rejects_indefinite():
    L = O.load(64)
    A = np.eye(6)
    A[2, 2] = -1
    x = np.zeros(6)
    assert L.or_ldlt6(O.P(A), O.P(np.ones(6)), O.P(x)) != 0


def test_se3_exp_vs_expm():
    L = O.load(64)
    rng = np.random.default_rng(4)
    for scale in (1e-9, 1e-6, 1e-3, 0.1, 1.0, 2.5):
        xi = rng.normal(size=6) * scale
        T = np.zeros(16)
        L.or_se3_exp(O.P(xi), O.P(T))
        M = np.array([[0, -xi[2], xi[1], xi[3]], [xi[2], 0, -xi[0], xi[4]], [-xi[1], xi[0], 0, xi[5]], [0, 0, 0, 0]])
        np.testing.assert_allclose(T.reshape(4, 4), sl.expm(M), atol=1e-13)


SENS = dict(rows=120, cols=160, fx=140.0, fy=140.0, cx=79.5, cy=59.5)
# planes n.x = c in world coordinates (+y is down): floor, back wall, two side walls
PLANES = [(np.array([0.0, 1.0, 0.0]), 1.0), (np.array([0.0, 0.0, 1.0]), 2.5),
          (np.array([1.0, 0.0, 0.0]), -1.3), (np.array([1.0, 0.0, 0.0]), 1.6)]


def render_planes(T):
    v, u = np.mgrid[0:SENS["rows"], 0:SENS["cols"]]
    dc = np.stack([(u - SENS["cx"]) / SENS["fx"], (v - SENS["cy"]) / SENS["fy"], np.ones_like(u, float)], -1)
    d = dc @ T[:3, :3].T
    o = T[:3, 3]
    best = np.full(u.shape, np.inf)
    nrm = np.zeros(u.shape + (3,))
    for n, c in PLANES:
        den = d @ n
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (c - o @ n) / den
        ok = (t > 0) & (t < best)
        best = np.where(ok, t, best)
        nrm[ok] = n
    depth = best.astype(np.float32)
    vert = (dc * best[..., None]).astype(np.float32)
    ncam = (nrm @ T[:3, :3]).astype(np.float32)  # world normal into camera coordinates
    return depth, vert, ncam


def pose(rv, t):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(rv).as_matrix()
    T[:3, 3] = t
    return T


@pytest.mark.parametrize("bits", [32, 64])
def test_icp_recovers_known_pose(bits):
    A = pose([-0.35, 0.2, 0.0], [0.1, -0.1, 0.0])
    B = pose([-0.33, 0.215, 0.01], [0.118, -0.105, 0.012])
    dA, vA, nA = render_planes(A)
    dB, _, _ = render_planes(B)
    model = O.model_maps(vA, nA, np.zeros(dA.shape, np.float32))
    src = O.source_maps(dB, None, SENS, bits)
    icp = dict(lvl_iterations=(30, 10, 10), min_depth=0.1, max_depth=10.0, dist_max=0.2, photo_weight=0.0,
               regularization=0.02, max_increment=0.3)
    T = O.icp_track(src, model, SENS, icp, bits)
    true = np.linalg.inv(A) @ B
    E = np.linalg.inv(true) @ T
    assert np.linalg.norm(E[:3, 3]) < 2e-5
    assert Rotation.from_matrix(E[:3, :3]).magnitude() < 2e-5
    # residuals vanish at the true pose, not at the identity
    _, b0, s0 = O.icp_system(src, model, SENS, 0, np.eye(4), icp, bits)
    _, b1, s1 = O.icp_system(src, model, SENS, 0, true, icp, bits)
    assert np.sqrt(s1[1] / s1[0]) < 1e-4 < np.sqrt(s0[1] / s0[0])


def test_level_stop_rule():
    """gnSolver.cu:413-415: a level ends after the step that changes the
    summed twist by at most minIncrement, the change being measured from the
    sum after the last step that did not end a level (so the step that ended
    a level also counts towards the next level's first change).  Restated
    here step by step from the oracle's own linearisation and step; the
    oracle's or_icp_track must end every level at the same iteration and
    give the same increment, and a zero threshold must run every iteration."""
    A = pose([-0.35, 0.2, 0.0], [0.1, -0.1, 0.0])
    B = pose([-0.34, 0.21, 0.005], [0.11, -0.103, 0.006])
    dA, vA, nA = render_planes(A)
    dB, _, _ = render_planes(B)
    model = O.model_maps(vA, nA, np.zeros(dA.shape, np.float32))
    src = O.source_maps(dB, None, SENS, 64)
    L = O.load(64)
    for min_inc in (0.0, 1e-3, 5e-3):
        icp = dict(lvl_iterations=(10, 5, 4), min_depth=0.1, max_depth=10.0, dist_max=0.2, photo_weight=0.0,
                   regularization=0.02, max_increment=0.3, min_increment=min_inc)
        thr = float(np.float32(min_inc))
        T = np.eye(4)
        acc, prev = np.zeros(6), np.zeros(6)
        ran = []
        for lvl in (2, 1, 0):
            n = 0
            for it in range(1, icp["lvl_iterations"][lvl] + 1):
                Am, b, _ = O.icp_system(src, model, SENS, lvl, T, icp, 64)
                xi = np.zeros(6)
                L.or_icp_step(O.P(np.ascontiguousarray(Am.ravel())), O.P(np.ascontiguousarray(b)), it,
                              icp["regularization"], O.P(xi))
                E = np.zeros(16)
                L.or_se3_exp(O.P(xi), O.P(E))
                T = E.reshape(4, 4) @ T
                acc = acc + xi
                n += 1
                if np.linalg.norm(acc - prev) <= thr:
                    break
                prev = acc.copy()
            ran.append(n)
        To = O.icp_track(src, model, SENS, icp, 64)
        print("threshold", min_inc, "iterations per level (2, 1, 0):", ran)
        if min_inc == 0.0:
            assert ran == [4, 5, 10]
        else:
            assert sum(ran) < 19
        assert np.allclose(To, T, rtol=0, atol=1e-12)
