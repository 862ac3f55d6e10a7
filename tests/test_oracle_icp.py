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


@pytest.mark.parametrize("bits", [32, 64])
def test_planar_depth_gives_plane_normal(bits):
    """rgbdImage.cu:22-49,126-145: the vertex map of a planar depth map is
    the plane, and its normal map (cross of the horizontal and vertical
    vertex differences) is the plane normal, oriented away from the camera
    (cross(right, down) = +z on a fronto-parallel plane).  A reversed cross
    product flips it; a cx/cy or fx/fy exchange in the back-projection bends
    the vertex map off the plane (fx != fy, cx != cy here)."""
    s = dict(rows=60, cols=80, fx=70.0, fy=55.0, cx=41.5, cy=27.0)
    n = np.array([0.25, -0.35, -1.0])
    n /= np.linalg.norm(n)
    c = -1.4  # plane n.X = c in camera coordinates, facing the camera
    v, u = np.mgrid[0:s["rows"], 0:s["cols"]]
    ray = np.stack([(u - s["cx"]) / s["fx"], (v - s["cy"]) / s["fy"], np.ones(u.shape)], -1)
    depth = (c / (ray @ n)).astype(np.float32)
    lv = O.source_maps(depth, None, s, bits)[0]
    vert, nrm = lv[0], lv[1]
    # vertices lie on the plane (float rounding of the depth only)
    assert np.abs(vert.reshape(-1, 3).astype(np.float64) @ n - c).max() < 2e-6
    inner = nrm[1:-1, 2:-2].reshape(-1, 3).astype(np.float64)
    assert np.isfinite(inner).all()
    assert np.abs(inner + n).max() < 2e-5  # -n: away from the camera


@pytest.mark.parametrize("bits", [32, 64])
def test_normal_stencil_on_cubic_surface(bits):
    """rgbdImage.cu:126-145 takes V[i + 2] - V[i - 2] along the row and
    V[i + W] - V[i - W] along the column.  On the surface
    X(u, v) = (a u, b v, z0 + al u^3 + be v^3) these differences are exactly
    (4a, 0, al((u+2)^3 - (u-2)^3)) and (0, 2b, be((v+1)^3 - (v-1)^3)), whose
    cross product gives the normal in closed form; a stencil of +-1 columns
    or +-2 rows, or a transposed pair, gives another direction (the cubic
    terms make central differences depend on the step)."""
    rows, cols = 40, 50
    a, b, z0, al, be = 0.01, 0.012, 1.5, 1e-5, -1.5e-5
    v, u = np.mgrid[0:rows, 0:cols].astype(np.float64)
    u, v = u - 25, v - 20  # a surface centred in the image
    vert = np.stack([a * u, b * v, z0 + al * u ** 3 + be * v ** 3], -1).astype(np.float32)
    nrm = np.zeros_like(vert)
    O.load(bits).or_normal_map(O.P(vert), rows, cols, O.P(nrm))

    def normal(hu, hv):  # cross of the +-hu column and +-hv row differences
        du = np.stack([np.full(u.shape, 2 * hu * a), np.zeros(u.shape), al * ((u + hu) ** 3 - (u - hu) ** 3)], -1)
        dv = np.stack([np.zeros(u.shape), np.full(u.shape, 2 * hv * b), be * ((v + hv) ** 3 - (v - hv) ** 3)], -1)
        n = np.cross(du, dv)
        return (n / np.linalg.norm(n, axis=-1, keepdims=True))[1:-1, 2:-2]

    want = normal(2, 1)
    got = nrm[1:-1, 2:-2].astype(np.float64)
    assert np.abs(got - want).max() < 2e-5
    # the alternative stencils give other directions, by far more than that
    for hu, hv in ((1, 1), (2, 2), (1, 2)):
        assert np.abs(normal(hu, hv) - want).max() > 1e-3


def twist_exp(xi):
    M = np.array([[0, -xi[2], xi[1], xi[3]], [xi[2], 0, -xi[0], xi[4]], [-xi[1], xi[0], 0, xi[5]], [0, 0, 0, 0]])
    return sl.expm(M)


def render_planes_s(T, s):
    """render_planes for the sensor s."""
    v, u = np.mgrid[0:s["rows"], 0:s["cols"]]
    dc = np.stack([(u - s["cx"]) / s["fx"], (v - s["cy"]) / s["fy"], np.ones_like(u, float)], -1)
    d = dc @ T[:3, :3].T
    o = T[:3, 3]
    best = np.full(u.shape, np.inf)
    nrm = np.zeros(u.shape + (3,))
    for n, c in PLANES:
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (c - o @ n) / (d @ n)
        ok = (t > 0) & (t < best)
        best = np.where(ok, t, best)
        nrm[ok] = n
    return best.astype(np.float32), (dc * best[..., None]).astype(np.float32), (nrm @ T[:3, :3]).astype(np.float32)


def c_round(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def brute_force_system(src, tgt, s, Tinc, icp, lin, h=1e-6):
    """The normal equations summed pixel by pixel in FP64 with Jacobians by
    central differences of each residual under T <- exp(xi) Tinc:
    geometric r = (p - q).n at the projectively associated target pixel,
    photometric r = I_t(pi(p)) - I_s with I_t the continuous linear image
    lin = (i0, di/du, di/dv) that the target intensities sample, looked up
    at the nearest pixel for the residual value itself.  Association gates
    as gnSolver.cu:25-196 with the build's documented changes (depth range,
    bounds, normal angle |cos| > cos(pi/10), point distance)."""
    sv, sn, si = src
    tv, tn, ti = tgt
    rows, cols = s["rows"], s["cols"]
    V = sv.reshape(-1, 3).astype(np.float64)
    N = sn.reshape(-1, 3).astype(np.float64)
    TV, TN = tv.reshape(-1, 3).astype(np.float64), tn.reshape(-1, 3).astype(np.float64)
    R, t = Tinc[:3, :3], Tinc[:3, 3]
    A = np.zeros((6, 6)); b = np.zeros(6); ng = nph = 0
    E = [(twist_exp(h * e), twist_exp(-h * e)) for e in np.eye(6)]
    proj = lambda p: (s["fx"] * p[0] / p[2] + s["cx"], s["fy"] * p[1] / p[2] + s["cy"])
    for i in range(rows * cols):
        if not (icp["min_depth"] <= V[i, 2] <= icp["max_depth"]):
            continue
        if not np.isfinite(N[i]).all() or not N[i].any():
            continue
        p = R @ V[i] + t
        pu, pv = (int(c_round(x)) for x in proj(p))
        if not (0 <= pu < cols and 0 <= pv < rows):
            continue
        j = pv * cols + pu
        q, n = TV[j], TN[j]
        if not (icp["min_depth"] <= q[2] <= icp["max_depth"]) or not np.isfinite(n).all() or not n.any():
            continue
        if not abs(n @ (R @ N[i])) > np.cos(np.pi / 10):
            continue
        if not np.linalg.norm(p - q) <= float(np.float32(icp["dist_max"])):
            continue
        moved = [(Ep[:3, :3] @ p + Ep[:3, 3], Em[:3, :3] @ p + Em[:3, 3]) for Ep, Em in E]
        J = np.array([((pp - q) @ n - (pm - q) @ n) / (2 * h) for pp, pm in moved])
        r = (p - q) @ n
        A += np.outer(J, J); b += J * r; ng += 1
        if icp["photo_weight"] > 0 and 1 <= pu < cols - 1 and 1 <= pv < rows - 1:
            It = lambda x: lin[0] + lin[1] * proj(x)[0] + lin[2] * proj(x)[1]
            Jp = np.array([(It(pp) - It(pm)) / (2 * h) for pp, pm in moved])
            rp = float(ti.reshape(-1)[j]) - float(si.reshape(-1)[i])
            A += icp["photo_weight"] * np.outer(Jp, Jp); b += icp["photo_weight"] * Jp * rp; nph += 1
    return A, b, ng, nph


@pytest.mark.parametrize("bits", [32, 64])
def test_icp_system_vs_brute_force(bits):
    """or_icp_system's A = sum J J^T and b = sum J r, geometric plus
    photometric (gnSolver.cu:25-196; the photometric term is the build's,
    the reference's being weighted 0), against brute_force_system.  The
    target intensity is the linear image 1/4 + u/128 - v/256 (exact in
    float), whose central-difference gradient is exactly its derivative, so
    the oracle's analytic photometric Jacobian (image gradient chained
    through the pinhole projection, then through exp(xi)) must equal the
    finite differences of the continuous residual.  fx != fy and the two
    image slopes differ, so an exchanged focal length or gradient axis, a
    missing 1/2 in the central difference, a flipped sign of either
    residual or Jacobian, or a reversed cross product all fail."""
    s = dict(rows=48, cols=64, fx=60.0, fy=52.0, cx=31.5, cy=22.0)
    Ta = pose([-0.35, 0.2, 0.0], [0.1, -0.1, 0.0])
    Tb = pose([-0.34, 0.21, 0.006], [0.108, -0.104, 0.006])
    _, va, na = render_planes_s(Ta, s)
    db, _, _ = render_planes_s(Tb, s)
    v, u = np.mgrid[0:s["rows"], 0:s["cols"]]
    lin = (0.25, 1 / 128, -1 / 256)
    it = (lin[0] + lin[1] * u + lin[2] * v).astype(np.float32)
    isrc = (0.5 + 0.1 * np.sin(0.3 * u) * np.cos(0.2 * v)).astype(np.float32)
    sv, sn, _ = O.source_maps(db, None, s, bits)[0]
    src = [(sv, sn, isrc)]
    tgt = [(va, na, it)]
    Tinc = twist_exp(np.array([0.004, -0.003, 0.002, 0.005, -0.004, 0.003])) @ np.linalg.inv(Ta) @ Tb
    icp = dict(min_depth=0.1, max_depth=10.0, dist_max=0.2, photo_weight=0.05)
    A, b, sums = O.icp_system(src, tgt, s, 0, Tinc, icp, bits)
    Ab, bb, ng, nph = brute_force_system((sv, sn, isrc), (va, na, it), s, Tinc, icp, lin)
    assert sums[0] == ng > 1500 and sums[2] == nph > 1500
    rt = 1e-7 if bits == 64 else 3e-5
    assert np.abs(A - Ab).max() <= rt * np.abs(Ab).max()
    assert np.abs(b - bb).max() <= rt * np.abs(bb).max()
    # the photometric part alone (difference of the two weights) as well
    A0, b0, _ = O.icp_system(src, tgt, s, 0, Tinc, dict(icp, photo_weight=0.0), bits)
    Ab0, bb0, _, _ = brute_force_system((sv, sn, isrc), (va, na, it), s, Tinc, dict(icp, photo_weight=0.0), lin)
    assert np.abs((A - A0) - (Ab - Ab0)).max() <= rt * np.abs(Ab - Ab0).max()
    assert np.abs((b - b0) - (bb - bb0)).max() <= rt * np.abs(bb - bb0).max()


@pytest.mark.parametrize("bits", [32, 64])
def test_photometric_term_recovers_in_plane_motion(bits):
    """On a single textured plane the point-to-plane term cannot see motion
    within the plane; the photometric term must recover it.  The camera
    moves 4 pixels along x and 4 along y at the plane's depth (1, 2 and 4
    pixels at the three pyramid levels, so the nearest-pixel intensity
    lookup is exact at the true pose).  With photo_weight = 0 the in-plane
    translation stays unrecovered.  With it, the pose converges into the
    term's fixed-point set: the residual reads the nearest target pixel, so
    every pose whose projections round to the true pixels has zero
    photometric residual, i.e. each in-plane axis within half a pixel of the
    truth (a flipped sign or a wrong gradient axis moves it away instead)."""
    s = dict(rows=96, cols=128, fx=100.0, fy=90.0, cx=63.5, cy=47.5)
    Z = 1.2
    tx, ty = 4 * Z / s["fx"], 4 * Z / s["fy"]
    v, u = np.mgrid[0:s["rows"], 0:s["cols"]]
    X, Y = (u - s["cx"]) / s["fx"] * Z, (v - s["cy"]) / s["fy"] * Z
    tex = lambda x, y: 0.5 + 0.2 * np.sin(2 * np.pi * x / 0.6) + 0.15 * np.cos(2 * np.pi * y / 0.5)
    depth = np.full(u.shape, Z, np.float32)
    va = np.stack([X, Y, depth], -1).astype(np.float32)
    na = np.zeros_like(va)
    na[..., 2] = -1
    ia = tex(X, Y).astype(np.float32)
    ib = tex(X + tx, Y + ty).astype(np.float32)  # camera B = A translated by (tx, ty, 0)
    src = [(lv[0], lv[1], np.ascontiguousarray(ib[:: 1 << l, :: 1 << l]))
           for l, lv in enumerate(O.source_maps(depth, None, s, bits))]
    tgt = O.model_maps(va, na, ia)
    true = np.eye(4)
    true[:2, 3] = tx, ty
    err = {}
    for w in (0.0, 0.01):
        icp = dict(lvl_iterations=(10, 10, 10), min_depth=0.1, max_depth=10.0, dist_max=0.2, photo_weight=w,
                   regularization=0.02, max_increment=0.3)
        T = O.icp_track(src, tgt, s, icp, bits)
        err[w] = (np.linalg.inv(true) @ T)[:3, 3]
    print("translation error, geometric only / with photometric:", err)
    assert np.hypot(*err[0.0][:2]) > 0.9 * np.hypot(tx, ty)
    px = np.array([Z / s["fx"], Z / s["fy"]])
    assert (np.abs(err[0.01][:2]) < 0.5 * px).all()
    assert abs(err[0.01][2]) < 1e-4
