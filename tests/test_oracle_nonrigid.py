"""Pins the oracle's deformation graph, dual-quaternion blending and
non-rigid Gauss-Newton (oracle/dm_oracle_nr.c) by independent means:
numpy brute-force neighbours and cell sampling, Kavan's published blending
formula written here from scipy rotations, closed forms (a common rigid
motion, the node-vertex rule), finite differences of the energy for the
Gauss-Newton gradient and Hessian, a dense numpy solve for the PCG, and
recovery of known deformations rendered by the synthetic generator.

Reference: include/geometry/defGraph/defGraph.cu, include/blender/dqBlender/
dqBlender.cu:162-186, include/math/dualQuat.h:293-302,
include/solver/nonRigidICP/nonRigidICP.cu:494-515.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_lib as O
from dynamap_amd import _native as N

SENS = dict(rows=240, cols=320, fx=262.5, fy=262.5, cx=159.5, cy=119.5)


def cylinder(nt, ny):
    S = N.synth()
    V = np.zeros((nt * ny, 3), np.float32)
    Nn = np.zeros_like(V)
    S.dms_cylinder_mesh(nt, ny, N.ptr(V), N.ptr(Nn))
    return V, Nn


def cyl_pose(f=0):
    S = N.synth()
    P = np.zeros((f + 1, 16))
    S.dms_cylinder_trajectory(f + 1, 11, N.ptr(P))
    return P[f].reshape(4, 4)


def cyl_depth(frame, pose, sens=SENS):
    S = N.synth()
    d = np.zeros((sens["rows"], sens["cols"]), np.float32)
    S.dms_cylinder_render(float(frame), N.ptr(np.ascontiguousarray(pose)), sens["rows"], sens["cols"], sens["fx"],
                          sens["fy"], sens["cx"], sens["cy"], 1, 0, N.ptr(d))
    return d


def random_cloud(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= 0.2
    return p.astype(np.float32), (p / 0.2).astype(np.float32)


def test_sampling_and_neighbours():
    V, Nn = random_cloud(3000, 1)
    r, dgw = np.float32(0.05), 0.04
    g = O.OracleGraph(V, Nn, float(r), dgw, bits=32)
    assert g.status == 0
    # per cell the vertex closest to the cell centre, float32 arithmetic
    c = np.floor(V / r).astype(np.int64)
    e = V - (c.astype(np.float32) + np.float32(0.5)) * r
    d2 = (np.float32(0) + e[:, 0] * e[:, 0]) + e[:, 1] * e[:, 1]
    d2 = d2 + e[:, 2] * e[:, 2]
    order = np.lexsort((np.arange(len(V)), d2, c[:, 2], c[:, 1], c[:, 0]))
    cs = c[order]
    head = np.ones(len(V), bool)
    head[1:] = np.any(cs[1:] != cs[:-1], axis=1)
    expect = np.sort(order[head])
    vid, nbr, w = g.nodes()
    assert np.array_equal(vid, expect)
    # 4 nearest nodes (brute force, FP64 distances), excluding self for nodes
    P = V[vid].astype(np.float64)
    D = np.linalg.norm(P[:, None] - P[None], axis=2)
    np.fill_diagonal(D, np.inf)
    assert np.array_equal(nbr, np.argsort(D, axis=1, kind="stable")[:, :4])
    dn = np.take_along_axis(D, nbr.astype(np.int64), 1)
    assert np.allclose(w, np.exp(-np.maximum(dn, 1e-4) ** 2 / dgw ** 2), rtol=1e-5, atol=1e-7)
    ids, vw = g.vertex_nn()
    Dv = np.linalg.norm(V[:, None].astype(np.float64) - P[None], axis=2)
    assert np.array_equal(ids, np.argsort(Dv, axis=1, kind="stable")[:, :4])


def kavan(T_nodes, ids, w, V, Nn):
    """Dual-quaternion linear blending (Kavan et al. 2007) from scipy rotations."""
    q = Rotation.from_matrix(T_nodes[:, :3, :3]).as_quat()  # x, y, z, w
    q = np.concatenate([q[:, 3:], q[:, :3]], 1)              # w, x, y, z
    q[q[:, 0] < 0] *= -1

    def qmul(a, b):
        w1, x1, y1, z1 = a.T
        w2, x2, y2, z2 = b.T
        return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], 1)

    t = np.concatenate([np.zeros((len(q), 1)), T_nodes[:, :3, 3]], 1)
    d = 0.5 * qmul(t, q)
    sgn = np.sign(np.einsum("vkc,vc->vk", q[ids], q[ids[:, 0]]))
    sgn[sgn == 0] = 1
    ww = w * sgn
    br = np.einsum("vk,vkc->vc", ww, q[ids])
    bd = np.einsum("vk,vkc->vc", ww, d[ids])
    n = np.linalg.norm(br, axis=1, keepdims=True)
    br, bd = br / n, bd / n
    conj = br * np.array([1, -1, -1, -1])
    tt = 2 * qmul(bd, conj)[:, 1:]
    R = Rotation.from_quat(np.concatenate([br[:, 1:], br[:, :1]], 1))
    return R.apply(V.astype(np.float64)) + tt, R.apply(Nn.astype(np.float64))


def random_transforms(n, seed, rot=0.2, trans=0.03):
    rng = np.random.default_rng(seed)
    T = np.tile(np.eye(4), (n, 1, 1))
    T[:, :3, :3] = Rotation.from_rotvec(rng.normal(size=(n, 3)) * rot).as_matrix()
    T[:, :3, 3] = rng.normal(size=(n, 3)) * trans
    return T


def test_blend_matches_kavan():
    V, Nn = random_cloud(2000, 2)
    g = O.OracleGraph(V, Nn, 0.06, 0.05, bits=64)
    T = random_transforms(g.nn, 3)
    g.set_transforms(T)
    vb, nb = g.blend()
    ids, w = g.vertex_nn()
    kv, kn = kavan(T, ids, w, V, Nn)
    vid, _, _ = g.nodes()
    isnode = np.zeros(len(V), bool)
    isnode[vid] = True
    # non-node vertices: Kavan's formula; node vertices: their node's transform
    assert np.abs(vb[~isnode] - kv[~isnode]).max() < 2e-6
    assert np.abs(nb[~isnode] - kn[~isnode]).max() < 2e-6
    own = np.einsum("nij,nj->ni", T[:, :3, :3], V[vid].astype(np.float64)) + T[:, :3, 3]
    assert np.abs(vb[vid] - own).max() < 2e-6


def test_blend_common_rigid_motion():
    V, Nn = random_cloud(1500, 4)
    for bits in (32, 64):
        g = O.OracleGraph(V, Nn, 0.07, 0.05, bits=bits)
        T = random_transforms(1, 5, rot=0.5, trans=0.2)[0]
        g.set_transforms(np.tile(T, (g.nn, 1, 1)))
        vb, nb = g.blend()
        assert np.abs(vb - (V @ T[:3, :3].T + T[:3, 3])).max() < (3e-6 if bits == 32 else 5e-8)
        assert np.abs(nb - Nn @ T[:3, :3].T).max() < (3e-6 if bits == 32 else 5e-8)


def test_dq_of_transform_unit():
    for T in random_transforms(20, 6, rot=1.5, trans=0.5):
        q = O.dq_of_transform(T)
        r, d = q[:4], q[4:]
        assert abs(np.dot(r, r) - 1) < 1e-12 and abs(np.dot(r, d)) < 1e-12 and r[0] >= 0


@pytest.fixture(scope="module")
def small_cyl():
    V, Nn = cylinder(60, 40)
    return V, Nn


def energy(g, depth, pose, **kw):
    _, _, _, e = g.system(depth, SENS, pose, pcg_iterations=0, **kw)
    return e[0] + e[1]


def perturb(T, i, k, eps):
    xi = np.zeros(6)
    xi[k] = eps
    E = np.zeros(16)
    O.load(64).or_se3_exp(O.P(xi), O.P(E))
    T2 = T.copy()
    T2[i] = E.reshape(4, 4) @ T[i]
    return T2


def flat_patch(nx, ny, z=1.0):
    x, y = np.meshgrid(np.linspace(-0.25, 0.25, nx), np.linspace(-0.18, 0.18, ny))
    V = np.stack([x.ravel(), y.ravel(), np.full(x.size, z)], 1).astype(np.float32)
    return V, np.tile(np.array([0, 0, -1], np.float32), (len(V), 1))


def plane_depth(a, b, c, sens=SENS):
    """Depth of the plane z = a + b x + c y (camera frame) along every pixel ray."""
    v, u = np.mgrid[0:sens["rows"], 0:sens["cols"]]
    rx = (u - sens["cx"]) / sens["fx"]
    ry = (v - sens["cy"]) / sens["fy"]
    return (a / (1 - b * rx - c * ry)).astype(np.float32)


def fd_gradient_check(g, depth, T0, kw, eps, idx):
    H, b, _, e = g.system(depth, SENS, np.eye(4), **kw)
    out = []
    for i, k in idx:
        g.set_transforms(perturb(T0, i, k, eps))
        ep = energy(g, depth, np.eye(4), **kw)
        g.set_transforms(perturb(T0, i, k, -eps))
        em = energy(g, depth, np.eye(4), **kw)
        out.append((-0.25 * (ep - em) / eps, b[6 * i + k]))
    g.set_transforms(T0)
    return np.array(out), b, H, e


def test_gn_gradient_and_hessian_by_finite_differences():
    """rhs = -J^T r equals -1/2 dE/dxi by central differences of the energy the
    oracle evaluates, for the data term (planar target, so projective
    re-association cannot make E jump; all nodes at one rigid offset, where
    the blended warp is rigid and the linear-blend Jacobian exact) and the
    ARAP term (random node transforms); H matches finite differences of the
    rhs on a zero-residual data problem (Gauss-Newton = Newton there)."""
    V, Nn = flat_patch(50, 36)
    g = O.OracleGraph(V, Nn, 0.08, 0.1, bits=64)
    rng = np.random.default_rng(0)
    idx = [(int(rng.integers(g.nn)), int(rng.integers(6))) for _ in range(16)]
    # data term
    T0 = np.tile(random_transforms(1, 7, rot=0.01, trans=0.004)[0], (g.nn, 1, 1))
    g.set_transforms(T0)
    d, b, _, e = fd_gradient_check(g, plane_depth(1.01, 0.03, -0.02), T0, dict(damping=0.0, lambda_reg=0.0), 1e-3,
                                   idx)
    assert e[2] > 1000 and np.abs(b).max() > 0.1
    assert np.abs(d[:, 0] - d[:, 1]).max() < 1e-4 * np.abs(b).max()
    # ARAP term
    T1 = random_transforms(g.nn, 8, rot=0.02, trans=0.003)
    g.set_transforms(T1)
    d, b, _, e = fd_gradient_check(g, plane_depth(1.01, 0.03, -0.02), T1, dict(damping=0.0, dist_max=1e-9), 1e-4,
                                   idx)
    assert e[2] == 0 and e[1] > 0
    assert np.abs(d[:, 0] - d[:, 1]).max() < 1e-7 * np.abs(b).max()
    # Hessian of the zero-residual data problem
    T2 = np.tile(np.eye(4), (g.nn, 1, 1))
    g.set_transforms(T2)
    depth = plane_depth(1.0, 0.0, 0.0)
    kw = dict(damping=0.0, lambda_reg=0.0, pcg_iterations=0)
    H, b0, _, _ = g.system(depth, SENS, np.eye(4), **kw)
    assert np.abs(b0).max() < 1e-5
    for i, k in idx[:6]:
        g.set_transforms(perturb(T2, i, k, 1e-4))
        _, bp, _, _ = g.system(depth, SENS, np.eye(4), **kw)
        g.set_transforms(perturb(T2, i, k, -1e-4))
        _, bm, _, _ = g.system(depth, SENS, np.eye(4), **kw)
        col = -(bp - bm) / 2e-4
        assert np.abs(col - H[:, 6 * i + k]).max() <= 1e-3 * np.abs(H[:, 6 * i + k]).max()


def test_pcg_matches_dense_solve(small_cyl):
    V, Nn = small_cyl
    g = O.OracleGraph(V, Nn, 0.12, 0.15, bits=64)
    assert 10 <= g.nn <= 128
    pose = cyl_pose(0)
    depth = cyl_depth(6, pose)
    H, b, x, e = g.system(depth, SENS, pose, pcg_iterations=500, pcg_tol=1e-12)
    assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() > 0
    xd = np.linalg.solve(H, b)
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max() + 1e-12


def test_identity_target_keeps_mesh():
    V, Nn = cylinder(120, 80)
    g = O.OracleGraph(V, Nn, 0.05, 0.08)
    pose = cyl_pose(0)
    depth = cyl_depth(0, pose)
    st = g.solve(depth, SENS, pose)
    assert st[0] > 1000
    vb, _ = g.blend()
    vis = V[:, 2] < -0.05
    # on the surface: radial error ~0; sliding along the axis is not observable
    rad = np.sqrt(vb[:, 0] ** 2 + vb[:, 2] ** 2)
    assert np.abs(rad[vis] - 0.15).max() < 1e-4
    assert np.abs(vb[vis] - V[vis]).max() < 3e-3


def test_rigid_motion_recovered():
    """The object moved rigidly (rendered from the inversely moved camera):
    every visible vertex lands on the moved canonical surface."""
    V, Nn = cylinder(120, 80)
    g = O.OracleGraph(V, Nn, 0.05, 0.08)
    pose = cyl_pose(0)
    M = np.eye(4)
    M[:3, :3] = Rotation.from_rotvec([0.0, 0.02, 0.01]).as_matrix()
    M[:3, 3] = [0.006, -0.004, 0.005]
    depth = cyl_depth(0, np.linalg.inv(M) @ pose)
    for _ in range(3):
        g.solve(depth, SENS, pose)
    vb, _ = g.blend()
    # distance to the moved cylinder surface (axis symmetric: compare radially)
    q = (vb - M[:3, 3]) @ M[:3, :3]
    rad = np.sqrt(q[:, 0] ** 2 + q[:, 2] ** 2)
    vis = V[:, 2] < -0.07
    assert np.abs(rad[vis] - 0.15).mean() < 2e-4
    assert np.abs(rad[vis] - 0.15).max() < 3e-3


def test_bend_recovered():
    V, Nn = cylinder(120, 80)
    g = O.OracleGraph(V, Nn, 0.05, 0.08)
    pose = cyl_pose(0)
    f = 10.0
    depth = cyl_depth(f, pose)
    b = lambda y: 0.05 * np.sin(2 * np.pi * f / 40) * ((0.3 - y) / 0.6) ** 2
    vis = V[:, 2] < -0.07
    err0 = np.abs(np.sqrt((V[:, 0] - b(V[:, 1])) ** 2 + V[:, 2] ** 2) - 0.15)[vis].mean()
    for _ in range(3):
        g.solve(depth, SENS, pose)
    vb, _ = g.blend()
    err = np.abs(np.sqrt((vb[:, 0] - b(vb[:, 1])) ** 2 + vb[:, 2] ** 2) - 0.15)[vis]
    assert err0 > 5e-3 and err.mean() < 5e-4


# ------------------------------------------------ given nodes, update, source gate, mesh target
def test_given_nodes_match_brute_force():
    """or_graph_create_nodes (defGraph::init with the reference's own node
    choice, defGraph.cu:673-702): neighbours of an arbitrary, unsorted node
    list by brute force; the sampled node list reproduces the sampled graph;
    defGraphUpdateNodes recomputes the same links and keeps the transforms."""
    V, Nn = random_cloud(2000, 7)
    rng = np.random.default_rng(3)
    ids = rng.choice(len(V), 150, replace=False).astype(np.int32)  # unsorted, as mode 0's random picks
    dgw = 0.05
    g = O.OracleGraph(V, Nn, dgw=dgw, node_ids=ids)
    assert g.status == 0 and g.nn == 150
    vid, nbr, w = g.nodes()
    assert np.array_equal(vid, ids)
    P = V[ids].astype(np.float64)
    D = np.linalg.norm(P[:, None] - P[None], axis=2)
    np.fill_diagonal(D, np.inf)
    assert np.array_equal(nbr, np.argsort(D, axis=1, kind="stable")[:, :4])
    dn = np.take_along_axis(D, nbr.astype(np.int64), 1)
    assert np.allclose(w, np.exp(-np.maximum(dn, 1e-4) ** 2 / dgw ** 2), rtol=1e-12, atol=1e-15)
    vids, vw = g.vertex_nn()
    Dv = np.linalg.norm(V[:, None].astype(np.float64) - P[None], axis=2)
    assert np.array_equal(vids, np.argsort(Dv, axis=1, kind="stable")[:, :4])
    # a node vertex blends with its own node only (the node-vertex rule)
    T = np.tile(np.eye(4), (g.nn, 1, 1))
    T[5, :3, 3] = [0.01, 0.02, -0.03]
    g.set_transforms(T)
    bv, _ = g.blend()
    assert np.allclose(bv[ids[5]], V[ids[5]] + T[5, :3, 3], atol=1e-7)
    g.update_nodes()
    assert np.array_equal(g.nodes()[1], nbr) and np.array_equal(g.vertex_nn()[0], vids)
    assert np.array_equal(g.transforms(), T)
    # the grid-sampled node list given explicitly gives the same graph
    gs = O.OracleGraph(V, Nn, 0.05, dgw)
    svid, snbr, sw = gs.nodes()
    ge = O.OracleGraph(V, Nn, dgw=dgw, node_ids=svid)
    evid, enbr, ew = ge.nodes()
    assert np.array_equal(evid, svid) and np.array_equal(enbr, snbr) and np.array_equal(ew, sw)
    assert np.array_equal(ge.vertex_nn()[0], gs.vertex_nn()[0])
    # repeated or out-of-range ids are rejected
    assert O.OracleGraph(V, Nn, dgw=dgw, node_ids=[1, 2, 3, 4, 4]).status == -3
    assert O.OracleGraph(V, Nn, dgw=dgw, node_ids=[1, 2, 3, 4, len(V)]).status == -3
    assert O.OracleGraph(V, Nn, dgw=dgw, node_ids=[1, 2, 3, 4]).status == -1


def plane_grid(n=60, z=1.0, step=0.01):
    """A fronto-parallel vertex grid at depth z (camera = world), normals to the camera."""
    xs = (np.arange(n) - n / 2 + 0.25) * step
    X, Y = np.meshgrid(xs, xs)
    V = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], 1).astype(np.float32)
    Nn = np.tile(np.array([0, 0, -1], np.float32), (len(V), 1))
    return V, Nn


def test_source_gate_selects_pixels():
    """or_nonrigid_solve_pair: a vertex takes part only where the source depth
    at its projected pixel is in range (nonRigidICP.cu:291-292).  On a wall
    seen head on, the residual count equals the number of vertices whose
    pixel (u = round(fx x / z + cx), v = round(fy y / z + cy)) the source
    keeps; an all-valid source changes nothing."""
    V, Nn = plane_grid()
    s = SENS
    depth = np.full((s["rows"], s["cols"]), 1.0, np.float32)
    u = np.floor(s["fx"] * V[:, 0].astype(np.float64) / V[:, 2] + s["cx"] + 0.5).astype(int)
    v = np.floor(s["fy"] * V[:, 1].astype(np.float64) / V[:, 2] + s["cy"] + 0.5).astype(int)
    src = depth.copy()
    src[:, : s["cols"] // 2] = 0.0     # left half of the source unobserved
    src[: s["rows"] // 2 - 3, :] = 9.0  # upper rows beyond max_depth
    keep = (src[v, u] >= 0.1) & (src[v, u] <= 5.0)
    assert 0 < keep.sum() < len(V)
    kw = dict(gn_iterations=1, pcg_iterations=5)
    g0 = O.OracleGraph(V, Nn, 0.05, 0.1)
    st0 = g0.solve(depth, s, np.eye(4), **kw)
    assert st0[0] == len(V)
    g1 = O.OracleGraph(V, Nn, 0.05, 0.1)
    st1 = g1.solve_pair(depth, depth, s, np.eye(4), **kw)
    assert np.array_equal(st1, st0) and np.array_equal(g1.transforms(), g0.transforms())
    g2 = O.OracleGraph(V, Nn, 0.05, 0.1)
    st2 = g2.solve_pair(depth, src, s, np.eye(4), **kw)
    assert st2[0] == keep.sum()
    # the gate reads the source, not the target: the swapped call keeps all
    g3 = O.OracleGraph(V, Nn, 0.05, 0.1)
    assert g3.solve_pair(depth, depth, s, np.eye(4), **kw)[0] == len(V)


def test_mesh_target_recovers_rigid_motion():
    """or_nonrigid_solve_mesh (nonRigidICP.cu:706-852, vertex v <-> v): the
    canonical mesh itself is a zero-residual target (no step); a rigidly
    moved copy is recovered by every node; vertices with a zero normal or a
    NaN position drop out."""
    V, Nn = cylinder(60, 40)
    pose = cyl_pose(3)
    Ti = np.linalg.inv(pose)
    cam = lambda P, R=Ti: (P @ R[:3, :3].T + R[:3, 3]).astype(np.float32)
    kw = dict(gn_iterations=5, pcg_iterations=60, pcg_tol=1e-10, dist_max=0.1, normal_cos=0.3, lambda_reg=1.0)
    g = O.OracleGraph(V, Nn, 0.04, 0.08)
    st = g.solve_mesh(cam(V.astype(np.float64)), (Nn.astype(np.float64) @ Ti[:3, :3].T).astype(np.float32), pose,
                      **kw)
    assert st[0] == len(V) and st[1] < 1e-12
    # (the target is rounded to float32 in the camera frame: steps of that size only)
    assert np.allclose(g.transforms(), np.eye(4), atol=1e-6)
    # a rigid motion M of the target (world frame) with no part along the
    # cylinder's own symmetries (rotation about and sliding along its y axis),
    # which point-to-plane residuals cannot see
    M = np.eye(4)
    M[:3, :3] = Rotation.from_rotvec([0.02, 0.0, 0.015]).as_matrix()
    M[:3, 3] = [0.01, 0.0, 0.012]
    Vt = (V.astype(np.float64) @ M[:3, :3].T + M[:3, 3])
    Nt = Nn.astype(np.float64) @ M[:3, :3].T
    mv, mn = cam(Vt), (Nt @ Ti[:3, :3].T).astype(np.float32)
    g = O.OracleGraph(V, Nn, 0.04, 0.08)
    e0 = O.OracleGraph(V, Nn, 0.04, 0.08).solve_mesh(mv, mn, pose, **dict(kw, gn_iterations=1))[1]
    for _ in range(3):
        st = g.solve_mesh(mv, mn, pose, **kw)
    assert st[1] < 1e-5 * e0
    T = g.transforms()
    assert np.abs(T - M).max() < 1e-3, np.abs(T - M).max()  # from a 1.5 cm, 1.4 degree motion
    bv, _ = g.blend()
    assert np.abs(bv - Vt).max() < 5e-4
    # dropped vertices
    mn2, mv2 = mn.copy(), mv.copy()
    mn2[:10] = 0
    mv2[10:15] = np.nan
    g = O.OracleGraph(V, Nn, 0.04, 0.08)
    assert g.solve_mesh(mv2, mn2, pose, gn_iterations=1)[0] == len(V) - 15
