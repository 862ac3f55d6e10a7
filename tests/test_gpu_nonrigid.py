"""GPU deformation graph, dual-quaternion blending and non-rigid Gauss-Newton
against the CPU oracle: graph (node set, neighbour ids) bit-exact vs the FP32
transcription, blending within float rounding of it, and the solve within a
stated tolerance of the FP64 oracle on the bending-cylinder scene.

Reference: include/geometry/defGraph/defGraph.cu, include/blender/dqBlender/
dqBlender.cu:162-218, include/solver/nonRigidICP/nonRigidICP.cu:557-704.
"""
import time

import numpy as np
import pytest

import dynamap_amd as dm
import oracle_lib as O
from test_oracle_nonrigid import cyl_depth, cyl_pose, cylinder, random_cloud, random_transforms, SENS

pytestmark = pytest.mark.gpu


def bend_err(vb, f):
    b = 0.05 * np.sin(2 * np.pi * f / 40) * ((0.3 - vb[:, 1]) / 0.6) ** 2
    return np.abs(np.sqrt((vb[:, 0] - b) ** 2 + vb[:, 2] ** 2) - 0.15)


@pytest.mark.parametrize("mesh", ["cloud", "cylinder"])
def test_graph_matches_transcription(mesh):
    V, Nn = random_cloud(5000, 3) if mesh == "cloud" else cylinder(200, 100)
    r, dgw = (0.05, 0.04) if mesh == "cloud" else (0.025, 0.05)
    g = dm.DefGraph(V, Nn, r, dgw)
    o = O.OracleGraph(V, Nn, r, dgw, bits=32)
    assert g.nodeNum == o.nn
    vid, nbr, w = g.nodes()
    ovid, onbr, ow = o.nodes()
    assert np.array_equal(vid, ovid)
    assert np.array_equal(nbr, onbr)
    assert np.allclose(w, ow, rtol=2e-7, atol=1e-12)
    ids, vw = g.vertex_neighbours()
    oids, ovw = o.vertex_nn()
    assert np.array_equal(ids, oids)
    assert np.allclose(vw, ovw, rtol=2e-7, atol=1e-12)
    assert g.num_blocks() >= g.nodeNum


def test_blend_matches_transcription():
    V, Nn = cylinder(200, 100)
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    o = O.OracleGraph(V, Nn, 0.025, 0.05, bits=32)
    T = random_transforms(g.nodeNum, 9, rot=0.1, trans=0.02)
    g.set_transforms(T)
    o.set_transforms(T)
    assert np.array_equal(g.transforms(), T)
    gv, gn = dm.DqBlender(g).blendMesh()
    ov, on = o.blend()
    dv, dn = np.abs(gv - ov).max(), np.abs(gn - on).max()
    print("blend max |gpu - fp32 oracle|", dv, dn)
    assert dv <= 1e-6 and dn <= 1e-6


@pytest.fixture(scope="module")
def bend_case():
    V, Nn = cylinder(200, 100)
    pose = cyl_pose(0)
    return V, Nn, pose, cyl_depth(10, pose)


def test_solve_matches_fp64_oracle(bend_case):
    V, Nn, pose, depth = bend_case
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    o = O.OracleGraph(V, Nn, 0.025, 0.05, bits=64)
    solver = dm.NonRigidICP()
    sg = solver.solve(g, depth, SENS, pose)
    so = o.solve(depth, SENS, pose)
    gv, _ = dm.DqBlender(g).blendMesh()
    ov, on = o.blend()
    vis = V[:, 2] < -0.07
    # Compared along the surface normal: sliding along the cylinder (axial,
    # about the axis) is not observable by a point-to-plane term and is
    # decided by rounding (the FP32 and FP64 oracles differ there by mm too).
    dn = np.abs(np.sum((gv - ov) * on, 1))[vis]
    print("stats gpu", sg, "oracle", so, "normal-direction |gpu - fp64| mean / p95 / max", dn.mean(),
          np.percentile(dn, 95), dn.max())
    assert abs(sg["residuals"] - so[0]) <= 0.01 * so[0]
    assert dn.mean() < 1e-4 and np.percentile(dn, 95) < 3e-4
    assert bend_err(gv, 10)[vis].mean() < 3e-4 and bend_err(ov, 10)[vis].mean() < 3e-4


def test_solve_is_deterministic(bend_case):
    """Two graphs built from the same mesh and solved on the same frames give
    bit-equal node transforms: every PCG dot product is a fixed-order sum of
    workgroup partials, and no workgroup overwrites a partial that another
    one of the same launch still reads."""
    V, Nn, pose, depth = bend_case
    out = []
    for _ in range(2):
        g = dm.DefGraph(V, Nn, 0.025, 0.05)
        solver = dm.NonRigidICP()
        for f in (4, 8, 12):
            solver.solve(g, cyl_depth(f, cyl_pose(f)), SENS, cyl_pose(f))
        out.append(g.transforms())
        g.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_identity_target_keeps_surface():
    V, Nn = cylinder(200, 100)
    pose = cyl_pose(0)
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    st = dm.NonRigidICP().solve(g, cyl_depth(0, pose), SENS, pose)
    assert st["residuals"] > 1000
    gv, _ = dm.DqBlender(g).blendMesh()
    vis = V[:, 2] < -0.05
    rad = np.sqrt(gv[:, 0] ** 2 + gv[:, 2] ** 2)
    assert np.abs(rad[vis] - 0.15).max() < 2e-4


def test_bending_sequence_200k():
    """The deforming-cylinder workload at its full mesh size (200 000
    vertices, 640x480, 2.5 cm nodes): ten frames of increasing bend, one
    solve (5 Gauss-Newton iterations) per frame."""
    V, Nn = cylinder(500, 400)
    sens = dict(rows=480, cols=640, fx=525.0, fy=525.0, cx=319.5, cy=239.5)
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    solver = dm.NonRigidICP()
    poses = [cyl_pose(f) for f in range(10)]
    frames = [cyl_depth(f, poses[f], sens) for f in range(10)]
    vis = V[:, 2] < -0.07
    t = time.time()
    for f in range(10):
        st = solver.solve(g, frames[f], sens, poses[f])
    el = time.time() - t
    gv, _ = dm.DqBlender(g).blendMesh()
    err = bend_err(gv, 9)[vis]
    print(f"nodes {g.nodeNum} blocks {g.num_blocks()} {10 / el:.1f} frames/s, surface err mean "
          f"{1000 * err.mean():.3f} mm max {1000 * err.max():.3f} mm, last {st}")
    assert err.mean() < 5e-4


def test_bending_full_length_120():
    """The deforming-cylinder workload at its full length: 120 frames (three
    bend periods of 40 frames, through both bend directions), one solve per
    frame; the blended surface follows the analytic bend at every 20th
    frame."""
    V, Nn = cylinder(500, 400)
    sens = dict(rows=480, cols=640, fx=525.0, fy=525.0, cx=319.5, cy=239.5)
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    solver = dm.NonRigidICP()
    vis = V[:, 2] < -0.07
    errs = []
    for f in range(120):
        pose = cyl_pose(f)
        solver.solve(g, cyl_depth(f, pose, sens), sens, pose)
        if f % 20 == 19:
            gv, _ = dm.DqBlender(g).blendMesh()
            errs.append(float(bend_err(gv, f)[vis].mean()))
    print("surface error mean (mm) at frames 19, 39, ..., 119:", [round(1000 * e, 4) for e in errs])
    g.close()
    assert max(errs) < 5e-4


def test_given_nodes_and_update_match_transcription():
    """dm_graph_create_nodes (defGraph::init with the reference's own node
    choice, defGraph.cu:673-702) and dm_graph_update_nodes
    (defGraphUpdateNodes, :483-545): neighbour ids bit-exact vs the FP32
    transcription for an unsorted random node list; the update keeps
    transforms and links."""
    V, Nn = cylinder(200, 100)
    ids = np.random.default_rng(11).choice(len(V), 700, replace=False).astype(np.int32)
    g = dm.DefGraph(V, Nn, dgw=0.05, node_ids=ids)
    o = O.OracleGraph(V, Nn, dgw=0.05, node_ids=ids, bits=32)
    assert g.nodeNum == o.nn == 700
    vid, nbr, w = g.nodes()
    ovid, onbr, ow = o.nodes()
    assert np.array_equal(vid, ids) and np.array_equal(nbr, onbr)
    assert np.allclose(w, ow, rtol=2e-7, atol=1e-12)
    vids, vw = g.vertex_neighbours()
    assert np.array_equal(vids, o.vertex_nn()[0])
    T = random_transforms(g.nodeNum, 4, rot=0.05, trans=0.01)
    g.set_transforms(T)
    g.defGraphUpdateNodes()
    assert np.array_equal(g.nodes()[1], nbr) and np.array_equal(g.vertex_neighbours()[0], vids)
    assert np.array_equal(g.transforms(), T)
    with pytest.raises(dm._native.DmError):
        dm.DefGraph(V, Nn, dgw=0.05, node_ids=[1, 2, 3, 4, 4])


def test_pair_and_mesh_solves_match_fp64_oracle(bend_case):
    """dm_nonrigid_solve_pair (source depth gates the pixels,
    nonRigidICP.cu:291-292) and dm_nonrigid_solve_mesh (vertex-for-vertex
    mesh target, nonRigidICP.cu:706-852) against the FP64 oracle."""
    V, Nn, pose, depth = bend_case
    src = depth.copy()
    src[:, : SENS["cols"] // 2] = 0.0
    solver = dm.NonRigidICP()
    # the first linearisation (identity transforms): the same gated residuals
    g1 = dm.DefGraph(V, Nn, 0.025, 0.05)
    s1 = solver.solve(g1, depth, SENS, pose, source_depth=src, gn_iterations=1)
    o1 = O.OracleGraph(V, Nn, 0.025, 0.05, bits=64).solve_pair(depth, src, SENS, pose, gn_iterations=1)
    full = O.OracleGraph(V, Nn, 0.025, 0.05, bits=64).solve(depth, SENS, pose, gn_iterations=1)
    print("pair, first linearisation: gpu", s1, "oracle", o1, "ungated oracle residuals", full[0])
    assert o1[0] < 0.8 * full[0]
    assert abs(s1["residuals"] - o1[0]) <= 0.002 * o1[0]
    assert abs(s1["data_energy"] - o1[1]) <= 0.01 * o1[1]
    # the full solve: where the gated frame has data, both fit the bend
    g = dm.DefGraph(V, Nn, 0.025, 0.05)
    o = O.OracleGraph(V, Nn, 0.025, 0.05, bits=64)
    solver.reserve(g, SENS["rows"] * SENS["cols"])
    sg = solver.solve(g, depth, SENS, pose, source_depth=src)
    so = o.solve_pair(depth, src, SENS, pose)
    print("pair: gpu", sg, "oracle", so)
    gv, _ = dm.DqBlender(g).blendMesh()
    ov, on = o.blend()
    Ti = np.linalg.inv(pose)
    cam = V.astype(np.float64) @ Ti[:3, :3].T + Ti[:3, 3]
    u = SENS["fx"] * cam[:, 0] / cam[:, 2] + SENS["cx"]
    seen = (V[:, 2] < -0.07) & (u > SENS["cols"] // 2 + 20)  # well inside the kept half
    assert seen.sum() > 1000
    eg, eo = bend_err(gv, 10)[seen], bend_err(ov, 10)[seen]
    print(f"gated solve, surface error where data is kept: gpu {1000 * eg.mean():.4f} mm, "
          f"oracle {1000 * eo.mean():.4f} mm")
    assert eg.mean() < 3e-4 and eo.mean() < 3e-4
    # mesh target: the analytic bend of frame 10 in the camera frame at pose
    b = 0.05 * np.sin(2 * np.pi * 10 / 40) * ((0.3 - V[:, 1].astype(np.float64)) / 0.6) ** 2
    Vt = V.astype(np.float64).copy()
    Vt[:, 0] += b
    Ti = np.linalg.inv(pose)
    mv = (Vt @ Ti[:3, :3].T + Ti[:3, 3]).astype(np.float32)
    mn = (Nn.astype(np.float64) @ Ti[:3, :3].T).astype(np.float32)
    g2 = dm.DefGraph(V, Nn, 0.025, 0.05)
    o2 = O.OracleGraph(V, Nn, 0.025, 0.05, bits=64)
    sg2 = solver.solve_mesh(g2, mv, mn, pose)
    so2 = o2.solve_mesh(mv, mn, pose)
    print("mesh: gpu", sg2, "oracle", so2)
    assert sg2["residuals"] == so2[0] > 0.9 * len(V)
    gv2, _ = dm.DqBlender(g2).blendMesh()
    ov2, on2 = o2.blend()
    dn2 = np.abs(np.sum((gv2 - ov2) * on2, 1))
    assert dn2.mean() < 1e-4
    # five Gauss-Newton steps from the rest shape close most of the bend
    # (up to 5 cm); the target surface is reached along its normal
    e0 = np.abs(np.sum((V - Vt) * Nn, 1)).mean()
    e1 = np.abs(np.sum((ov2 - Vt) * on2, 1)).mean()
    print(f"mesh target: normal-direction distance to the target {1000 * e0:.3f} mm -> {1000 * e1:.3f} mm")
    assert e1 < 0.1 * e0


def test_4096_node_graph_matches_fp64_oracle():
    """The deforming-cylinder workload's 4096-node reading: the 200 000-vertex
    mesh with nodes on a 13.3 mm grid (4048 nodes, the grid rule's closest
    count to 4096) and a 2.5 cm weight radius, five frames of the bend, one
    solve per frame, against the FP64 oracle on the same frames."""
    V, Nn = cylinder(500, 400)
    sens = dict(rows=480, cols=640, fx=525.0, fy=525.0, cx=319.5, cy=239.5)
    g = dm.DefGraph(V, Nn, 0.0133, 0.025, max_nodes=8192)
    o = O.OracleGraph(V, Nn, 0.0133, 0.025, max_nodes=8192, bits=64)
    assert g.nodeNum == o.nn == 4048
    assert np.array_equal(g.nodes()[0], o.nodes()[0])
    solver = dm.NonRigidICP()
    for f in range(5, 10):
        pose = cyl_pose(f)
        d = cyl_depth(f, pose, sens)
        sg = solver.solve(g, d, sens, pose)
        so = o.solve(d, sens, pose)
    gv, _ = dm.DqBlender(g).blendMesh()
    ov, on = o.blend()
    vis = V[:, 2] < -0.07
    dn = np.abs(np.sum((gv - ov) * on, 1))[vis]
    eg, eo = bend_err(gv, 9)[vis], bend_err(ov, 9)[vis]
    print(f"nodes {g.nodeNum} blocks {g.num_blocks()}: normal-direction |gpu - fp64| mean {dn.mean():.2e} "
          f"p95 {np.percentile(dn, 95):.2e}; surface error gpu {1000 * eg.mean():.4f} mm, oracle "
          f"{1000 * eo.mean():.4f} mm; last stats gpu {sg} oracle {so}")
    assert abs(sg["residuals"] - so[0]) <= 0.01 * so[0]
    assert dn.mean() < 1e-4 and np.percentile(dn, 95) < 3e-4
    assert eg.mean() < 3e-4


def test_solve_many_equals_one_at_a_time():
    """dm_nonrigid_solve_many (the graphs' solves enqueued together on their
    own streams) against the same solves one after another: node transforms
    and stats bit-equal, over two frames."""
    V, Nn = cylinder(200, 100)
    frames, poses = [], []
    for f in range(6):
        P = cyl_pose(f)
        frames.append(cyl_depth(f, P))
        poses.append(P)
    solver = dm.NonRigidICP()
    seq = [dm.DefGraph(V, Nn, 0.025, 0.05) for _ in range(3)]
    many = [dm.DefGraph(V, Nn, 0.025, 0.05) for _ in range(3)]
    st_seq, st_many = [], []
    for step in range(2):
        idx = [k + 3 * step for k in range(3)]  # graph k sees frame k (+3 on the second step)
        st_seq.append([solver.solve(seq[k], frames[i], SENS, poses[i]) for k, i in enumerate(idx)])
        st_many.append(solver.solve_many(many, [frames[i] for i in idx], SENS, [poses[i] for i in idx]))
    for k in range(3):
        assert np.array_equal(seq[k].transforms(), many[k].transforms()), f"graph {k}"
    assert st_seq == st_many
    with pytest.raises(RuntimeError, match="twice"):
        solver.solve_many([many[0], many[0]], frames[:2], SENS, poses[:2])
    for g in seq + many:
        g.close()


def test_batch_equals_one_at_a_time():
    """dm_nonrigid_batch (every kernel of the solve launched once for all
    graphs, graph = blockIdx.y, replayed from one captured HIP graph)
    against the same solves one after another on their own: node transforms
    bit-equal and stats equal, over three calls.  The graphs differ in size
    (node radius and vertex count), so each kernel's batched grid is the
    largest graph's and the smaller graphs' extra workgroups must return;
    the third call uses 160x120 frames, for which the batch captures its
    solve again."""
    meshes = [cylinder(200, 100), cylinder(200, 100), cylinder(120, 60)]
    radii = [0.025, 0.035, 0.03]
    frames, poses = [], []
    for f in range(9):
        P = cyl_pose(f)
        frames.append(cyl_depth(f, P))
        poses.append(P)
    solver = dm.NonRigidICP()
    seq = [dm.DefGraph(V, Nn, r, 0.05) for (V, Nn), r in zip(meshes, radii)]
    bat = [dm.DefGraph(V, Nn, r, 0.05) for (V, Nn), r in zip(meshes, radii)]
    assert len({g.nodeNum for g in bat}) == 3
    batch = solver.batch(bat)
    half = dict(rows=120, cols=160, fx=131.25, fy=131.25, cx=79.5, cy=59.5)
    st_seq, st_bat = [], []
    for step in range(3):
        idx = [k + 3 * step for k in range(3)]
        sens = SENS if step < 2 else half
        fr = [frames[i] if step < 2 else cyl_depth(i, poses[i], half) for i in idx]
        st_seq.append([solver.solve(seq[k], fr[k], sens, poses[i]) for k, i in enumerate(idx)])
        st_bat.append(batch.solve(fr, sens, [poses[i] for i in idx]))
        for k in range(3):
            assert np.array_equal(seq[k].transforms().view(np.uint64), bat[k].transforms().view(np.uint64)), \
                f"call {step} graph {k}"
    assert st_seq == st_bat
    assert all(s["residuals"] > 0 for row in st_bat for s in row)
    with pytest.raises(RuntimeError, match="twice"):
        solver.batch([bat[0], bat[1], bat[0]])
    batch.close()
    for g in seq + bat:
        g.close()
