"""Serving loop (no reference counterpart): every session's result is the
same alone and in the mix, batched or per frame, with or without Poisson
arrivals, under HBM-budget admission, and after the launch-shape autotuner
(the per-frame path is builder::buildModelLS, include/builder/builder.cu:70-139).
Poses are compared bit for bit."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs

pytestmark = pytest.mark.gpu
N = dm._native
NF = 6


def small(name):
    w = configs.WORKLOADS[name]
    t = dict(w["tsdf"], num_buckets=1 << 16, num_blocks=1 << 15)
    return w, t


@pytest.fixture(scope="module")
def traces():
    out = []
    for s, name in enumerate(["rigid-320x240-2cm"] * 3 + ["rigid-640x480-8mm"] * 3):
        d, c, g = configs.make_sequence(name, NF, session=s)
        out.append((name, d, c, g))
    return out


def server(traces, sel=None, budget=1 << 40):
    srv = dm.Server(8, 480, 640, hbm_budget_bytes=budget)
    for k, (name, d, c, g) in enumerate(traces):
        if sel is not None and k not in sel:
            continue
        w, t = small(name)
        srv.add_trace(t, w["icp"], w["sensor"], g[0], d, c)
    return srv


def test_batched_equals_per_frame_and_alone(traces):
    srv = server(traces)
    pb = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.reset()
    pd = srv.run(N.DM_RUN_PER_FRAME_DEV, 0, NF, want_poses=True)["poses"]
    srv.reset()
    ph = srv.run(N.DM_RUN_PER_FRAME_HOST, 0, NF, want_poses=True)["poses"]
    srv.close()
    assert np.array_equal(pb, pd) and np.array_equal(pb, ph)
    for k in (1, 4):  # alone
        s1 = server(traces, sel=[k])
        p1 = s1.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"][0]
        s1.close()
        assert np.array_equal(p1, pb[k])
    gt = np.stack([t[3] for t in traces])
    assert np.abs(pb[..., :3, 3] - gt[..., :3, 3]).max() < 5e-3


def test_poisson_arrivals_and_budget_admission(traces):
    srv = server(traces)
    ref = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.close()
    rng = np.random.default_rng(3)
    gaps = rng.exponential(1.0 / 30.0, size=(len(traces), NF))
    arrivals = np.cumsum(gaps, axis=1) - gaps[:, :1]
    srv = server(traces)
    r = srv.run(N.DM_RUN_BATCHED, 0, NF, arrivals=arrivals, want_poses=True)
    assert np.all(r["completions"] >= arrivals - 1e-9)
    assert np.array_equal(r["poses"], ref)
    # a budget for about two sessions: admitted in arrival order as others retire
    budget = 2 * max(srv.trace_bytes(t) for t in range(len(traces))) + (1 << 20)
    srv.close()
    sb = server(traces, budget=budget)
    rb = sb.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)
    sb.close()
    assert np.array_equal(rb["poses"], ref)


def test_autotune_keeps_results(traces):
    srv = server(traces)
    ref = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.reset()
    srv.run(N.DM_RUN_BATCHED, 0, 2)
    tune = srv.autotune(frame=2, reps=2)
    print("tuning", tune)
    assert tune["ray_tile"] in ("8x32", "16x16", "32x8", "64x4")
    assert tune["integrate_workgroups"] in (1024, 2048, 4096)
    # the ICP variants are timed at every level but the configured one stays
    assert tune["icp_reduce"] == "shuffle"
    assert all(len(tune["icp_reduce_ms"][k]) == 3 and min(tune["icp_reduce_ms"][k]) > 0 for k in ("shuffle", "mfma"))
    p = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.close()
    assert np.array_equal(p, ref)


def test_icp_matrix_core_variant(traces):
    """The FP32 matrix-core linearisation (dm_server_set_shapes) against the
    shuffle-tree one: the same poses up to summation order, and batched ==
    per-frame within the variant.  Launch shapes that do not touch the ICP
    leave the poses bit-equal."""
    srv = server(traces)
    ref = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.set_shapes(ray_tile_width=8, integrate_workgroups=1024, icp_reduce="shuffle")
    same = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    srv.set_shapes(ray_tile_width=16, integrate_workgroups=2048, icp_reduce="mfma")
    pm = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    pd = srv.run(N.DM_RUN_PER_FRAME_DEV, 0, NF, want_poses=True)["poses"]
    with pytest.raises(RuntimeError):
        srv.set_shapes(ray_tile_width=12)
    srv.close()
    assert np.array_equal(same, ref)
    assert np.array_equal(pm, pd)
    dt = np.abs(pm[..., :3, 3] - ref[..., :3, 3]).max()
    dr = np.abs(pm[..., :3, :3] - ref[..., :3, :3]).max()
    print(f"matrix cores vs shuffle tree: max translation {dt:.3g} m, max rotation entry {dr:.3g}")
    assert dt < 5e-4 and dr < 5e-4


def test_continued_runs_and_their_checks(traces):
    srv = server(traces)
    ref = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"]
    # the full run retired every session: a continuation has nothing to continue
    with pytest.raises(RuntimeError, match="no live session"):
        srv.run(N.DM_RUN_BATCHED, 2, NF - 2)
    a = srv.run(N.DM_RUN_BATCHED, 0, 2, want_poses=True)["poses"]
    with pytest.raises(RuntimeError, match="stopped"):
        srv.run(N.DM_RUN_PER_FRAME_DEV, 3, NF - 3)
    b = srv.run(N.DM_RUN_PER_FRAME_DEV, 2, NF - 2, want_poses=True)["poses"]
    srv.close()
    assert np.array_equal(np.concatenate([a, b], 1), ref)


def test_server_session_equals_standalone_session(traces):
    """A served trace and a standalone Session with the same parameters
    (including the ICP's per-level stop threshold) give bit-equal poses."""
    k = 4
    name, d, c, g = traces[k]
    w, t = small(name)
    assert w["icp"].get("min_increment", 0.0) > 0.0
    srv = server(traces, sel=[k])
    pb = srv.run(N.DM_RUN_BATCHED, 0, NF, want_poses=True)["poses"][0]
    srv.close()
    s = dm.Session(t, w["icp"], w["sensor"], g[0])
    ps = np.stack([s.process_frame(d[f], None if c is None else c[f]) for f in range(NF)])
    s.close()
    assert np.array_equal(pb, ps)
