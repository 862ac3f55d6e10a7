#!/usr/bin/env python3
"""Serving benchmark: many concurrent rigid reconstruction sessions on one MI355X.

Workload "serve-16": 16 sessions, 8 x (320x240, 2 cm voxels, depth-only ICP)
+ 8 x (640x480, 8 mm voxels, geometric + photometric ICP), each a seeded
synthetic room scene and trajectory.  All frames are resident in HBM before
any timed region.

A step is one batch of the serving loop: the next frame of every session
(16 frames) through tracking (3-level ICP, 10/5/4 iterations), block
allocation, integration and raycasting.

Reported:
  value         aggregate frames/s of the continuously batched loop (closed
                loop: every frame available), timed over --steps batches after
                --warmup batches
  per_frame     the same frames through the per-frame path (one frame at a
                time, host 6x6 solves as the reference does; and with device
                solves), and the batched/per-frame ratio
  latency_ms    p50 / p99 frame latency with Poisson arrivals (30 Hz mean per
                session), batched and per-frame
  ate_mm        trajectory agreement: batched vs per-frame (same results
                alone and in the mix), GPU vs the FP64 CPU oracle on the
                CPU-baseline sample, GPU vs ground truth
  roofline      the dominant kernel's algorithmic bytes / measured duration
  cpu_baseline  the FP64 CPU oracle pipeline on a bounded sample
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=300, help="frames per session trace")
    ap.add_argument("--no-8mm-line", action="store_true", help="skip the 16 x 640x480 / 8 mm serving line")
    ap.add_argument("--sessions", type=int, default=16)
    ap.add_argument("--poisson-hz", type=float, default=30.0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-frame", action="store_true")
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--no-nonrigid", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--icp-reduce", choices=("shuffle", "mfma"), default="shuffle",
                    help="ICP linearisation kernel: shuffle-tree reduction or FP32 matrix-core F^T F")
    ap.add_argument("--nonrigid-frames", type=int, default=20)
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--no-stress", action="store_true", help="skip the 65-session stress mix")
    ap.add_argument("--stress-frames", type=int, default=24)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ate(A, B):
    return float(np.sqrt(np.mean(np.sum((A[..., :3, 3] - B[..., :3, 3]) ** 2, -1))))


def bench_nonrigid(dm, nframes, with_cpu, node_radius=0.0133, dgw=0.025):
    """Deforming-cylinder workload: 200 000-vertex canonical mesh, 640x480
    depth of a sinusoidally bending cylinder; one solve (5 Gauss-Newton
    iterations, block-Jacobi PCG) per frame.  The default graph is the
    workload's 4096-node reading: nodes on a 13.3 mm grid (4048 nodes, the
    grid rule's closest count) with a 2.5 cm weight radius."""
    N = dm._native
    S = N.synth()
    nt, ny = 500, 400
    V = np.zeros((nt * ny, 3), np.float32)
    Nn = np.zeros_like(V)
    S.dms_cylinder_mesh(nt, ny, N.ptr(V), N.ptr(Nn))
    sens = dict(rows=480, cols=640, fx=525.0, fy=525.0, cx=319.5, cy=239.5)
    poses = np.zeros((nframes, 16))
    S.dms_cylinder_trajectory(nframes, 20261202 + 3000, N.ptr(poses))
    frames = []
    for f in range(nframes):
        d = np.zeros((480, 640), np.float32)
        S.dms_cylinder_render(float(f), N.ptr(np.ascontiguousarray(poses[f])), 480, 640, 525.0, 525.0, 319.5,
                              239.5, 20261202 + 3000 + f, 1, N.ptr(d))
        frames.append(d)
    t0 = time.time()
    g = dm.DefGraph(V, Nn, node_radius, dgw, max_nodes=8192)
    t_graph = time.time() - t0
    solver = dm.NonRigidICP()
    solver.solve(g, frames[0], sens, poses[0].reshape(4, 4))  # capture
    g.set_transforms(np.tile(np.eye(4), (g.nodeNum, 1, 1)))
    t0 = time.time()
    for f in range(nframes):
        st = solver.solve(g, frames[f], sens, poses[f].reshape(4, 4))
    el = time.time() - t0
    vb, _ = dm.DqBlender(g).blendMesh()
    f = nframes - 1
    b = 0.05 * np.sin(2 * np.pi * f / 40) * ((0.3 - vb[:, 1]) / 0.6) ** 2
    err = np.abs(np.sqrt((vb[:, 0] - b) ** 2 + vb[:, 2] ** 2) - 0.15)[V[:, 2] < -0.07]
    res = dict(workload=f"bending cylinder, 200k-vertex mesh, 640x480, {g.nodeNum} nodes ({1000 * node_radius:g} mm "
                        f"grid, {1000 * dgw:g} mm weight radius), 5 GN iterations x 40 PCG per frame",
               frames=nframes, fps=nframes / el, ms_per_frame=1000 * el / nframes, nodes=g.nodeNum,
               blocks=g.num_blocks(), graph_build_ms=1000 * t_graph, surface_err_mm=1000 * float(err.mean()),
               last_stats=st)
    log(f"non-rigid: {res['fps']:.1f} frames/s ({res['ms_per_frame']:.2f} ms), nodes {g.nodeNum}, "
        f"surface err {res['surface_err_mm']:.4f} mm")
    # several deforming objects tracked at once: one graph each (same mesh and
    # graph rule), object k on frame (f + 10 k) of the sequence, all of a
    # round's solves enqueued together (dm_nonrigid_solve_many)
    K = 4
    gs = [dm.DefGraph(V, Nn, node_radius, dgw, max_nodes=8192) for _ in range(K)]
    pick = lambda f: [(f + 10 * k) % nframes for k in range(K)]
    solver.solve_many(gs, [frames[i] for i in pick(0)], sens, [poses[i].reshape(4, 4) for i in pick(0)])  # capture
    for gk in gs:
        gk.set_transforms(np.tile(np.eye(4), (gk.nodeNum, 1, 1)))
    t0 = time.time()
    for f in range(nframes):
        solver.solve_many(gs, [frames[i] for i in pick(f)], sens, [poses[i].reshape(4, 4) for i in pick(f)])
    elk = time.time() - t0
    res["concurrent"] = dict(graphs=K, rounds=nframes, fps_aggregate=K * nframes / elk,
                             ms_per_round=1000 * elk / nframes,
                             note="one solve per graph per round, enqueued on the graphs' own streams together")
    log(f"non-rigid, {K} graphs at once: {K * nframes / elk:.1f} frames/s aggregate "
        f"({1000 * elk / nframes:.2f} ms per round of {K})")
    # the same rounds through a batch: every kernel launched once for all K
    # graphs (graph = blockIdx.y), one captured HIP graph per round
    batch = solver.batch(gs)
    batch.solve([frames[i] for i in pick(0)], sens, [poses[i].reshape(4, 4) for i in pick(0)])  # capture
    for gk in gs:
        gk.set_transforms(np.tile(np.eye(4), (gk.nodeNum, 1, 1)))
    t0 = time.time()
    for f in range(nframes):
        batch.solve([frames[i] for i in pick(f)], sens, [poses[i].reshape(4, 4) for i in pick(f)])
    elb = time.time() - t0
    batch.close()
    for gk in gs:
        gk.close()
    res["batched"] = dict(graphs=K, rounds=nframes, fps_aggregate=K * nframes / elb,
                          ms_per_round=1000 * elb / nframes,
                          note="one solve per graph per round, each kernel launched once for all graphs")
    log(f"non-rigid, {K} graphs batched: {K * nframes / elb:.1f} frames/s aggregate "
        f"({1000 * elb / nframes:.2f} ms per round of {K})")
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        o = O.OracleGraph(V, Nn, node_radius, dgw, max_nodes=8192, bits=64)
        t0 = time.time()
        k = 0
        while k < nframes and (k < 2 or time.time() - t0 < 5.0):
            o.solve(frames[k], sens, poses[k].reshape(4, 4))
            k += 1
        res["cpu_baseline"] = dict(value=k / (time.time() - t0), unit="frames/s", kind="port",
                                   cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                                   sample=f"FP64 oracle non-rigid solve, frames 0..{k - 1}")
    g.close()
    return res


def bench_mesh(dm, with_cpu, reps=5):
    """Marching cubes of a 640x480 / 8 mm room volume (30 integrated frames),
    every allocated block meshed; the extraction call includes its host
    copies of positions, colours and edge ids."""
    from dynamap_amd import configs
    name = "rigid-640x480-8mm"
    w = configs.WORKLOADS[name]
    depth, rgb, gt = configs.make_sequence(name, 30, session=8)
    vol = dm.TsdfVolume(w["tsdf"], w["sensor"]["rows"], w["sensor"]["cols"])
    for f in range(30):
        vol.IntegrateScan(depth[f], None if rgb is None else rgb[f], w["sensor"], gt[f].astype(np.float32))
    m = vol.ExtractMesh(max_triangles=1 << 23)
    t0 = time.time()
    for _ in range(reps):
        m = vol.ExtractMesh(max_triangles=1 << 23)
    el = (time.time() - t0) / reps
    nb = vol.GetNumAllocatedBlocks()
    res = dict(workload="640x480 / 8 mm room, 30 frames integrated, all allocated blocks", blocks=nb,
               triangles=len(m), ms_per_extraction=1000 * el, triangles_per_s=len(m) / el,
               voxel_bytes=nb * 512 * 8)
    log(f"mesh: {len(m)} triangles from {nb} blocks in {1000 * el:.2f} ms per extraction (host copies included)")
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        ent = vol.GetHashEntries()
        _, gs, gc, gw = vol.blocks()
        occ = ent[ent[:, 3] >= 0]
        mir = O.OracleVolume(w["tsdf"], 32)
        mir.set_blocks(occ[:, :3], gs[occ[:, 3]].astype(np.float64), gw[occ[:, 3]], gc[occ[:, 3]])
        t0 = time.time()
        op, _, oe = O.mc_extract(mir)
        elc = time.time() - t0
        res["bit_exact_vs_oracle"] = bool(len(op) == len(m) and np.array_equal(oe, m.edge_ids)
                                          and np.array_equal(op.view(np.uint32), m.triangles.view(np.uint32)))
        L32 = O.load(32)
        L32.or_get_threads.restype = C.c_int
        res["cpu_baseline"] = dict(value=1000 * elc, unit="ms per extraction", kind="port",
                                   cores=int(L32.or_get_threads()), sample="FP32 oracle marching cubes, same volume")
        log(f"mesh: oracle {1000 * elc:.1f} ms, bit-exact {res['bit_exact_vs_oracle']}")
    vol.close()
    return res


def bench_stress(dm, nframes, poisson_hz, warm=3):
    """Large stress mix on one GPU: 32 x (320x240, 2 cm) + 32 x (640x480,
    8 mm) room sessions plus one 640x480 / 4 mm session of the 8 x 8 x 3 m
    multi-room scene (2^22 buckets, 2^21 blocks), all in one serving loop.
    Reports the closed-loop aggregate rate, the per-frame device path on the
    same frames, Poisson-arrival latencies and trajectory agreement."""
    from dynamap_amd import configs
    N = dm._native
    names = (["rigid-320x240-2cm"] * 32 + ["rigid-640x480-8mm"] * 32 + ["large-640x480-4mm"])
    t0 = time.time()
    srv = dm.Server(len(names), 480, 640, hbm_budget_bytes=250 << 30, use_graph=True)
    gts = []
    for s, name in enumerate(names):
        depth, rgb, gt = configs.make_sequence(name, nframes, session=s)
        w = configs.WORKLOADS[name]
        srv.add_trace(w["tsdf"], w["icp"], w["sensor"], gt[0], depth, rgb)
        gts.append(gt)
        del depth, rgb
    gts = np.stack(gts)
    hbm = sum(srv.trace_bytes(t) for t in range(len(names)))
    log(f"stress: {len(names)} sessions x {nframes} frames set up in {time.time() - t0:.1f}s, "
        f"{hbm / 2**30:.1f} GiB of session state (frames excluded)")
    ns, K = len(names), nframes - warm
    srv.run(N.DM_RUN_BATCHED, 0, warm)
    r = srv.run(N.DM_RUN_BATCHED, warm, K, want_poses=True)
    srv.run(N.DM_RUN_PER_FRAME_DEV, 0, warm)
    rd = srv.run(N.DM_RUN_PER_FRAME_DEV, warm, K, want_poses=True)
    rng = np.random.default_rng(configs.SEED_BASE + 5000)
    gaps = rng.exponential(1.0 / poisson_hz, size=(ns, nframes))
    arrivals = np.cumsum(gaps, axis=1) - gaps[:, :1]
    rp = srv.run(N.DM_RUN_BATCHED, 0, nframes, arrivals=arrivals)
    L = (rp["completions"] - arrivals) * 1000.0
    pb, pd = r["poses"], rd["poses"]
    res = dict(workload="stress-65: 32 x 320x240 @ 2 cm + 32 x 640x480 @ 8 mm + 1 x 640x480 @ 4 mm multi-room "
                        "(2^22 buckets)", sessions=ns, frames_per_session=nframes, timed_frames=K,
               session_state_gib=hbm / 2**30,
               batched_fps=ns * K / r["wall"], ms_per_batch=1000 * r["wall"] / K,
               per_frame_device_fps=ns * K / rd["wall"],
               poisson=dict(hz_per_session=poisson_hz, p50=float(np.percentile(L, 50)),
                            p99=float(np.percentile(L, 99)), max=float(L.max()),
                            large_session_p99=float(np.percentile(L[-1], 99)), batches=rp["batches"]),
               max_abs_pose_diff_batched_vs_per_frame_device=float(np.abs(pb - pd).max()),
               ate_mm_vs_ground_truth={"2cm": 1000 * ate(pb[:32], gts[:32, warm:]),
                                       "8mm": 1000 * ate(pb[32:64], gts[32:64, warm:]),
                                       "4mm_large": 1000 * ate(pb[64], gts[64, warm:])})
    res["ratio_batched_vs_per_frame_device"] = res["batched_fps"] / res["per_frame_device_fps"]
    log(f"stress: batched {res['batched_fps']:.1f} frames/s ({res['ms_per_batch']:.2f} ms/batch), per-frame device "
        f"{res['per_frame_device_fps']:.1f} frames/s, poisson p50 {res['poisson']['p50']:.2f} p99 "
        f"{res['poisson']['p99']:.2f} ms")
    srv.close()
    return res


def serve_line(dm, traces, nf, K, W, args, label):
    """One serving workload on a fresh server: every trace has nf frames.
    Frames 0 .. nf-K-W-1 build each session's map (untimed), W warmup
    batches follow, then K timed batches end at the trace's last frame, so
    the timed frames see the map at its steady state.  The per-frame paths
    run the same frames after the same untimed history; Poisson arrivals
    (args.poisson_hz per session) cover all nf frames of fresh sessions."""
    from dynamap_amd import configs
    N = dm._native
    ns = len(traces)
    F0 = nf - K - W
    if F0 < 0:
        raise SystemExit(f"--steps + --warmup must not exceed the {nf} frames of a trace")
    srv = dm.Server(ns, 480, 640, hbm_budget_bytes=250 << 30, use_graph=True)
    for name, depth, rgb, gt in traces:
        w = configs.WORKLOADS[name]
        srv.add_trace(w["tsdf"], w["icp"], w["sensor"], gt[0], depth, rgb)
    out = {}
    # launch shapes chosen by the autotuner on this workload's steady state
    tuning = "defaults (16x16 raycast tiles, 2048 integration workgroups per session)"
    shapes = dict(ray_tile_width=16, integrate_workgroups=2048)
    if not args.no_autotune:
        srv.run(N.DM_RUN_BATCHED, 0, F0 + W)
        tuning = srv.autotune(frame=F0 + W, reps=3)
        srv.reset()
        log(label, "autotune", tuning)
        shapes = dict(ray_tile_width=int(tuning["ray_tile"].split("x")[0]),
                      integrate_workgroups=tuning["integrate_workgroups"])
    if args.icp_reduce != "shuffle":
        srv.set_shapes(icp_reduce=args.icp_reduce, **shapes)
        tuning = dict(tuning, icp_reduce=args.icp_reduce) if isinstance(tuning, dict) else \
            f"{tuning}, ICP on the matrix cores"
    # batched closed loop
    srv.run(N.DM_RUN_BATCHED, 0, F0 + W)
    r = srv.run(N.DM_RUN_BATCHED, F0 + W, K, want_poses=True)
    wall_b = r["wall"]
    fps_b = ns * K / wall_b
    poses_b = r["poses"]
    log(f"{label} batched: {fps_b:.1f} frames/s, {r['batches']} batches, {1000 * wall_b / K:.3f} ms/batch "
        f"(frames {F0 + W}..{nf - 1})")
    out.update(fps=fps_b, wall=wall_b, tuning=tuning, timed_frames=[F0 + W, nf - 1], untimed_history=F0,
               ms_per_batch=1000 * wall_b / K)
    if not args.no_per_frame:
        per = {}
        for mode, key in ((N.DM_RUN_PER_FRAME_HOST, "host_solve"), (N.DM_RUN_PER_FRAME_DEV, "device_solve")):
            srv.reset()
            srv.run(N.DM_RUN_BATCHED, 0, F0)
            srv.run(mode, F0, W)
            rp = srv.run(mode, F0 + W, K, want_poses=True)
            per[key] = dict(fps=ns * K / rp["wall"], ms_per_frame=1000 * rp["wall"] / (ns * K))
            per[key]["poses"] = rp["poses"]
            log(f"{label} per-frame {key}: {per[key]['fps']:.1f} frames/s")
        out["per_frame"] = {k: {kk: vv for kk, vv in v.items() if kk != "poses"} for k, v in per.items()}
        out["per_frame"]["ratio_vs_host_solve"] = fps_b / per["host_solve"]["fps"]
        out["per_frame"]["ratio_vs_device_solve"] = fps_b / per["device_solve"]["fps"]
        out["ate_mm"] = dict(
            batched_vs_per_frame_device=1000 * ate(poses_b, per["device_solve"]["poses"]),
            batched_vs_per_frame_host=1000 * ate(poses_b, per["host_solve"]["poses"]),
            max_abs_pose_diff_batched_vs_per_frame_device=float(np.abs(poses_b - per["device_solve"]["poses"]).max()),
            max_abs_pose_diff_batched_vs_per_frame_host=float(np.abs(poses_b - per["host_solve"]["poses"]).max()))
    gt_all = np.stack([t[3][F0 + W:nf] for t in traces])
    out.setdefault("ate_mm", {})["gpu_vs_ground_truth"] = 1000 * ate(poses_b, gt_all)
    if not args.no_poisson:
        rng = np.random.default_rng(configs.SEED_BASE + 4000)
        gaps = rng.exponential(1.0 / args.poisson_hz, size=(ns, nf))
        arrivals = np.cumsum(gaps, axis=1) - gaps[:, :1]
        lat = {}
        for mode, key in ((N.DM_RUN_BATCHED, "batched"), (N.DM_RUN_PER_FRAME_HOST, "per_frame_host_solve")):
            srv.reset()
            rr = srv.run(mode, 0, nf, arrivals=arrivals)
            L = (rr["completions"] - arrivals) * 1000.0
            lat[key] = dict(p50=float(np.percentile(L, 50)), p99=float(np.percentile(L, 99)),
                            max=float(L.max()), served_fps=ns * nf / rr["wall"], batches=rr["batches"],
                            samples=int(L.size))
            log(f"{label} poisson {key}: p50 {lat[key]['p50']:.2f} ms p99 {lat[key]['p99']:.2f} ms "
                f"over {L.size} frames")
        out["latency_ms"] = lat
    return srv, out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sessions are independent and "
              f"would shard by session across GPUs, which this build does not launch", file=sys.stderr)
        sys.exit(2)
    import dynamap_amd as dm
    from dynamap_amd import configs

    W, K = args.warmup, args.steps
    nf = args.frames
    ns = args.sessions
    names = ["rigid-320x240-2cm"] * (ns // 2) + ["rigid-640x480-8mm"] * (ns - ns // 2)
    t0 = time.time()
    traces = []
    for s, name in enumerate(names):
        depth, rgb, gt = configs.make_sequence(name, nf, session=s)
        traces.append((name, depth, rgb, gt))
    log(f"generated {ns} x {nf} frames in {time.time() - t0:.1f}s")

    srv, line = serve_line(dm, traces, nf, K, W, args, "serve-16")
    fps_b, wall_b, tuning = line.pop("fps"), line.pop("wall"), line.pop("tuning")
    out = dict(line)
    # stage profile of one batch at the traces' last frame (steady state)
    srv.reset()
    srv.run(dm._native.DM_RUN_BATCHED, 0, nf - 1)
    stage_ms, work = srv.profile(nf - 1, reps=5)
    log("stage ms", stage_ms, work)

    # ---- roofline of the dominant kernel
    stage_kernels = {"integrate": stage_ms["integrate"], "raycast": stage_ms["raycast"],
                     "alloc_rays": stage_ms["alloc_rays"]}
    dom = max(stage_kernels, key=stage_kernels.get)
    P = work["pixels"]
    # pixels with a colour image (the 8 mm sessions): the traversal reads
    # their colour and writes every pixel's surface distance and colour (8 B)
    # for the integration
    P_rgb = sum(t[1].shape[1] * t[1].shape[2] for t in traces if t[2] is not None)
    nvis = work["visible_blocks"]
    vb = 8
    if dom == "integrate":
        bytes_ = nvis * 512 * 2 * vb + P * (4 + 3)
    elif dom == "raycast":
        bytes_ = nvis * 512 * 4 + P * (4 + 12 + 12 + 3)
    else:
        bytes_ = 4 * P + 3 * P_rgb + 8 * P + 16 * work["unique_blocks"]
    t_s = stage_kernels[dom] * 1e-3
    achieved = bytes_ / t_s / 1e9
    roofline = dict(bound="hbm", kernel=dom, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=None, algorithmic_bytes=int(bytes_),
                    launch_ms=stage_kernels[dom], vb=vb, Nvis=nvis, pixels=P,
                    stage_ms={k: round(v, 4) for k, v in stage_ms.items()})
    # every stage against the same roofline (algorithmic bytes of one batched
    # launch over all sessions / its HIP-event duration)
    alg = {"alloc_all": 4 * P + 3 * P_rgb + 8 * P + 16 * work["unique_blocks"] + 16 * work["new_blocks"],
           "integrate": nvis * 512 * 2 * vb + P * (4 + 3),
           "raycast": nvis * 512 * 4 + P * (4 + 12 + 12 + 3),
           "icp_l0": 64 * P}
    roofline["stages"] = {k: dict(ms=round(stage_ms[k], 4), algorithmic_bytes=int(b),
                                  achieved_gbs=round(b / (stage_ms[k] * 1e-3) / 1e9, 1),
                                  frac=round(b / (stage_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
                          for k, b in alg.items()}
    roofline["Nuniq"], roofline["Nalloc"] = work["unique_blocks"], work["new_blocks"]
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_serve16.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")

    # ---- CPU baseline: FP64 oracle pipeline on a bounded sample of one 640x480 session
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        s_idx = ns // 2
        name, depth, rgb, gt = traces[s_idx]
        w = configs.WORKLOADS[name]
        L64 = O.load(64)
        L64.or_get_threads.restype = C.c_int
        cores = int(L64.or_get_threads())

        def oracle_run(threads, budget_s, min_frames):
            L64.or_set_threads(threads)
            pl = O.OraclePipeline(w["tsdf"], w["icp"], w["sensor"], gt[0], bits=64)
            t1 = time.time()
            ps = []
            while len(ps) < nf and (time.time() - t1 < budget_s or len(ps) < min_frames):
                k = len(ps)
                ps.append(pl.process_frame(depth[k], rgb[k] if rgb is not None else None))
            return ps, time.time() - t1

        poses_o, el = oracle_run(cores, args.cpu_seconds, 3)
        f = len(poses_o)
        p1, el1 = oracle_run(1, min(10.0, args.cpu_seconds), 2)
        L64.or_set_threads(cores)
        cpu = dict(value=f / el, unit="frames/s", cores=cores, kind="port",
                   sample=f"FP64 oracle pipeline (track, integrate, raycast), session {s_idx} "
                          f"(640x480, 8 mm), frames 0..{f - 1}, {el:.1f} s",
                   value_1_thread=len(p1) / el1, sample_1_thread=f"frames 0..{len(p1) - 1}, {el1:.1f} s")
        # GPU (per-frame device path) poses for the same frames vs the oracle
        s1 = dm.Session(w["tsdf"], w["icp"], w["sensor"], gt[0])
        g = np.stack([s1.process_frame(depth[k], rgb[k]) for k in range(f)])
        s1.close()
        out["ate_mm"]["gpu_vs_fp64_oracle"] = 1000 * ate(g, np.stack(poses_o))
        out["ate_mm"]["fp64_oracle_frames"] = f

    # ---- non-rigid workload: bending cylinder, 200k-vertex mesh, 640x480
    if not args.no_nonrigid:
        out["nonrigid"] = bench_nonrigid(dm, args.nonrigid_frames, not args.no_cpu)
    # ---- marching cubes of one room volume
    if not args.no_mesh:
        out["mesh"] = bench_mesh(dm, not args.no_cpu)
    srv.close()
    # ---- the same protocol on 16 sessions that are all 640x480 / 8 mm (the
    # workload's 8 such sessions plus 8 more seeded ones)
    if not args.no_8mm_line:
        t0 = time.time()
        t8 = [t for t in traces if t[0] == "rigid-640x480-8mm"]
        for s in range(ns, ns + 16 - len(t8)):
            depth, rgb, gt = configs.make_sequence("rigid-640x480-8mm", nf, session=s)
            t8.append(("rigid-640x480-8mm", depth, rgb, gt))
        log(f"8 mm line: {len(t8)} traces ready in {time.time() - t0:.1f}s")
        srv8, line8 = serve_line(dm, t8, nf, K, W, args, "serve-16-8mm")
        srv8.close()
        del t8
        line8.pop("wall")
        fps8, tune8 = line8.pop("fps"), line8.pop("tuning")
        out["serve16_8mm"] = dict(workload="16 x 640x480 @ 8 mm (2^20 buckets), geometric + photometric ICP",
                                  batched_fps=fps8, tuning=tune8, **line8)
    # ---- 65-session stress mix, after the serve-16 sessions are released
    if not args.no_stress:
        out["stress"] = bench_stress(dm, args.stress_frames, args.poisson_hz)

    res = {
        "metric": "aggregate frames/s + p50/p99 frame latency, batched vs per-frame, ATE≤1mm",
        "value": fps_b,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1000 * wall_b / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded analytic RGB-D room scenes, noise + dropout)",
        "config": {"workload": "serve-16: 8 x 320x240 @ 2 cm + 8 x 640x480 @ 8 mm", "sessions": ns,
                   "frames_per_session": nf, "timed_frames": out.pop("timed_frames"),
                   "untimed_history_frames": out.pop("untimed_history"), "icp_iterations": [10, 5, 4],
                   "batch": "one frame per session", "tuning": tuning},
    }
    res.update(out)
    res["roofline"] = roofline
    res["cpu_baseline"] = cpu
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
