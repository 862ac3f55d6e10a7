"""Large-volume case: one session at 4 mm voxels in the 8 x 8 x 3 m
multi-room scene (2^22 buckets), tracked on the GPU; the full-volume
marching cubes is bit-exact (count, order, edge ids, positions) against the
FP32 transcription run on the same blocks.

Reference: include/geometry/mesh/meshExtractor/meshExtractor.cu:22-242,
include/builder/builder.cu:70-139.
"""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O
from test_gpu_mesh import check_same, mirror

pytestmark = pytest.mark.gpu


def test_large_multiroom_mesh_bit_exact():
    name = "large-640x480-4mm"
    w = configs.WORKLOADS[name]
    nf = 16
    depth, rgb, gt = configs.make_sequence(name, nf)
    s = dm.Session(w["tsdf"], w["icp"], w["sensor"], gt[0])
    poses = np.stack([s.process_frame(depth[f], rgb[f]) for f in range(nf)])
    ate = np.sqrt(np.mean(np.sum((poses[:, :3, 3] - gt[:, :3, 3]) ** 2, 1)))
    vol = s.volume
    n = vol.GetNumAllocatedBlocks()
    print("blocks", n, "ate mm", 1000 * ate, "counters", vol.counters())
    assert ate < 5e-3 and n > 10000
    m = vol.ExtractMesh(max_triangles=1 << 26)
    assert m.num_triangles_ == len(m) > 100000
    check_same((m.triangles, m.colors, m.edge_ids), O.mc_extract(mirror(vol)))
    s.close()


def test_large_1000_frames():
    """The large session at its full 1000 frames: every key fits (no drop),
    the first frames agree with the FP64 oracle within 1 mm ATE, the
    full-volume marching cubes after frame 300 is bit-exact against the FP32
    transcription run on all of the GPU volume's blocks, the full-volume mesh
    after frame 1000 fits the extractor, and marching cubes over a 1.5 m box
    around the first view's surface is bit-exact there too.  The trajectory's
    drift against the ground truth (about 3 cm mean after 1000 frames) is
    the oracle's too: the depth noise at 4 m (sigma 2.6 cm) exceeds the
    2 cm truncation of this workload."""
    name = "large-640x480-4mm"
    w = configs.WORKLOADS[name]
    nf, no, nm = 1000, 10, 300
    depth, rgb, gt = configs.make_sequence(name, nf)
    s = dm.Session(w["tsdf"], w["icp"], w["sensor"], gt[0])
    G = [s.process_frame(depth[f], rgb[f]) for f in range(nm)]
    vol = s.volume
    m = vol.ExtractMesh(max_triangles=1 << 26)
    print(f"after {nm} frames: blocks {vol.GetNumAllocatedBlocks()}, triangles {len(m)}", flush=True)
    assert m.num_triangles_ == len(m) > 100000
    check_same((m.triangles, m.colors, m.edge_ids), O.mc_extract(mirror(vol)))
    del m
    G = np.stack(G + [s.process_frame(depth[f], rgb[f]) for f in range(nm, nf)])
    cnt = vol.counters()
    print("counters", cnt, flush=True)
    assert cnt["dropped_bucket"] == 0 and cnt["dropped_pool"] == 0 and cnt["dropped_keyset"] == 0
    assert np.all(np.isfinite(G))
    ora = O.OraclePipeline(w["tsdf"], w["icp"], w["sensor"], gt[0], bits=64)
    R = np.stack([ora.process_frame(depth[f], rgb[f]) for f in range(no)])
    a = float(np.sqrt(np.mean(np.sum((G[:no, :3, 3] - R[:, :3, 3]) ** 2, 1))))
    print("ATE gpu-vs-fp64 over the first", no, "frames (mm)", 1000 * a, flush=True)
    assert a < 1e-3
    m = vol.ExtractMesh(max_triangles=1 << 25)
    print("full-volume triangles", m.num_triangles_, flush=True)
    assert m.num_triangles_ == len(m) > 1000000
    del m
    # box around the surface seen at the centre of the first frame
    S = w["sensor"]
    r, c = S["rows"] // 2, S["cols"] // 2
    win = depth[0][r - 10:r + 11, c - 10:c + 11]
    d = float(np.median(win[(win > 0.3) & (win < 5.0)]))
    pc = np.array([(c - S["cx"]) * d / S["fx"], (r - S["cy"]) * d / S["fy"], d, 1.0])
    p = (gt[0] @ pc)[:3]
    lo, hi = p - 0.75, p + 0.75
    gm = vol.ExtractMesh(lo, hi, max_triangles=1 << 23)
    assert gm.num_triangles_ == len(gm) > 10000
    # the blocks the box's cells read (one block of margin), in table order
    ent = vol.GetHashEntries()
    occ = ent[ent[:, 3] >= 0]
    bs = w["tsdf"]["voxel_size"] * 8
    blo, bhi = np.floor(lo / bs) - 1, np.floor(hi / bs) + 1
    sel = occ[np.all((occ[:, :3] >= blo) & (occ[:, :3] <= bhi), 1)]
    xyz, sdf, col, wt = vol.blocks()
    ov = O.OracleVolume(w["tsdf"], 32)
    assert ov.set_blocks(sel[:, :3], sdf[sel[:, 3]].astype(np.float64), wt[sel[:, 3]], col[sel[:, 3]]) == 0
    del xyz, sdf, col, wt
    print("box triangles", len(gm), "blocks mirrored", len(sel), flush=True)
    check_same((gm.triangles, gm.colors, gm.edge_ids), O.mc_extract(ov, (lo, hi)))
    s.close()
