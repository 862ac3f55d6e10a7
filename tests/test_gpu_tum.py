"""A TUM RGB-D sequence through the GPU session end to end
(include/utils/utils_Dataset.cpp:58-247 reading, builder::buildModelLS per
frame, builder::writePoseToTxt include/builder/builder.cu:243-262 writing):
a synthetic 640x480 RGB-D sequence is written in the TUM layout (16-bit
depth PNGs at 1/5000 m, rgb PNGs, associations, ground truth), read back,
tracked and fused, and its trajectory written as TUM lines.  The poses equal
those of the same quantised frames processed from memory, bit for bit, and
ATE / RPE against groundtruth.txt stay within the synthetic bound."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
from dynamap_amd import io as dio

pytestmark = pytest.mark.gpu


def test_tum_sequence_end_to_end(tmp_path):
    name, n = "rigid-640x480-8mm", 30
    w = configs.WORKLOADS[name]
    depth, rgb, gt = configs.make_sequence(name, n)
    root = str(tmp_path / "seq")
    print(f"case: {name}, {n} frames written in the TUM layout", flush=True)
    dio.write_tum_sequence(root, depth, rgb, gt)
    ds = dio.TumDataset(root)
    assert len(ds) == n
    t0 = ds.timestamps()[0]
    T0 = ds.groundtruth(t0)
    traj = str(tmp_path / "estimate.txt")
    s1 = dm.Session(w["tsdf"], w["icp"], w["sensor"], T0)
    P1 = s1.run_dataset(ds, trajectory=traj)
    s1.close()
    # the same frames from memory, quantised as the 16-bit PNGs hold them
    q = np.clip(np.round(depth.astype(np.float64) * 5000.0), 0, 65535).astype(np.uint16)
    s2 = dm.Session(w["tsdf"], w["icp"], w["sensor"], T0)
    P2 = np.stack([s2.process_frame(q[k].astype(np.float32) / 5000.0, rgb[k]) for k in range(n)])
    s2.close()
    assert np.array_equal(P1.view(np.uint64), P2.view(np.uint64)), "dataset and in-memory poses differ"
    # the written trajectory: one TUM line per frame, 4 decimals
    st, est = dio.read_trajectory(traj)
    assert len(st) == n and np.abs(st - ds.timestamps()).max() < 1e-4
    assert np.abs(est[:, :3, 3] - P1[:, :3, 3]).max() <= 5e-5 + 1e-9
    gst, gtp = dio.read_trajectory(str(tmp_path / "seq" / "groundtruth.txt"))
    a = dio.ate(P1, gtp, align=False)
    r_t, r_r = dio.rpe(P1, gtp)
    print(f"ATE {1000 * a:.3f} mm, RPE {1000 * r_t:.3f} mm / {np.degrees(r_r):.4f} deg per frame")
    assert a < 3e-3 and r_t < 2e-3
