"""Snapshot / restore of a session: after restoring into a fresh session,
the next frames give bit-identical poses and volumes to the original's."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs

pytestmark = pytest.mark.gpu


def test_restore_continues_bit_identically():
    name = "rigid-640x480-8mm"
    w = configs.WORKLOADS[name]
    depth, rgb, gt = configs.make_sequence(name, 10)
    props = dict(w["tsdf"], num_buckets=1 << 18, num_blocks=1 << 16)
    a = dm.Session(props, w["icp"], w["sensor"], gt[0])
    for f in range(6):
        a.process_frame(depth[f], rgb[f])
    snap = a.volume.snapshot()
    pa = np.stack([a.process_frame(depth[f], rgb[f]) for f in range(6, 10)])
    b = dm.Session(props, w["icp"], w["sensor"], np.eye(4))
    b.volume.restore(snap)
    pb = np.stack([b.process_frame(depth[f], rgb[f]) for f in range(6, 10)])
    assert np.array_equal(pa, pb)
    # pool indices are internal: compare the blocks by coordinates
    xa, sa, ca, wa = a.volume.blocks()
    xb, sb, cb, wb = b.volume.blocks()
    ia, ib = np.lexsort(xa.T[::-1]), np.lexsort(xb.T[::-1])
    assert np.array_equal(xa[ia], xb[ib])
    assert np.array_equal(sa[ia].view(np.uint32), sb[ib].view(np.uint32))
    assert np.array_equal(wa[ia], wb[ib]) and np.array_equal(ca[ia], cb[ib])
    # a volume with other properties refuses the snapshot
    c = dm.Session(dict(props, num_blocks=1 << 15), w["icp"], w["sensor"], np.eye(4))
    with pytest.raises(dm._native.DmError):
        c.volume.restore(snap)
    for s in (a, b, c):
        s.close()
