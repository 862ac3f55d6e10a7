"""GPU block deletion and pool recycling (HashTable::DeleteBlock,
include/dataStructure/hashTable/hashTable.cu:112-134; Heap::Append,
include/dataStructure/heap/heap.cu:19-22) against the FP32 transcription:
the hash table slot for slot, the live block set and voxels, also after
later frames re-insert blocks into freed pool indices, and across a
snapshot / restore."""
import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import configs
import oracle_lib as O
from test_gpu_config_sizes import compare_tables

pytestmark = pytest.mark.gpu
W = configs.WORKLOADS["rigid-320x240-2cm"]
SENS = W["sensor"]
PROPS = dict(W["tsdf"], num_buckets=1 << 12, num_blocks=1 << 14)


def same_voxels(vol, ora):
    gx, gs, gc, gw = vol.blocks()
    ox, os_, oc, ow = ora.blocks()
    assert len(gx) == len(ox) == vol.GetNumAllocatedBlocks() == ora.num_alloc()
    gi, oi = np.lexsort(gx.T[::-1]), np.lexsort(ox.T[::-1])
    assert np.array_equal(gx[gi], ox[oi])
    assert np.array_equal(gw[gi], ow[oi])
    assert np.abs(gs[gi] - os_[oi].astype(np.float32)).max() <= 1e-6


@pytest.fixture(scope="module")
def seq():
    return configs.make_sequence("rigid-320x240-2cm", 8)


def test_delete_and_reinsert_match_transcription(seq):
    depth, _, gt = seq
    vol = dm.TsdfVolume(PROPS, SENS["rows"], SENS["cols"])
    ora = O.OracleVolume(PROPS, 32)
    for f in range(3):
        Tf = gt[f].astype(np.float32)
        vol.IntegrateScan(depth[f], None, SENS, Tf)
        ora.integrate_scan(depth[f], None, SENS, Tf.astype(np.float64))
    xyz = ora.blocks()[0]
    rng = np.random.default_rng(2)
    pick = xyz[rng.choice(len(xyz), len(xyz) // 3, replace=False)]
    req = np.vstack([pick, pick[:7], [[99999, 1, 1]]]).astype(np.int32)
    print(f"case: delete {len(pick)} of {len(xyz)} blocks", flush=True)
    assert vol.DeleteBlocks(req) == len(pick) == ora.delete_blocks(req)
    compare_tables(vol, ora, PROPS)
    same_voxels(vol, ora)
    assert vol.counters()["allocated"] == len(xyz) - len(pick)
    for f in range(3, 6):
        Tf = gt[f].astype(np.float32)
        vol.IntegrateScan(depth[f], None, SENS, Tf)
        ora.integrate_scan(depth[f], None, SENS, Tf.astype(np.float64))
        compare_tables(vol, ora, PROPS)
    same_voxels(vol, ora)


def test_recycle_snapshot_and_continue(seq):
    depth, _, gt = seq
    vol = dm.TsdfVolume(PROPS, SENS["rows"], SENS["cols"])
    ora = O.OracleVolume(PROPS, 32)
    for f in range(4):
        Tf = gt[f].astype(np.float32)
        vol.IntegrateScan(depth[f], None, SENS, Tf)
        ora.integrate_scan(depth[f], None, SENS, Tf.astype(np.float64))
    c = gt[3][:3, 3]
    nd = vol.Recycle(c, 1.5)
    print(f"case: recycle beyond 1.5 m: {nd} blocks", flush=True)
    assert nd == ora.recycle(c, 1.5) > 0
    compare_tables(vol, ora, PROPS)
    same_voxels(vol, ora)
    # a snapshot carries the free stack: the restored volume continues identically
    snap = vol.snapshot()
    vol2 = dm.TsdfVolume(PROPS, SENS["rows"], SENS["cols"])
    vol2.restore(snap)
    for f in range(4, 8):
        Tf = gt[f].astype(np.float32)
        for v in (vol, vol2):
            v.IntegrateScan(depth[f], None, SENS, Tf)
        ora.integrate_scan(depth[f], None, SENS, Tf.astype(np.float64))
    compare_tables(vol, ora, PROPS)
    compare_tables(vol2, ora, PROPS)
    same_voxels(vol, ora)
    a, b = vol.blocks(), vol2.blocks()
    ia, ib = np.lexsort(a[0].T[::-1]), np.lexsort(b[0].T[::-1])
    for x, y in zip(a, b):
        assert np.array_equal(x[ia], y[ib])
    # the raycast of the recycled volume only sees what is left
    d, _, _, _ = vol.Raycast(gt[7].astype(np.float32), SENS)
    assert (d > 0).mean() > 0.5
