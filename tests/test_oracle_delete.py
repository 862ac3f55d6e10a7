"""Pins the oracle's block deletion and pool recycling
(HashTable::DeleteBlock, include/dataStructure/hashTable/hashTable.cu:112-134;
Heap::Append, include/dataStructure/heap/heap.cu:19-22) against an
independent restatement of the bucket rule written here: every bucket keeps
its occupied slots as a prefix, a deletion closes the bucket up in order, and
freed pool indices are reused before the pool grows."""
import numpy as np

import oracle_lib as O
from dynamap_amd import configs

W = configs.WORKLOADS["rigid-320x240-2cm"]
SENS = W["sensor"]
PROPS = dict(W["tsdf"], num_buckets=1 << 12, num_blocks=1 << 14)


def keys_of(xyz):
    M = (1 << 21) - 1
    x = xyz.astype(np.int64) & M
    u = np.uint64
    return (u(1) << u(63)) | (x[:, 0].astype(u) << u(42)) | (x[:, 1].astype(u) << u(21)) | x[:, 2].astype(u)


def expected_after_delete(keys, counts, S, gone):
    """Stable removal of the keys in `gone` from each bucket's prefix."""
    k = keys.reshape(-1, S).copy()
    c = counts.copy()
    gone = set(int(g) for g in gone)
    for b in np.nonzero(c)[0]:
        row = [int(x) for x in k[b, :c[b]] if int(x) not in gone]
        k[b, :] = 0
        k[b, :len(row)] = row
        c[b] = len(row)
    return k.reshape(-1), c


def filled_volume(nframes=3):
    depth, _, gt = configs.make_sequence("rigid-320x240-2cm", nframes)
    ora = O.OracleVolume(PROPS, 32)
    for f in range(nframes):
        ora.integrate_scan(depth[f], None, SENS, gt[f])
    return ora, depth, gt


def test_delete_closes_buckets_and_frees_pool():
    ora, depth, gt = filled_volume()
    keys0, counts0 = ora.table()
    xyz, sdf, rgb, w = ora.blocks()
    n0 = len(xyz)
    assert n0 > 500 and counts0.max() >= 3  # buckets with several entries
    rng = np.random.default_rng(1)
    pick = xyz[rng.choice(n0, n0 // 3, replace=False)]
    # absent blocks and repeats are skipped
    req = np.vstack([pick, pick[:10], [[100000, 5, 5], [123, -456, 789]]]).astype(np.int32)
    assert ora.delete_blocks(req) == len(pick)
    keys1, counts1 = ora.table()
    ek, ec = expected_after_delete(keys0, counts0, PROPS["bucket_size"], keys_of(pick))
    assert np.array_equal(counts1, ec) and np.array_equal(keys1, ek)
    assert ora.num_alloc() == n0 - len(pick)
    # survivors keep their voxels; deleted blocks are gone
    xyz1, sdf1, _, w1 = ora.blocks()
    before = {tuple(k): i for i, k in enumerate(xyz)}
    for j in range(0, len(xyz1), 37):
        i = before[tuple(xyz1[j])]
        assert np.array_equal(w1[j], w[i]) and np.array_equal(sdf1[j], sdf[i])
    assert not set(map(tuple, pick)) & set(map(tuple, xyz1))
    # a later frame re-inserts what its rays cross, into freed pool indices
    # first: the table equals a fresh volume's after the same frames...
    ora.integrate_scan(depth[2], None, SENS, gt[2])
    assert ora.num_alloc() > n0 - len(pick)
    # ...and the pool high-water mark did not grow while freed indices remained
    fresh = O.OracleVolume(PROPS, 32)
    for f in range(3):
        fresh.allocate(depth[f], SENS, gt[f])
    assert ora.num_alloc() <= fresh.num_alloc()


def test_recycle_selects_by_block_centre():
    ora, depth, gt = filled_volume()
    xyz, _, _, _ = ora.blocks()
    c = gt[2][:3, 3]
    r = 1.5
    bm = 8 * PROPS["voxel_size"]
    cen = (xyz.astype(np.float64) + 0.5) * bm
    far = np.linalg.norm(cen - c, axis=1) > r
    assert 0 < far.sum() < len(xyz)
    keys0, counts0 = ora.table()
    assert ora.recycle(c, r) == far.sum()
    keys1, counts1 = ora.table()
    ek, ec = expected_after_delete(keys0, counts0, PROPS["bucket_size"], keys_of(xyz[far]))
    assert np.array_equal(keys1, ek) and np.array_equal(counts1, ec)
    left = ora.blocks()[0]
    assert np.array_equal(np.sort(keys_of(left)), np.sort(keys_of(xyz[~far])))
    assert ora.recycle(c, r) == 0  # idempotent
