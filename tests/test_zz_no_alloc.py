"""dm_mesh_extract_noalloc (no reference counterpart: marching cubes for
callers that cannot allocate device memory) against dm_mesh_extract, the
entry point it shares its kernels with.  Two calls in a row on one device,
the second on a smaller volume, so stale scratch from the first call would
show; plus a box and a triangle cap.  Triangle count, order, positions,
colours and edge ids must be bit-identical."""
import ctypes as C

import numpy as np
import pytest

import dynamap_amd as dm
from dynamap_amd import _native as N
from dynamap_amd import configs

pytestmark = pytest.mark.gpu

CHUNK_BLOCKS = 1 << 14  # blocks per chunk of the no-alloc extraction


class DeviceArray:
    """Device buffer from the HIP runtime the library itself links."""

    def __init__(self, nbytes):
        N.lib()  # resolves libamdhip64.so.7 for the library first
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.p = C.c_void_p()
        self.nbytes = max(int(nbytes), 16)
        assert self.rt.hipMalloc(C.byref(self.p), self.nbytes) == 0
        # a fill pattern that no extraction writes, so untouched entries show
        assert self.rt.hipMemset(self.p, 0xA5, self.nbytes) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        if count:
            assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        if self.p:
            self.rt.hipFree(self.p)
            self.p = C.c_void_p()


def extract_noalloc(vol, cap, box=None):
    lib = N.lib()
    lo = hi = None
    if box is not None:
        lo = np.ascontiguousarray(box[0], np.float64)
        hi = np.ascontiguousarray(box[1], np.float64)
    n = C.c_int64()
    # count first (no outputs), then extract into buffers of that size
    N.check(lib.dm_mesh_extract_noalloc(vol._h, N.ptr(lo), N.ptr(hi), 0, None, None, None, C.byref(n)))
    total = n.value
    m = min(total, cap)
    pos, col, eid = DeviceArray(m * 36), DeviceArray(m * 36), DeviceArray(m * 24)
    try:
        N.check(lib.dm_mesh_extract_noalloc(vol._h, N.ptr(lo), N.ptr(hi), cap, pos.p, col.p, eid.p, C.byref(n)))
        assert n.value == total
        return (pos.get(np.float32, m * 9).reshape(m, 3, 3), col.get(np.float32, m * 9).reshape(m, 3, 3),
                eid.get(np.uint64, m * 3).reshape(m, 3), total)
    finally:
        for b in (pos, col, eid):
            b.free()


def check_equal(vol, cap, box=None):
    ref = vol.ExtractMesh(None if box is None else box[0], None if box is None else box[1], max_triangles=cap)
    pos, col, eid, total = extract_noalloc(vol, cap, box)
    print(f"  no-alloc extraction: {total} triangles ({len(pos)} written), cap {cap}", flush=True)
    assert total == ref.num_triangles_
    assert len(pos) == len(ref.triangles)
    assert np.array_equal(eid, ref.edge_ids), "edge ids / triangle order differ"
    assert np.array_equal(pos.view(np.uint32), ref.triangles.view(np.uint32)), "positions differ"
    assert np.array_equal(col.view(np.uint32), ref.colors.view(np.uint32)), "colours differ"
    return total


def integrate(name, props, nframes):
    w = configs.WORKLOADS[name]
    depth, rgb, gt = configs.make_sequence(name, nframes)
    vol = dm.TsdfVolume(props, w["sensor"]["rows"], w["sensor"]["cols"])
    for f in range(nframes):
        vol.IntegrateScan(depth[f], None if rgb is None else rgb[f], w["sensor"], gt[f].astype(np.float32))
    return vol


def test_noalloc_matches_allocating_extraction():
    # first call: 640x480 room at 4 mm voxels (2^20 buckets), several chunks
    props = dict(configs.TSDF_4MM, num_buckets=1 << 20, num_blocks=1 << 20)
    print("case: 640x480 room, 4 mm voxels, 2^20 buckets, 4 frames", flush=True)
    big = integrate("rigid-640x480-8mm", props, 4)
    nblk = big.GetNumAllocatedBlocks()
    print(f"  {nblk} allocated blocks", flush=True)
    assert nblk > 2 * CHUNK_BLOCKS
    t_big = check_equal(big, 1 << 24)
    assert t_big > 0
    # second call, smaller input: the 320x240 room at 2 cm (one chunk)
    print("case: 320x240 room, 2 cm voxels, 3 frames", flush=True)
    small = integrate("rigid-320x240-2cm", configs.TSDF_2CM, 3)
    assert small.GetNumAllocatedBlocks() < CHUNK_BLOCKS
    t_small = check_equal(small, 1 << 22)
    assert 0 < t_small < t_big
    # a box and a cap below the total, on the big volume again
    print("case: box and cap", flush=True)
    check_equal(big, 1 << 24, box=([-1.0, -1.0, 0.0], [1.0, 0.5, 2.5]))
    check_equal(big, t_big // 3)


def test_noalloc_rejects_too_many_buckets():
    props = dict(configs.TSDF_2CM, num_buckets=1 << 21, num_blocks=1 << 10)
    vol = dm.TsdfVolume(props, 240, 320)
    n = C.c_int64()
    rc = N.lib().dm_mesh_extract_noalloc(vol._h, None, None, 0, None, None, None, C.byref(n))
    assert rc == N.DM_E_CAPACITY
    assert b"2^20" in N.lib().dm_last_error()
