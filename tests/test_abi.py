"""The C ABI library loads without a GPU and exports every function that
include/*.h declares; argument validation fails cleanly before any device
work; the C++ class header compiles against the C ABI."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

import dynamap_amd as dm
from dynamap_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names |= set(re.findall(r"DM_API\s+[\w\s\*]+?\b(dm_\w+)\s*\(", txt))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(N.HIP_LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python bindings cover all of them
    assert set(names) <= set(N.PROTOTYPES), set(names) - set(N.PROTOTYPES)


def test_invalid_arguments_rejected_without_device():
    lib = N.lib()
    p = N.dm_tsdf_props()
    p.voxel_size, p.num_buckets, p.bucket_size, p.num_blocks, p.block_size = 0.01, 16, 10, 16, 4
    p.max_sdf_weight, p.truncation_distance, p.max_sensor_depth, p.min_sensor_depth = 64, 0.03, 5.0, 0.3
    h = ctypes.c_void_p()
    assert lib.dm_volume_create(ctypes.byref(p), 240, 320, ctypes.byref(h)) == N.DM_E_INVALID
    assert b"block_size" in lib.dm_last_error()
    assert lib.dm_volume_create(None, 240, 320, ctypes.byref(h)) == N.DM_E_INVALID
    with pytest.raises(N.DmError):
        dm.TsdfVolume(dict(voxel_size=-1.0, num_buckets=16, bucket_size=10, num_blocks=16, block_size=8,
                           max_sdf_weight=64, truncation_distance=0.03, max_sensor_depth=5.0,
                           min_sensor_depth=0.3), 10, 10)


def test_device_count_without_gpu():
    n = ctypes.c_int32(-1)
    assert N.lib().dm_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_cpp_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "dynamap/dynamap.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
