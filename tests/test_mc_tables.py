"""The marching-cubes tables of the build (product library and oracle) are
Bourke's tables as the reference uses them
(include/geometry/mesh/meshExtractor/lookupTables.h): compared through the
digests and per-case triangle counts in tests/golden/mc_tables.json."""
import ctypes
import hashlib
import json
import os

import numpy as np

import oracle_lib as O
from dynamap_amd import _native as N

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mc_tables.json")))


def decode(words):
    rows = []
    for w in words:
        w = int(w)
        r = [(w >> (4 * k)) & 0xF for k in range(16)]
        r = [e if e != 0xF else -1 for e in r]
        k = r.index(-1) if -1 in r else 16
        rows.append(r[:k] + [-1] * (16 - k))
    return rows


def check(words):
    rows = decode(words)
    flat = [x for r in rows for x in r]
    assert hashlib.sha256(",".join(map(str, flat)).encode()).hexdigest() == GOLD["tri_sha256"]
    edge = [sum(1 << e for e in set(x for x in r if x >= 0)) for r in rows]
    assert hashlib.sha256(",".join(map(str, edge)).encode()).hexdigest() == GOLD["edge_sha256"]
    assert [sum(1 for x in r if x >= 0) // 3 for r in rows] == GOLD["triangles_per_case"]


def test_oracle_table():
    check(O.mc_table(32))


def test_product_table():
    lib = N.lib()
    out = np.zeros(256, np.uint64)
    assert lib.dm_mc_tri_table(out.ctypes.data_as(ctypes.c_void_p)) == 0
    check(out)
