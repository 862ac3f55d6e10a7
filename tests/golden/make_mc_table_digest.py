"""Writes tests/golden/mc_tables.json: digests and per-case triangle counts of
the marching-cubes tables the reference uses
(include/geometry/mesh/meshExtractor/lookupTables.h, Paul Bourke's published
edge and triangle tables), read here as text.  The build's own tables are
checked against these values; the reference file itself is not copied.
Run: python tests/golden/make_mc_table_digest.py
"""
import hashlib
import json
import os
import re

SRC = "/root/reference/include/geometry/mesh/meshExtractor/lookupTables.h"
txt = open(SRC).read()
tri = [int(x) for x in re.findall(r"-?\d+", re.search(r"triTable\[256\]\[16\]\s*=\s*\{(.*?)\};", txt, re.S).group(1))]
edge = [int(x, 16) for x in re.findall(r"0x[0-9a-fA-F]+", re.search(r"edgeTable\[256\]\s*=\s*\{(.*?)\};", txt, re.S).group(1))]
assert len(tri) == 256 * 16 and len(edge) == 256
rows = [tri[16 * i:16 * i + 16] for i in range(256)]
ntri = [(r.index(-1) if -1 in r else 16) // 3 for r in rows]
out = dict(
    tri_sha256=hashlib.sha256(",".join(map(str, tri)).encode()).hexdigest(),
    edge_sha256=hashlib.sha256(",".join(map(str, edge)).encode()).hexdigest(),
    triangles_per_case=ntri,
    note="sha256 of the comma-joined decimal values in row-major order")
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mc_tables.json")
json.dump(out, open(path, "w"), indent=1)
print("wrote", path)
