"""Writes tests/golden/hash_vectors.json: inputs and expected outputs of the
voxel-block hash of include/dataStructure/hashTable/hashTable.cu:152-163,
computed here with Python integers (int32 wrap of each product and of the
XOR made explicit, C remainder truncating toward zero), independently of the
C oracle.  Run: python tests/golden/make_hash_vectors.py
"""
import json
import os


def wrap32(v):
    return ((v + (1 << 31)) % (1 << 32)) - (1 << 31)


def c_mod(a, b):
    r = abs(a) % b
    return -r if a < 0 else r


def ref_hash(x, y, z, buckets, bucket_size):
    h = wrap32(wrap32(x * 73856093) ^ wrap32(y * 19349669) ^ wrap32(z * 83492791))
    r = c_mod(h, buckets)
    if r < 0:
        r += buckets
    return r * bucket_size


keys = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, -1),
        (1, 2, 3), (-7, 13, -21), (29, 30, 31), (100, -200, 300), (1000, 1000, 1000), (-1000, 500, -250),
        (65535, -65536, 12345), (1048575, -1048576, 7), (46341, 46341, 46341), (-46341, 29, 77),
        (524287, 524287, -524288), (3, -5, 8)]
vectors = []
for buckets, bs in ((2500, 10), (1 << 16, 10), (1 << 20, 10), (1 << 22, 10), (512, 1), (7, 3)):
    for k in keys:
        vectors.append(dict(key=list(k), num_buckets=buckets, bucket_size=bs, slot=ref_hash(*k, buckets, bs)))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_vectors.json")
with open(out, "w") as f:
    json.dump(vectors, f, indent=0)
print("wrote", out, len(vectors))
