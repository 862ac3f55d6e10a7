"""Pins the oracle's marching cubes (include/geometry/mesh/meshExtractor/
meshExtractor.cu:125-242): a wall is extracted at its depth; a sphere seen
from six sides gives a closed, consistently welded surface of Euler
characteristic 2 whose vertices lie on the sphere."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_lib as O

S = dict(rows=160, cols=160, fx=180.0, fy=180.0, cx=79.5, cy=79.5)
PROPS = dict(voxel_size=0.01, num_buckets=1 << 14, bucket_size=10, num_blocks=1 << 14, max_sdf_weight=64,
             truncation_distance=0.03, max_sensor_depth=5.0, min_sensor_depth=0.1)


def look_at(eye, target):
    f = np.asarray(target, float) - eye
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0]) if abs(f[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    x = np.cross(up, f)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    T = np.eye(4)
    T[:3, :3] = np.stack([x, y, f], 1)
    T[:3, 3] = eye
    return T


def sphere_depth(T, c, r):
    v, u = np.mgrid[0:S["rows"], 0:S["cols"]]
    dc = np.stack([(u - S["cx"]) / S["fx"], (v - S["cy"]) / S["fy"], np.ones_like(u, float)], -1)
    d = dc @ T[:3, :3].T
    oc = T[:3, 3] - c
    a = (d * d).sum(-1)
    b = d @ oc
    cc = oc @ oc - r * r
    disc = b * b - a * cc
    t = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / a, 0.0)
    return t.astype(np.float32)


def observed_corners(vol):
    """Dense flag grid: corner C (at (C - 0.5) * vs) has all 8 voxels C-1..C observed,
    i.e. the strict trilinear sample at it exists."""
    xyz, sdf, rgb, w = vol.blocks()
    g = (xyz[:, None, :] * 8 + np.stack(np.unravel_index(np.arange(512), (8, 8, 8)), 1)[None]).reshape(-1, 3)
    lo = g.min(0) - 2
    shape = tuple(g.max(0) - lo + 3)
    obs = np.zeros(shape, bool)
    gi = g - lo
    obs[gi[:, 0], gi[:, 1], gi[:, 2]] = w.reshape(-1) > 0
    ok = obs[1:, 1:, 1:] & obs[:-1, 1:, 1:] & obs[1:, :-1, 1:] & obs[1:, 1:, :-1]
    ok &= obs[:-1, :-1, 1:] & obs[:-1, 1:, :-1] & obs[1:, :-1, :-1] & obs[:-1, :-1, :-1]
    return ok, lo + 1  # ok[i] is corner C = i + lo + 1


def edge_corner(E):
    off = 1 << 19
    E = E.astype(np.uint64)
    ax = (E >> np.uint64(60)).astype(np.int64)
    C = np.stack([((E >> np.uint64(s)) & np.uint64((1 << 20) - 1)).astype(np.int64) - off for s in (40, 20, 0)], 1)
    return ax, C


@pytest.mark.parametrize("bits", [32, 64])
def test_sphere_closed_surface(bits):
    """Surface is a 2-manifold wherever every cell around an edge has its strict
    samples (the rule of meshExtractor.cu:33-41 leaves holes next to unobserved
    voxels); vertices lie on the sphere and on their edge."""
    c, r = np.array([0.013, -0.021, 0.007]), 0.22
    vol = O.OracleVolume(PROPS, bits)
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 1), (-1, -1, 1)]
    for d in dirs:
        eye = c + 0.9 * np.asarray(d, float) / np.linalg.norm(d)
        T = look_at(eye, c)
        for rep in range(2):
            vol.integrate_scan(sphere_depth(T, c, r), None, S, T)
    pos, col, eid = O.mc_extract(vol)
    assert len(pos) > 1000
    rad = np.linalg.norm(pos.reshape(-1, 3) - c, axis=1)
    assert np.abs(rad - r).max() < 6e-3 and np.median(np.abs(rad - r)) < 2e-3
    # each vertex lies on the grid edge its id names
    ax, C = edge_corner(eid.reshape(-1))
    p = pos.reshape(-1, 3).astype(np.float64)
    c0 = (C - 0.5) * PROPS["voxel_size"]
    along = p[np.arange(len(p)), ax] - c0[np.arange(len(p)), ax]
    assert np.all((along > -1e-6) & (along < PROPS["voxel_size"] + 1e-6))
    for k in range(3):
        m = ax != k
        assert np.abs(p[m, k] - c0[m, k]).max() < 1e-6
    tri, nv = O.weld(eid)
    assert np.all(tri[:, 0] != tri[:, 1]) and np.all(tri[:, 1] != tri[:, 2]) and np.all(tri[:, 0] != tri[:, 2])
    e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    ue, cnt = np.unique(e, axis=0, return_counts=True)
    assert cnt.max() == 2
    # every open mesh edge touches a cell that lost a strict sample
    v2e = np.zeros(nv, np.uint64)
    v2e[tri.reshape(-1)] = eid.reshape(-1)
    ok, base = observed_corners(vol)
    open_edges = ue[cnt == 1]
    assert len(open_edges) < 0.1 * len(ue)
    for vtx in open_edges.reshape(-1)[:4000]:
        a, Cv = edge_corner(v2e[vtx:vtx + 1])
        a, Cv = int(a[0]), Cv[0]
        o = [j for j in range(3) if j != a]
        Q = np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [0, 1], indexing="ij")).reshape(3, -1).T
        corners = np.repeat(Cv[None], len(Q), 0)
        corners[:, o[0]] += Q[:, 0]
        corners[:, o[1]] += Q[:, 1]
        corners[:, a] += Q[:, 2]
        idx = corners - base
        assert not ok[idx[:, 0], idx[:, 1], idx[:, 2]].all(), "hole next to fully observed cells"


def test_sphere_analytic_closed():
    """A TSDF written from the sphere's signed distance (every voxel observed)
    meshes to a closed genus-0 surface: Euler characteristic 2."""
    c, r = np.array([0.013, -0.021, 0.007]), 0.15
    vol = O.OracleVolume(PROPS, 64)
    lo = np.floor((c - r - 0.05) / 0.08).astype(int)
    hi = np.floor((c + r + 0.05) / 0.08).astype(int)
    blocks = np.array(np.meshgrid(*[np.arange(lo[k], hi[k] + 1) for k in range(3)], indexing="ij")).reshape(3, -1).T
    g = (blocks[:, None, :] * 8 + np.stack(np.unravel_index(np.arange(512), (8, 8, 8)), 1)[None])
    d = np.linalg.norm(g * 0.01 - c, axis=-1) - r
    vol.set_blocks(blocks.astype(np.int32), np.clip(d, -0.03, 0.03), np.full(d.shape, 7, np.uint8))
    pos, col, eid = O.mc_extract(vol)
    rad = np.linalg.norm(pos.reshape(-1, 3) - c, axis=1)
    assert np.abs(rad - r).max() < 2e-3
    tri, nv = O.weld(eid)
    e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    ue, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2), np.bincount(cnt)
    assert nv - len(ue) + len(tri) == 2


def test_wall_vertices_at_depth():
    S2 = dict(rows=120, cols=160, fx=150.0, fy=150.0, cx=80.0, cy=60.0)
    vol = O.OracleVolume(PROPS, 64)
    for k in range(3):
        vol.integrate_scan(np.full((120, 160), 1.0, np.float32), None, S2, np.eye(4))
    pos, col, eid = O.mc_extract(vol, box=([-0.2, -0.2, 0.5], [0.2, 0.2, 1.5]))
    assert len(pos) > 100
    assert np.abs(pos[..., 2] - 1.0).max() < 2e-3
    assert np.all(np.abs(pos[..., :2]) <= 0.2 + 0.011)
