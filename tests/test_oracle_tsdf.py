"""Pins the oracle's integration and raycast with closed forms.

Integration (include/geometry/tsdf/tsdf.cu:267-344, merge
include/geometry/geom_Types.h:79-102): for a fronto-parallel wall at depth D
seen from the identity pose, a voxel on the optical axis at depth z has
sdf = D - z inside the band, +trunc in front of it, nothing behind; repeated
identical frames keep the sdf and raise the weight to min(n, max_weight);
two frames average.  Raycast (tsdf.cu:470-526): the wall and a sphere are
recovered at their analytic depth; wall normals face the camera.
"""
import numpy as np
import pytest

import oracle_lib as O

S = dict(rows=120, cols=160, fx=150.0, fy=150.0, cx=80.0, cy=60.0)
PROPS = dict(voxel_size=0.01, num_buckets=1 << 14, bucket_size=10, num_blocks=1 << 14, max_sdf_weight=64,
             truncation_distance=0.03, max_sensor_depth=5.0, min_sensor_depth=0.1)


def wall(D):
    return np.full((S["rows"], S["cols"]), D, np.float32)


def block_of(xyz, blocks):
    idx = {tuple(b): k for k, b in enumerate(blocks)}
    return idx


def axis_voxels(vol):
    """(z index, sdf, weight, rgb) of voxels on the optical axis (x = y = 0)."""
    xyz, sdf, rgb, w = vol.blocks()
    out = []
    for k, (bx, by, bz) in enumerate(xyz):
        if bx != 0 or by != 0:
            continue
        for lz in range(8):
            vi = 0 * 64 + 0 * 8 + lz  # local (0,0,lz): x-major index x*64 + y*8 + z
            out.append((bz * 8 + lz, sdf[k, vi], w[k, vi], rgb[k, vi]))
    return sorted(out)


@pytest.mark.parametrize("bits", [32, 64])
def test_wall_closed_form(bits):
    D, tau, vs = 1.0, 0.03, 0.01
    vol = O.OracleVolume(PROPS, bits)
    rgb = np.zeros((S["rows"], S["cols"], 3), np.uint8)
    rgb[..., 0], rgb[..., 1], rgb[..., 2] = 10, 200, 77
    vol.integrate_scan(wall(D), rgb, S, np.eye(4))
    seen = 0
    for gz, sdf, w, c in axis_voxels(vol):
        z = gz * vs
        # K (tsdf.cu:292-298): a block is skipped when the projection of its
        # corner + half a block edge (in camera coordinates) leaves the image
        zc = (gz // 8) * 8 * vs + 0.04
        if 150.0 * 0.04 / zc + 80.0 >= 159.5:
            assert w == 0
            continue
        if min(abs(z - (D - tau)), abs(z - (D + tau))) < 1e-6:
            continue  # on a band edge: the strict comparisons decide by rounding
        if z < D - tau:
            assert w == 1 and sdf == pytest.approx(tau, abs=1e-7)
        elif z < D + tau:
            assert w == 1 and sdf == pytest.approx(D - z, abs=2e-6 if bits == 32 else 1e-12)
            assert tuple(c) == (10, 200, 77)
            seen += 1
        else:
            assert w == 0
    assert seen >= 5


def test_weight_saturation_and_average():
    vol = O.OracleVolume(PROPS, 64)
    for n in range(70):
        vol.integrate_scan(wall(1.0), None, S, np.eye(4))
    vals = {gz: (sdf, w) for gz, sdf, w, c in axis_voxels(vol)}
    assert vals[98][1] == 64 and vals[98][0] == pytest.approx(0.02, abs=1e-12)
    vol2 = O.OracleVolume(PROPS, 64)
    vol2.integrate_scan(wall(1.0), None, S, np.eye(4))
    vol2.integrate_scan(wall(1.01), None, S, np.eye(4))
    vals = {gz: (sdf, w) for gz, sdf, w, c in axis_voxels(vol2)}
    # z = 0.99: ((1.0 - z) + (1.01 - z)) / 2 with the depths as float32
    want = ((1.0 - 0.99) + (float(np.float32(1.01)) - 0.99)) / 2
    assert vals[99][1] == 2 and vals[99][0] == pytest.approx(want, abs=1e-12)


def sphere_depth(c, r, D):
    v, u = np.mgrid[0:S["rows"], 0:S["cols"]]
    d = np.stack([(u - S["cx"]) / S["fx"], (v - S["cy"]) / S["fy"], np.ones_like(u, float)], -1)
    a = (d * d).sum(-1)
    b = (d * -np.asarray(c)).sum(-1)
    cc = np.dot(c, c) - r * r
    disc = b * b - a * cc
    t = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / a, np.inf)
    return np.minimum(t, D), disc


@pytest.mark.parametrize("bits", [32, 64])
def test_raycast_wall_and_sphere(bits):
    c, r, D = np.array([0.05, -0.02, 0.9]), 0.25, 1.6
    depth, disc = sphere_depth(c, r, D)
    vol = O.OracleVolume(PROPS, bits)
    for k in range(3):
        vol.integrate_scan(depth.astype(np.float32), None, S, np.eye(4))
    d, vert, nrm, _ = vol.raycast(np.eye(4), S, want_rgb=False)
    hit = d > 0
    assert hit.mean() > 0.95
    # interior pixels away from the silhouette
    on_sphere = disc > 0.02 * np.abs(disc).max()
    on_wall = (disc < -0.05 * np.abs(disc).max())
    err_s = np.abs(d - depth)[on_sphere & hit]
    err_w = np.abs(d - depth)[on_wall & hit]
    assert np.median(err_s) < 1.5e-3 and np.percentile(err_s, 95) < 4e-3
    assert np.median(err_w) < 1.0e-3
    n_w = nrm[on_wall & hit & np.isfinite(nrm).all(-1)]
    assert np.median(n_w[:, 2]) < -0.99  # wall normals face the camera
    # vertex = K^-1 (u, v, 1) * depth
    v_, u_ = np.mgrid[0:S["rows"], 0:S["cols"]]
    np.testing.assert_allclose(vert[..., 2][hit], d[hit], rtol=0, atol=0)
    np.testing.assert_allclose(vert[..., 0][hit], ((u_ - S["cx"]) * d / S["fx"])[hit], atol=1e-6)


def linear_volume(bits, a, hole=None):
    """Blocks -2..1 on each axis, every voxel observed, voxel g (global index
    triple) holding sdf = a[:3].(g vs) + a[3] and colour 100 + g, so that both
    are linear functions of the voxel centre g vs (tsdf.cu:189-239 puts voxel
    g's centre at g vs: WorldToGlobalVoxelHashing rounds half away from zero).
    `hole`: a global voxel index left unobserved (weight 0, sdf 5)."""
    vs = PROPS["voxel_size"]
    vol = O.OracleVolume(PROPS, bits)
    rng = range(-2, 2)
    bl = np.array([(x, y, z) for x in rng for y in rng for z in rng], np.int32)
    loc = np.arange(512)
    lxyz = np.stack([loc // 64, (loc // 8) % 8, loc % 8], -1)  # x-major, z fastest
    g = bl[:, None, :] * 8 + lxyz[None]
    sdf = (g * vs) @ np.asarray(a[:3]) + a[3]
    w = np.ones(g.shape[:2], np.uint8)
    if hole is not None:
        m = (g == np.asarray(hole)).all(-1)
        w[m] = 0
        sdf[m] = 5.0
    assert vol.set_blocks(bl, sdf, w, (100 + g).astype(np.uint8)) == 0
    return vol


@pytest.mark.parametrize("bits", [32, 64])
def test_trilinear_exact_on_linear_field(bits):
    """GetInterpolatedVoxel (tsdf.cu:47-175) reproduces a linear field
    exactly wherever no coordinate is a whole number of voxels: its corners
    are the voxels at floor(p / vs) + {0, 1} per axis and its weights
    frac(p / vs), so sum(w * f(corner)) = f(p).  A swapped weight (1 - f for
    f), corners offset by half a voxel, or axes exchanged each miss by up to
    a[k] * vs.

    K (reference quirk): the corners come from rounding p - vs/2 half away
    from zero (WorldToGlobalVoxelHashing, tsdf.cu:189-193) while the weights
    use floor(p / vs).  On an axis with p / vs = k <= 0 exactly, the rounding
    gives corner k - 1 and the weight puts all of that axis on it, so the
    sample is f(p) - a[axis] * vs per such axis; k = 0 is exact in both
    builds."""
    vs = PROPS["voxel_size"]
    a = np.array([0.9, -0.45, 0.3, 0.02])
    vol = linear_volume(bits, a)
    pts = np.random.default_rng(11).uniform(-0.14, 0.14, (4000, 3))
    sdf, col, ok = vol.interp(pts)
    assert ok.all()
    pf = pts.astype(np.float32).astype(np.float64) if bits == 32 else pts
    want = pf @ a[:3] + a[3]
    tol = 1e-12 if bits == 64 else 2e-6
    assert np.abs(sdf - want).max() < tol
    want_c = 100 + pf / vs
    assert np.abs(col - want_c).max() < (1e-9 if bits == 64 else 2e-4)
    # coordinates exactly 0 on one, two or three axes
    q = np.random.default_rng(13).uniform(-0.13, 0.13, (600, 3))
    zero = np.zeros((600, 3), bool)
    zero[:200, 0] = True
    zero[200:400, 1:] = True
    zero[400:] = True
    q[zero] = 0.0
    sdf, col, ok = vol.interp(q)
    assert ok.all()
    qf = q.astype(np.float32).astype(np.float64) if bits == 32 else q
    want = qf @ a[:3] + a[3] - (zero * a[:3] * vs).sum(-1)
    assert np.abs(sdf - want).max() < tol
    assert np.abs(col - (100 + qf / vs - zero)).max() < (1e-9 if bits == 64 else 2e-4)


@pytest.mark.parametrize("bits", [32, 64])
def test_trilinear_unobserved_corner_takes_centre_voxel(bits):
    """tsdf.cu:47-175: a corner voxel with weight 0 contributes the values of
    the centre voxel (the voxel containing p) instead of its own; an
    unobserved centre voxel makes the sample unobserved.  On a linear field
    the result is therefore f(p) + c_h (f(centre) - f(hole)) with c_h the
    trilinear weight of the hole's corner."""
    vs = PROPS["voxel_size"]
    a = np.array([0.9, -0.45, 0.3, 0.02])
    hole = (3, -2, 1)
    vol = linear_volume(bits, a, hole=hole)
    rng = np.random.default_rng(12)
    # points whose 8 corners include the hole but whose centre voxel is another one
    pts = (np.asarray(hole) + rng.uniform(-0.98, 0.98, (2000, 3))) * vs
    centre = np.floor(pts / vs + 0.5).astype(int)
    corner0 = np.floor(pts / vs).astype(int)
    sdf, col, ok = vol.interp(pts)
    at_hole = (centre == hole).all(-1)
    assert not ok[at_hole].any() and ok[~at_hole].all()
    f = lambda g: (g * vs) @ a[:3] + a[3]
    wt = pts / vs - corner0
    d = np.asarray(hole) - corner0
    in_cell = ((d == 0) | (d == 1)).all(-1)
    c_h = np.prod(np.where(d == 1, wt, 1 - wt), -1) * in_cell
    want = pts @ a[:3] + a[3] + c_h * (f(centre) - f(np.asarray(hole)))
    sel = ~at_hole
    tol = 1e-12 if bits == 64 else 5e-6
    assert np.abs(sdf[sel] - want[sel]).max() < tol
    assert (c_h[sel] > 0.05).sum() > 200  # the rule was exercised
