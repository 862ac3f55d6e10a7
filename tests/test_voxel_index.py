"""The raycast's trilinear corner voxels, world_to_voxel(q - vs/2) and
world_to_voxel(q - vs/2 + vs) (GetInterpolatedVoxel, tsdf.cu:47-175, with
WorldToGlobalVoxelHashing, tsdf.cu:189-193: int(x / vs + sign(x) * 0.5)),
are taken as floor(vq) and floor(vq) + 1 from vq = q / vs whenever the
fraction of vq is more than |vq| * 1e-6 + 1e-6 away from 0 and 1, and
evaluated exactly otherwise.  Checked here on 2e7 float32 positions per
voxel size (uniform, near the voxel and half-voxel boundaries where the
rounding decides, and large magnitudes) against the exact float32
evaluation."""
import numpy as np

F = np.float32


def exact(x, vs):
    t = (x / vs).astype(F) + (np.sign(x).astype(F) * F(0.5))
    return np.trunc(t.astype(F)).astype(np.int64)


def test_corner_voxels_from_the_quotient():
    rng = np.random.default_rng(7)
    for vs in (F(0.004), F(0.008), F(0.02), F(0.05), F(0.0123)):
        n = 4_000_000
        k = rng.integers(-20000, 20000, n).astype(F)
        q = np.concatenate([
            rng.uniform(-60, 60, n).astype(F),
            (k * vs).astype(F),                                  # on voxel centres
            ((k + F(0.5)) * vs).astype(F),                       # half-way
            np.nextafter((k * vs).astype(F), F(np.inf)),
            np.nextafter((k * vs).astype(F), F(-np.inf)),
            rng.uniform(-2e4, 2e4, n).astype(F),
            np.array([0.0, -0.0, vs / 2, -vs / 2], F),
        ])
        half = F(vs / F(2))
        dual = (q - half).astype(F)
        i0 = exact(dual, vs)
        i1 = exact((dual + vs).astype(F), vs)
        vq = (q / vs).astype(F)
        fl = np.floor(vq)
        w = (vq - fl).astype(F)
        tol = (np.abs(vq) * F(1e-6) + F(1e-6)).astype(F)
        fast = (w > tol) & (w < (F(1) - tol)) & (np.abs(vq) < F(1e8))
        assert fast[:n].mean() > 0.95  # most uniform positions take the fast form
        assert np.array_equal(fl[fast].astype(np.int64), i0[fast]), vs
        assert np.array_equal(fl[fast].astype(np.int64) + 1, i1[fast]), vs
