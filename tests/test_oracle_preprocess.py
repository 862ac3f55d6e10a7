"""Pins the oracle's depth preprocessing (oracle/dm_oracle_pre.c) by closed
forms: a constant image is a fixed point; with a range sigma far above the
depth spread the bilateral filter is the normalised Gaussian blur
(numpy convolution over the valid pixels); a sharp step is preserved with a
small range sigma; pixels without depth stay 0 and do not leak; the pyramid
keeps a plane's samples (symmetric weights) and rejects the far side of a
discontinuity.  Reference: include/image/pyramid/pyramid.cu:200-282."""
import numpy as np

import oracle_lib as O


def test_constant_fixed_point():
    d = np.full((40, 50), 1.7, np.float32)
    for bits in (32, 64):
        assert np.abs(O.bilateral(d, 3, 0.03, 2.0, bits) - 1.7).max() < 1e-6
        assert np.abs(O.pyr_down(d, 0.03, bits) - 1.7).max() < 1e-6


def test_wide_range_sigma_is_gaussian_blur():
    rng = np.random.default_rng(0)
    d = (1.0 + 0.01 * rng.standard_normal((30, 36))).astype(np.float32)
    d[5, 7] = 0.0
    r, ss = 3, 1.5
    out = O.bilateral(d, r, 1e4, ss)
    i, j = np.mgrid[-r:r + 1, -r:r + 1]
    ws = np.exp(-(i * i + j * j) / (2 * ss * ss))
    pad = np.pad(d.astype(np.float64), r)
    valid = (pad > 0).astype(np.float64)
    num = np.zeros(d.shape)
    den = np.zeros(d.shape)
    for a in range(2 * r + 1):
        for b in range(2 * r + 1):
            num += ws[a, b] * (pad * valid)[a:a + d.shape[0], b:b + d.shape[1]]
            den += ws[a, b] * valid[a:a + d.shape[0], b:b + d.shape[1]]
    ref = np.where(d > 0, num / np.maximum(den, 1e-30), 0.0)
    assert np.abs(out - ref).max() < 1e-6
    assert out[5, 7] == 0.0


def test_step_preserved_with_small_range_sigma():
    d = np.full((20, 40), 1.0, np.float32)
    d[:, 20:] = 2.0
    out = O.bilateral(d, 4, 0.01, 3.0)
    assert np.abs(out[:, :20] - 1.0).max() < 1e-6 and np.abs(out[:, 20:] - 2.0).max() < 1e-6


def test_pyramid_plane_and_discontinuity():
    v, u = np.mgrid[0:32, 0:48]
    plane = (1.0 + 0.001 * u + 0.002 * v).astype(np.float32)
    out = O.pyr_down(plane, 0.03)
    # interior: symmetric weights reproduce the plane at (2u, 2v)
    assert np.abs(out[1:-1, 1:-1] - plane[2:-2:2, 2:-2:2][:out.shape[0] - 2, :out.shape[1] - 2]).max() < 1e-6
    step = np.full((16, 16), 1.0, np.float32)
    step[:, 8:] = 3.0
    o2 = O.pyr_down(step, 0.03)
    assert np.abs(o2[:, :4] - 1.0).max() < 1e-6 and np.abs(o2[:, 4:] - 3.0).max() < 1e-6
