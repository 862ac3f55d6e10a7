"""The integration's colour merge replaces the IEEE float division of
Voxel::mergeVoxel (include/geometry/geom_Types.h:85-90),
uchar((c*w + c_new) / (w + 1)), by trunc((n + 1/2) * rcp(m)) with
n = c*w + c_new, m = w + 1 and a reciprocal good to 1 ulp.  Checked
exhaustively over every channel value, new value and weight, with the
reciprocal rounded to nearest and moved one ulp either way, against integer
floor division and against the float quotient of the reference form."""
import numpy as np


def test_merge_channel_exhaustive():
    c = np.arange(256, dtype=np.int64)[:, None]
    cn = np.arange(256, dtype=np.int64)[None, :]
    for w in range(256):
        n = (c * w + cn).ravel()
        m = w + 1
        want = n // m
        # reference form: correctly rounded float32 quotient, truncated
        q = (n.astype(np.float32) / np.float32(m)).astype(np.float32)
        assert np.array_equal(q.astype(np.int64), want)
        r0 = np.float32(1.0) / np.float32(m)
        for r in (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(1))):
            got = ((n.astype(np.float32) + np.float32(0.5)) * np.float32(r)).astype(np.float32)
            assert np.array_equal(got.astype(np.int64), want), (w, r)
