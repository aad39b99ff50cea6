"""CPU: the oracle's restatement of write_color's resize to the output size
(image_writing.cpp:189-190; cv::resize INTER_LINEAR), used when Precision != 1.

OpenCV is absent, so its own operation order is parity unpinned; the rule
itself is pinned here by closed forms that a plausible mistake would fail:
  * an exact 2x downscale is the mean of each 2x2 block (OpenCV's area path);
  * a 3x downscale samples the centre pixel of each 3x3 block (half-pixel
    centres: f = (d + 0.5) * 3 - 0.5 = 3d + 1 exactly; without the offsets it
    would take the block's corner);
  * a 2x upscale weights neighbours 3/4 : 1/4 and repeats the edge pixels
    (f = d / 2 - 1/4, clamped at both ends);
  * a constant image stays constant at any size.
Integer images keep every product and sum exact in float, so the expected
values are computed exactly in float64 and compared bit for bit.
"""
import numpy as np

import oracle_bind as ob


def _img(h, w, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1024, (h, w, 3)).astype(np.float32)


def test_exact_2x_downscale_is_the_block_mean():
    src = _img(12, 18, 1)
    got = ob.resize_color(src, 9, 6)
    s = src.astype(np.float64)
    want = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2]) / 4
    assert np.array_equal(got, want.astype(np.float32))


def test_3x_downscale_takes_the_centre_pixel():
    src = _img(15, 21, 2)
    got = ob.resize_color(src, 7, 5)
    assert np.array_equal(got, src[1::3, 1::3])


def _up2_1d(n):
    """Weights of a 2x upscale along one axis: destination d <- (s0, s1, f)."""
    out = []
    for d in range(2 * n):
        f = (d + 0.5) * 0.5 - 0.5
        s0 = int(np.floor(f))
        f -= s0
        if s0 < 0:
            s0, f = 0, 0.0
        if s0 >= n - 1:
            s0, f = n - 1, 0.0
        out.append((s0, min(s0 + 1, n - 1), f))
    return out


def test_2x_upscale_quarter_weights_and_edges():
    src = _img(5, 7, 3)
    got = ob.resize_color(src, 14, 10)
    s = src.astype(np.float64)
    wx, wy = _up2_1d(7), _up2_1d(5)
    # every interior weight is 1/4 or 3/4, the edges repeat the border pixel
    assert {f for _, _, f in wx[1:-1]} == {0.25, 0.75} and wx[0] == (0, 1, 0.0) and wx[-1][2] == 0.0
    want = np.zeros((10, 14, 3))
    for y, (y0, y1, fy) in enumerate(wy):
        for x, (x0, x1, fx) in enumerate(wx):
            r0 = s[y0, x0] * (1 - fx) + s[y0, x1] * fx
            r1 = s[y1, x0] * (1 - fx) + s[y1, x1] * fx
            want[y, x] = r0 * (1 - fy) + r1 * fy
    assert np.array_equal(got, want.astype(np.float32))


def test_constant_stays_constant():
    src = np.full((9, 11, 3), 0.375, np.float32)
    for dw, dh in ((22, 18), (5, 4), (11, 9), (7, 13)):
        assert np.array_equal(ob.resize_color(src, dw, dh), np.full((dh, dw, 3), 0.375, np.float32)), (dw, dh)
