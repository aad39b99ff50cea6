"""CPU: the oracle's restatement of multispectral blending (BlendingMethod
"Multispectral": BlendedView.cpp:62-88, blending.cpp:62-104,371-394,
IntegralImage2D.h:738-770), pinned by brute force and by reduction:

  * the masked box blur equals a brute-force window sum over the finite
    values under the mask, truncated to an integer, divided by their count,
    with 0 outside the mask and windows clipped at the image border (values
    are multiples of 1/8, so every double sum is exact and the brute force
    needs no rounding model); a half-pixel shift of the window or a count that
    includes masked-out pixels would fail;
  * ColorSpace RGB blurs the three channels over one count (a pixel counts
    when all three channels are finite) and rounds the window sum to float
    instead of truncating it: brute force again, with a NaN in one channel
    removing the whole pixel and a sum of 2.5 kept as 2.5;
  * the split keeps low + high = input (Y) and leaves U, V in the low band;
  * a whole frame whose largest side is below 20 pixels blurs with radius 0
    (trunc(Y) for Y < 1 is 0), so with equal exponents the two bands add up
    exactly to simple blending: bit-identical to orc_synthesize.
"""
import numpy as np

import oracle_bind as ob
import workloads as W


def _brute_blur(y, mask, rad):
    h, w = y.shape
    out = np.zeros((h, w), np.float32)
    for i in range(h):
        for j in range(w):
            if mask[i, j] == 0:
                continue
            ys, ye = max(0, i - rad), min(h, i + rad + 1)
            xs, xe = max(0, j - rad), min(w, j + rad + 1)
            win = y[ys:ye, xs:xe].astype(np.float64)
            m = (mask[ys:ye, xs:xe] == 255) & np.isfinite(win)
            cnt = int(m.sum())
            s = int(win[m].sum())            # (int) of the window sum
            out[i, j] = np.float32(np.float32(s) / np.float32(cnt)) if cnt else 0.0
    return out


def test_blur_matches_brute_force():
    rng = np.random.default_rng(31)
    for h, w, rad in ((9, 13, 2), (16, 11, 3), (7, 7, 0), (12, 20, 5)):
        y = (rng.integers(0, 40, (h, w)) / 8.0).astype(np.float32)     # multiples of 1/8: exact sums
        y[rng.random((h, w)) < 0.05] = np.nan
        mask = np.where(rng.random((h, w)) < 0.8, 255, 0).astype(np.uint8)
        got = ob.blur_masked(y, mask, rad)
        want = _brute_blur(y, mask, rad)
        assert np.array_equal(got, want), (h, w, rad)
        assert np.all(got[mask == 0] == 0)


def test_blur_truncates_the_window_sum():
    # four 0.75 under a 2x2 window (rad 1 at a corner): sum 3.0 -> 3 / 4; with 0.625: 2.5 -> 2 / 4
    y = np.full((2, 2), 0.625, np.float32)
    mask = np.full((2, 2), 255, np.uint8)
    assert np.all(ob.blur_masked(y, mask, 1) == np.float32(2.0 / 4.0))


def _brute_blur_rgb(col, mask, rad):
    h, w = mask.shape
    out = np.zeros((h, w, 3), np.float32)
    for i in range(h):
        for j in range(w):
            if mask[i, j] == 0:
                continue
            ys, ye = max(0, i - rad), min(h, i + rad + 1)
            xs, xe = max(0, j - rad), min(w, j + rad + 1)
            win = col[ys:ye, xs:xe].astype(np.float64)
            m = (mask[ys:ye, xs:xe] == 255) & np.all(np.isfinite(win), axis=2)
            cnt = int(m.sum())
            if cnt:
                out[i, j] = (win[m].sum(axis=0).astype(np.float32) / np.float32(cnt)).astype(np.float32)
    return out


def test_blur_rgb_matches_brute_force():
    rng = np.random.default_rng(8)
    for h, w, rad in ((9, 13, 2), (11, 8, 4), (6, 6, 0)):
        col = (rng.integers(0, 40, (h, w, 3)) / 8.0).astype(np.float32)
        col[rng.random((h, w)) < 0.06, rng.integers(0, 3)] = np.inf
        mask = np.where(rng.random((h, w)) < 0.8, 255, 0).astype(np.uint8)
        got = ob.blur_masked_rgb(col, mask, rad)
        assert np.array_equal(got, _brute_blur_rgb(col, mask, rad)), (h, w, rad)
    # no integer truncation in the three-channel path: 4 x 0.625 = 2.5 -> 2.5 / 4
    col = np.full((2, 2, 3), 0.625, np.float32)
    assert np.all(ob.blur_masked_rgb(col, np.full((2, 2), 255, np.uint8), 1) == np.float32(0.625))


def test_split_rgb_blurs_every_channel():
    rng = np.random.default_rng(6)
    h, w = 25, 44
    col = rng.random((h, w, 3)).astype(np.float32)
    depth = rng.uniform(1, 5, (h, w)).astype(np.float32)
    shape = np.where(rng.random((h, w)) < 0.9, 1.0, 0.0).astype(np.float32)
    low, high = ob.split_frequencies(col, depth, shape, rgb=True)
    mask = np.where(shape / depth > 0, 255, 0).astype(np.uint8)
    assert np.array_equal(low, ob.blur_masked_rgb(col, mask, max(h, w) // 20))
    assert np.array_equal(high, col - low)


def test_split_low_plus_high_is_the_input():
    rng = np.random.default_rng(5)
    h, w = 30, 41
    col = rng.random((h, w, 3)).astype(np.float32)
    depth = rng.uniform(1, 5, (h, w)).astype(np.float32)
    shape = np.where(rng.random((h, w)) < 0.9, rng.uniform(1, 10000, (h, w)), 0).astype(np.float32)
    low, high = ob.split_frequencies(col, depth, shape)
    assert np.array_equal(low[..., 1:], col[..., 1:]) and np.all(high[..., 1:] == 0)
    assert np.array_equal(high[..., 0], col[..., 0] - low[..., 0])
    mask = np.where(shape / depth > 0, 255, 0).astype(np.uint8)
    assert np.array_equal(low[..., 0], ob.blur_masked(col[..., 0], mask, max(h, w) // 20))


def test_small_frame_with_equal_exponents_is_simple_blending():
    rg = W.rig("rects_640", scale=40)          # 16 x 12 sources and target: blur radius 0
    planes = W.render_sources(rg)
    t = rg["target"]
    src = W.oracle_sources(rg, planes, ob)
    cam = W.oracle_cam(t["cam"], ob)
    for a in (1.0, 5.0):
        simple = ob.synthesize(src, cam, [0.001, 0.0, 0.0], [1.0, 0.0, 0.0], t["bitdepth"], a)
        ms = ob.synthesize(src, cam, [0.001, 0.0, 0.0], [1.0, 0.0, 0.0], t["bitdepth"], 0.0, multispectral=(a, a))
        for k in ("y", "u", "v", "depth", "validity", "quality"):
            assert np.array_equal(simple[k].view(np.uint8), ms[k].view(np.uint8)), (a, k)
