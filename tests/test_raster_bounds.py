"""CPU: the raster's pixel-range narrowing (tri_tight in rvs_device.h) never
drops a pixel that the coverage test (transform.cpp:105-117, float32 as the
kernel evaluates it) accepts.

The narrowed range keeps the pixels whose centre lies within
dl = 0.07 + 4e-6 * s of the triangle's extent (s = max(W, H) + 2), when
den >= 2^-14 s^3.  Why that suffices: a centre more than dl outside the
x-extent W has an exact barycentric coordinate below -dl / (2 W) (the point's
x offset is the barycentric mean of the corners' offsets, at most two of which
are negative and none below -W), and the kernel's coordinates differ from the
exact ones by at most 2^-5 / s under the guard, so the computed one stays below
-dl / (2 W) + 2^-5 / s < -1e-6; likewise in y.  The binning passes narrow each
triangle's box the same way before they pick its tiles.

Random triangles of every shape -- tiny, stretched, near-degenerate slivers,
far off-screen vertices, corners on pixel centres, coordinates up to 8192 --
are tested exhaustively over their reference bounding box (max(0, floor(min))
.. ceil(max), as colorize_triangle clips it); every accepted pixel must lie in
the narrowed range whenever the narrowing applies, and a margin of zero is
caught.
"""
import numpy as np

f = np.float32


def _cover(A, B, C, den, x, y):
    px = x.astype(f) + f(0.5)
    py = y.astype(f) + f(0.5)
    n1 = (B[1] - C[1]) * (px - C[0]) + (C[0] - B[0]) * (py - C[1])
    n2 = (C[1] - A[1]) * (px - C[0]) + (A[0] - C[0]) * (py - C[1])
    l1 = n1 / den
    l2 = n2 / den
    l3 = f(1) - l1 - l2
    e = f(1e-6)
    return (l1 >= -e) & (l2 >= -e) & (l3 >= -e)


def _tight(A, B, C, den, zero_margin=False):
    """tri_tight's range, in the kernel's float operations (zero_margin: the
    same with dl = 0, a plausible mistake the test must catch)."""
    xs = np.stack([A[0], B[0], C[0]])
    ys = np.stack([A[1], B[1], C[1]])
    minx, maxx, miny, maxy = xs.min(0), xs.max(0), ys.min(0), ys.max(0)
    sz = np.maximum(maxx - minx, maxy - miny) + f(2)
    guard = den >= f(6.103515625e-05) * sz * sz * sz
    dl = f(0.07) + f(4e-6) * sz
    if zero_margin:
        dl = np.zeros_like(dl)
    xa = np.ceil(np.maximum(minx - f(0.5) - dl, f(-2)))
    xb = np.floor(np.minimum(maxx - f(0.5) + dl, f(1e9)))
    ya = np.ceil(np.maximum(miny - f(0.5) - dl, f(-2)))
    yb = np.floor(np.minimum(maxy - f(0.5) + dl, f(1e9)))
    return guard, xa, xb, ya, yb, minx, maxx, miny, maxy


def _triangles(rng, n):
    kind = rng.integers(0, 6, n)
    c = rng.uniform(0, 64, (n, 2)).astype(f)
    # kind 5: far from the origin (float spacing up to 2^-10 px)
    c = np.where((kind == 5)[:, None], c + rng.uniform(0, 8192, (n, 1)), c).astype(f)
    scale = np.where(kind == 0, 0.3, np.where(kind == 1, 3.0, np.where(kind == 2, 20.0, 8.0)))
    scale = np.where(kind >= 4, 1.0, scale)
    d1 = rng.normal(0, 1, (n, 2)) * scale[:, None]
    d2 = rng.normal(0, 1, (n, 2)) * scale[:, None]
    # kind 3: slivers -- second edge nearly parallel to the first
    d2 = np.where((kind == 3)[:, None], d1 * rng.uniform(0.2, 3.0, (n, 1)) + rng.normal(0, 1e-3, (n, 2)), d2)
    # kind 4: corners on pixel centres or a hair (4e-6 px) past them, small
    # integer offsets (edges through centres, the identity-warp ties; a
    # centre just outside the extent is still covered within the 1e-6 test)
    on = (kind == 4)[:, None]
    hair = np.where(rng.integers(0, 2, (n, 1)) == 1, f(4e-6), f(0))
    c = np.where(on, np.floor(c % 16) + f(0.5) + hair, c).astype(f)
    d1 = np.where(on, rng.integers(-8, 9, (n, 2)), d1)
    d2 = np.where(on, rng.integers(-8, 9, (n, 2)), d2)
    A = c.T
    B = (c + d1).astype(f).T
    C = (c + d2).astype(f).T
    return A, B, C


def _violations(rng, reps, zero_margin=False):
    """(pixels checked, triangles narrowed, covered pixels outside the range)."""
    checked = narrowed = bad = 0
    for rep in range(reps):
        A, B, C = _triangles(rng, 40000)
        den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
        # keep counter-clockwise (accepted) triangles; flip the others
        flip = den <= 0
        B2 = np.where(flip, C, B)
        C2 = np.where(flip, B, C)
        B, C = B2, C2
        den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
        keep = den > 0
        A, B, C, den = A[:, keep], B[:, keep], C[:, keep], den[keep]
        guard, xa, xb, ya, yb, minx, maxx, miny, maxy = _tight(A, B, C, den, zero_margin)
        # the kernel's box starts at column / row 0 (colorize_triangle clips it there)
        x0, x1 = np.maximum(np.floor(minx), 0), np.ceil(maxx)
        y0, y1 = np.maximum(np.floor(miny), 0), np.ceil(maxy)
        small = guard & (x1 - x0 <= 48) & (y1 - y0 <= 48) & (x1 >= x0) & (y1 >= y0)
        idx = np.nonzero(small)[0]
        narrowed += int(guard.sum())
        for dy in range(49):
            for dx in range(49):
                s = idx[(dx <= x1[idx] - x0[idx]) & (dy <= y1[idx] - y0[idx])]
                if len(s) == 0:
                    continue
                x = (x0[s] + dx).astype(np.int64)
                y = (y0[s] + dy).astype(np.int64)
                out = (x < xa[s]) | (x > xb[s]) | (y < ya[s]) | (y > yb[s])
                s, x, y = s[out], x[out], y[out]
                if len(s) == 0:
                    continue
                cov = _cover(A[:, s], B[:, s], C[:, s], den[s], x, y)
                checked += len(s)
                bad += int(cov.sum())
    return checked, narrowed, bad


def test_narrowed_range_keeps_every_covered_pixel():
    checked, narrowed, bad = _violations(np.random.default_rng(2024), 6)
    assert bad == 0, "%d covered pixels outside the narrowed range" % bad
    assert checked > 1_000_000 and narrowed > 100_000


def test_zero_margin_is_caught():
    """Without the margin, corners on pixel centres put covered centres on the
    extent's edge, half a pixel outside a range that stops at the extent: the
    harness above must see them."""
    _, _, bad = _violations(np.random.default_rng(7), 1, zero_margin=True)
    assert bad > 0


def test_slivers_keep_the_full_box():
    """A triangle of area ~1e-6 px^2 spanning 10 px fails the guard."""
    A = np.array([[0.0], [0.0]], f)
    B = np.array([[10.0], [1e-7]], f)
    C = np.array([[5.0], [-1e-7]], f)
    den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
    guard = _tight(A, B, C, den)[0]
    assert not guard.any()

