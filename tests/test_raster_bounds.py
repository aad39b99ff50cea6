"""CPU: the raster's pixel-range narrowing (tri_tight in rvs_device.h) never
drops a pixel that the coverage test (transform.cpp:105-117, float32 as the
kernel evaluates it) accepts.

Random triangles of every shape -- tiny, stretched, near-degenerate slivers,
far off-screen vertices -- are tested exhaustively over their reference
bounding box (max(0, floor(min)) .. ceil(max), as colorize_triangle clips it);
every accepted pixel must lie in the narrowed range whenever the narrowing
applies.
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


def _tight(A, B, C, den):
    xs = np.stack([A[0], B[0], C[0]])
    ys = np.stack([A[1], B[1], C[1]])
    minx, maxx, miny, maxy = xs.min(0), xs.max(0), ys.min(0), ys.max(0)
    sz = np.maximum(maxx - minx, maxy - miny) + f(2)
    guard = den >= f(6.103515625e-05) * sz * sz * sz
    xa = np.ceil(np.maximum(minx - f(1), f(-2)))
    xb = np.floor(np.minimum(maxx, f(1e9)))
    ya = np.ceil(np.maximum(miny - f(1), f(-2)))
    yb = np.floor(np.minimum(maxy, f(1e9)))
    return guard, xa, xb, ya, yb, minx, maxx, miny, maxy


def _triangles(rng, n):
    kind = rng.integers(0, 4, n)
    c = rng.uniform(0, 64, (n, 2)).astype(f)
    scale = np.where(kind == 0, 0.3, np.where(kind == 1, 3.0, np.where(kind == 2, 20.0, 8.0)))
    d1 = rng.normal(0, 1, (n, 2)) * scale[:, None]
    d2 = rng.normal(0, 1, (n, 2)) * scale[:, None]
    # kind 3: slivers -- second edge nearly parallel to the first
    d2 = np.where((kind == 3)[:, None], d1 * rng.uniform(0.2, 3.0, (n, 1)) + rng.normal(0, 1e-3, (n, 2)), d2)
    A = c.T
    B = (c + d1).astype(f).T
    C = (c + d2).astype(f).T
    return A, B, C


def test_narrowed_range_keeps_every_covered_pixel():
    rng = np.random.default_rng(2024)
    checked = narrowed = 0
    for rep in range(6):
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
        guard, xa, xb, ya, yb, minx, maxx, miny, maxy = _tight(A, B, C, den)
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
                assert not cov.any(), "covered pixel outside the narrowed range"
    assert checked > 1_000_000 and narrowed > 100_000


def test_slivers_keep_the_full_box():
    """A triangle of area ~1e-6 px^2 spanning 10 px fails the guard."""
    A = np.array([[0.0], [0.0]], f)
    B = np.array([[10.0], [1e-7]], f)
    C = np.array([[5.0], [-1e-7]], f)
    den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
    guard = _tight(A, B, C, den)[0]
    assert not guard.any()

