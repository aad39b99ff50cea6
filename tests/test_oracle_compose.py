"""CPU pins of the oracle's whole-frame composition (orc_synthesize) and of the
raster's seam triangles and prolongation branch.

orc_synthesize restates Pipeline::computeView's CPU branch (Pipeline.cpp:169-301):
per source in InputCameraNames order, decode with the first loaded view's
function-static scales (image_loading.cpp:134,193), reproject, raster
(transform.cpp:148-181), blend (BlendedView.cpp:180-212); then inpaint with the
mask `quality <= 0` (View.cpp:198-203) and encode.  Each test states a
property the composition must have, derived from those lines, and computes
the expected value independently (closed forms, single-source runs, exact L1
distances), so that a mistake in one composition step fails one of them.

The raster tests extend the brute force of test_oracle.py to the seam
triangles of a horizontally wrapping target (transform.cpp:170-175) and to the
prolongation rule (transform.cpp:119-142): a triangle whose three vertices all
have depth <= 0 overwrites a pixel whose prolongation flag is still set,
whatever its score, and clears the flag, after which only such triangles can
replace it.
"""
import numpy as np
import pytest

import oracle_bind as ob

F = 16.0          # focal length (power of two: the identity warp is exact)
NEAR, FAR = 2.0, 8.0


def persp(w, h, px=None, py=None):
    return ob.make_cam(0, w, h, f=(F, F), p=(w / 2.0 if px is None else px, h / 2.0 if py is None else py))


def yuv_planes(rng, w, h, bitdepth, lo=0.05, hi=0.95):
    m = (1 << bitdepth) - 1
    dt = np.uint8 if bitdepth <= 8 else np.uint16
    y = rng.integers(int(lo * m), int(hi * m), (h, w)).astype(dt)
    u = rng.integers(int(lo * m), int(hi * m), (h // 2, w // 2)).astype(dt)
    v = rng.integers(int(lo * m), int(hi * m), (h // 2, w // 2)).astype(dt)
    return y, u, v


def source(cam, y, u, v, depth_raw, bd_color, bd_depth, pos=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.0),
           znear=NEAR, zfar=FAR):
    return dict(cam=cam, pos=list(pos), rot=list(rot), bd_color=bd_color, bd_depth=bd_depth, znear=znear,
                zfar=zfar, has_invalid=True, y=y, u=u, v=v, depth=depth_raw)


def flat_depth(w, h, bd):
    """Raw depth at the maximum level everywhere: decodes to NEAR exactly
    (v = max * fl(1/max) = 1, FAR*NEAR / (1*(FAR-NEAR) + NEAR) = NEAR)."""
    dt = np.uint8 if bd <= 8 else np.uint16
    return np.full((h, w), (1 << bd) - 1, dt)


def l1_nearest_sets(known):
    """For every pixel: the exact L1 distance to the nearest known pixel, and
    the set of known pixels at that distance (brute force)."""
    ky, kx = np.nonzero(known)
    h, w = known.shape
    gy, gx = np.mgrid[0:h, 0:w]
    d = np.abs(gy[..., None] - ky) + np.abs(gx[..., None] - kx)
    dmin = d.min(-1)
    return dmin, d == dmin[..., None], ky, kx


# ----------------------------------------------------------------- identity

def test_identity_warp_is_exact_with_near_plane_depth():
    """Same camera for source and target, depth decoding to NEAR (a power of
    two) and f = 16: every reprojected vertex lands on its pixel centre, so the
    synthesized colour is the decoded source colour bit for bit, and the
    encoded planes round trip the raw Y plane except where trunc(fl(y/max)*max)
    loses a code."""
    rng = np.random.default_rng(11)
    w, h, bd = 24, 16, 10
    y, u, v = yuv_planes(rng, w, h, bd)
    cam = persp(w, h)
    out = ob.synthesize([source(cam, y, u, v, flat_depth(w, h, 16), bd, 16)], cam, [0, 0, 0], [0, 0, 0], bd, 1.0)
    dec = ob.decode_color(y, u, v, bd)
    assert np.array_equal(out["color"], dec)
    assert (out["depth"] == np.float32(NEAR)).all()
    s = np.float32(1.0) / np.float32(1023)
    ytrunc = (y.astype(np.float32) * s * np.float32(1023)).astype(np.int64)
    assert np.array_equal(out["y"].astype(np.int64), ytrunc)
    assert (np.abs(out["y"].astype(np.int64) - y) <= 1).all()


# ---------------------------------------------- inpainting of uncovered area

@pytest.mark.parametrize("n_src", [1, 2])
def test_uncovered_target_area_is_filled_from_the_l1_nearest_covered_pixel(n_src):
    """The target is wider and taller than the source (same intrinsics at the
    same pose), so its right and bottom bands are uncovered.  Every uncovered
    pixel takes the colour of a covered pixel at the exact minimum L1 distance;
    covered pixels keep their colour.  Some covered pixels have exactly the
    empty colour (0, 1, 0) that the blender writes where nothing covers: the
    hole mask is quality <= 0 (View.cpp:198-203), so they stay known, and the
    holes next to them copy (0, 1, 0)."""
    rng = np.random.default_rng(12)
    bd = 10
    sw, sh, tw, th = 20, 14, 29, 20
    ox, oy = 3, 2              # the source covers target columns ox.., rows oy..
    y, u, v = yuv_planes(rng, sw, sh, bd)
    # a 4x4 block of the source's bottom right corner decodes to (0, 1, 0)
    y[sh - 4:, sw - 4:] = 0
    u[(sh - 4) // 2:, (sw - 4) // 2:] = 1023
    v[(sh - 4) // 2:, (sw - 4) // 2:] = 0
    src_cam = persp(sw, sh, px=sw / 2.0, py=sh / 2.0)
    tgt_cam = persp(tw, th, px=sw / 2.0 + ox, py=sh / 2.0 + oy)
    srcs = [source(src_cam, y, u, v, flat_depth(sw, sh, 16), bd, 16)] * n_src
    out = ob.synthesize(srcs, tgt_cam, [0, 0, 0], [0, 0, 0], bd, 1.0)
    known = out["quality"] > 0
    # the source mesh spans pixel centres 0.5 .. sw-0.5 shifted by (ox, oy):
    # holes on all four sides
    inner = (slice(oy, oy + sh), slice(ox, ox + sw))
    kk = known.copy()
    assert kk[inner].all()
    kk[inner] = False
    assert not kk.any()
    col = out["color"]
    dec = ob.decode_color(y, u, v, bd)
    if n_src == 1:
        assert np.array_equal(col[inner], dec)
    else:   # two equal views: (c*q + c*q) * fl(1/(2q)) within an ulp of c
        assert np.allclose(col[inner], dec, rtol=2e-7, atol=0)
    assert (col[oy + sh - 4:oy + sh, ox + sw - 4:ox + sw] == np.array([0, 1, 0], np.float32)).all()
    dmin, nearest, ky, kx = l1_nearest_sets(known)
    holes = np.argwhere(~known)
    assert len(holes) == tw * th - sw * sh
    for (yy, xx) in holes:
        cand = nearest[yy, xx]
        opts = col[ky[cand], kx[cand]]
        assert (opts == col[yy, xx]).all(-1).any(), (yy, xx)
    # the corner hole (th-1, tw-1) is nearest to the empty-coloured block only
    assert (col[th - 1, tw - 1] == np.array([0, 1, 0], np.float32)).all()
    assert dmin[th - 1, tw - 1] == (th - sh - oy) + (tw - sw - ox)
    assert dmin[0, 0] == ox + oy


# -------------------------------------------- the first loaded view's scales

def test_sources_are_scaled_by_the_first_loaded_views_bit_depths():
    """read_color_YUV / read_depth_YUV keep function-static scales initialised
    by the first view loaded (image_loading.cpp:134,193).  Source 0 is 8-bit
    colour / 8-bit depth, source 1 10-bit colour / 16-bit depth: source 1's
    samples are scaled by fl(1/255) for both, which is what the pixels only
    source 1 covers show."""
    rng = np.random.default_rng(13)
    w, h = 24, 16
    y0, u0, v0 = yuv_planes(rng, 12, h, 8)
    y1, u1, v1 = yuv_planes(rng, w, h, 10, lo=0.01, hi=0.2)
    raw1 = np.full((h, w), 3000, np.uint16)
    cam0 = persp(12, h, px=w / 2.0, py=h / 2.0)       # covers target columns 0..11
    cam1 = persp(w, h)
    srcs = [source(cam0, y0, u0, v0, flat_depth(12, h, 8), 8, 8),
            source(cam1, y1, u1, v1, raw1, 10, 16)]
    out = ob.synthesize(srcs, cam1, [0, 0, 0], [0, 0, 0], 10, 1.0)
    only1 = np.zeros((h, w), bool)
    only1[:, 13:] = True
    s8 = np.float32(1.0) / np.float32(255)
    expect_y = y1.astype(np.float32) * s8
    # colour = (0 + c*q) * fl(1/q) with q of source 1: within an ulp of c
    assert np.allclose(out["color"][only1][:, 0], expect_y[only1], rtol=2e-7, atol=0)
    wrong_y = y1.astype(np.float32) * (np.float32(1.0) / np.float32(1023))
    assert not np.allclose(out["color"][only1][:, 0], wrong_y[only1], rtol=1e-2)
    # depth: v = raw * fl(1/255) (not 1/65535), far*near / (v*(far-near) + near)
    vv = np.float64(3000 * s8)
    expect_d = FAR * NEAR / (vv * (FAR - NEAR) + NEAR)
    assert np.allclose(out["depth"][only1], expect_d, rtol=1e-6)
    vv16 = 3000.0 / 65535.0
    assert abs(expect_d - FAR * NEAR / (vv16 * (FAR - NEAR) + NEAR)) > 1.0
    # where source 0 alone decodes at its own scales: Y = y0 * fl(1/255)
    both = np.zeros((h, w), bool)
    both[:, :12] = True
    assert (out["quality"][both] > 0).all()


# ----------------------------------------------------------- blending order

def _two_source_scene(rng, w=28, h=20):
    """Two perspective sources at +-0.04 m lateral offset looking at a tilted,
    bumpy surface: their triangle stretch (hence quality) differs pixel by
    pixel in both directions."""
    bd = 10
    srcs = []
    gy, gx = np.mgrid[0:h, 0:w]
    for k, lat in enumerate((0.04, -0.04)):
        y, u, v = yuv_planes(rng, w, h, bd)
        z = 3.0 + 0.08 * gx + 0.6 * np.sin(gy * 0.9 + k) * np.cos(gx * 0.7)
        inv = (1.0 / z - 1.0 / FAR) / (1.0 / NEAR - 1.0 / FAR)
        raw = np.round(65535 * inv).astype(np.uint16)
        srcs.append(source(persp(w, h), y, u, v, raw, bd, 16, pos=(0.0, lat, 0.0)))
    return srcs


@pytest.mark.parametrize("exp", [1.0, 5.0])
def test_swapping_input_order_changes_only_the_depth_rule(exp):
    """BlendedViewSimple with two views: colour, quality and validity are
    symmetric in the views (0 + a + b == 0 + b + a in float; max commutes),
    and the depth is the second view's where its quality exceeds the first's
    (BlendedView.cpp:206-207).  The single-source runs give each view's
    quality and depth, so the depth of either order is a closed form in them."""
    rng = np.random.default_rng(14)
    A, B = _two_source_scene(rng)
    w, h = 28, 20
    tgt = persp(w, h)
    args = (tgt, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 10, exp)
    ab = ob.synthesize([A, B], *args)
    ba = ob.synthesize([B, A], *args)
    a1 = ob.synthesize([A], *args)
    b1 = ob.synthesize([B], *args)
    for k in ("color", "quality", "validity", "y", "u", "v"):
        assert np.array_equal(ab[k], ba[k]), k
    qa, qb, da, db = a1["quality"], b1["quality"], a1["depth"], b1["depth"]
    assert np.array_equal(ab["depth"], np.where(qb > qa, db, da))
    assert np.array_equal(ba["depth"], np.where(qa > qb, da, db))
    assert np.array_equal(ab["validity"], np.maximum(a1["validity"], b1["validity"]))
    # the test has teeth: both orders of quality occur where the depths differ
    diff = da != db
    assert ((qb > qa) & diff & (qa > 0)).sum() > 20 and ((qa > qb) & diff & (qb > 0)).sum() > 20
    # where only one view covers, the blend is that view's colour within an ulp
    # ((c*w) * fl(1/w)); with exponent a, quality = (w)^(1/a)
    only_a = (qa > 0) & (qb <= 0)
    if only_a.any():
        assert np.allclose(ab["color"][only_a], a1["color"][only_a], rtol=3e-7, atol=1e-9)


@pytest.mark.parametrize("exp", [1.0, 3.0])
def test_blend_depth_rule_on_exact_quality_ties(exp):
    """BlendedView.cpp:206-207 copies the new view's depth only where its
    quality is strictly greater than the accumulated one (quality > quality):
    on an exact tie (shape / depth equal, depths different) the first view's
    depth stays, so the two input orders keep different depths there, while
    colour, quality and validity are the same for both orders."""
    h, w = 4, 6
    rng = np.random.default_rng(18)
    cA = rng.random((h, w, 3)).astype(np.float32)
    cB = rng.random((h, w, 3)).astype(np.float32)
    dA = np.full((h, w), 2.0, np.float32)
    sA = np.full((h, w), 4000.0, np.float32)
    dB = np.full((h, w), 1.0, np.float32)
    sB = np.full((h, w), 2000.0, np.float32)          # q = 2000 for both
    sB[0] = 3000.0                                     # row 0: B better
    sB[1] = 1000.0                                     # row 1: A better
    res = {}
    for name, order in (("ab", ((cA, dA, sA), (cB, dB, sB))), ("ba", ((cB, dB, sB), (cA, dA, sA)))):
        b = ob.Blender(w, h, exp)
        for c, d, sh in order:
            b.blend(c, d, sh)
        res[name] = b
    tie = np.zeros((h, w), bool)
    tie[2:] = True
    assert (res["ab"].depth[tie] == 2.0).all() and (res["ba"].depth[tie] == 1.0).all()
    assert (res["ab"].depth[0] == 1.0).all() and (res["ba"].depth[0] == 1.0).all()
    assert (res["ab"].depth[1] == 2.0).all() and (res["ba"].depth[1] == 2.0).all()
    for k in ("color", "quality", "validity"):
        assert np.array_equal(getattr(res["ab"], k), getattr(res["ba"], k)), k
    # colour: the q^a-weighted mean of the two colours
    qa, qb = sA / dA, sB / dB
    wa, wb = qa.astype(np.float64) ** exp, qb.astype(np.float64) ** exp
    expect = (wa[..., None] * cA + wb[..., None] * cB) / (wa + wb)[..., None]
    assert np.allclose(res["ab"].color, expect, rtol=1e-6, atol=1e-7)
    assert np.allclose(res["ab"].quality, (wa + wb) ** (1.0 / exp), rtol=1e-6)


def test_three_views_blend_quality_is_the_power_sum():
    """Three views with exponent 2: the carried quality after the fold is
    fl-close to sqrt(q1^2 + q2^2 + q3^2) of the single-view qualities, and the
    colour is the q^2-weighted mean of the single-view colours, where all three
    cover (the running mean re-weights the accumulated colour by (sum w)^(1/a)
    raised back to a)."""
    rng = np.random.default_rng(15)
    A, B = _two_source_scene(rng)
    C = dict(A)
    C["pos"] = [0.0, 0.0, 0.02]
    y, u, v = yuv_planes(rng, 28, 20, 10)
    C.update(y=y, u=u, v=v)
    tgt = persp(28, 20)
    args = (tgt, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], 10, 2.0)
    abc = ob.synthesize([A, B, C], *args)
    singles = [ob.synthesize([s], *args) for s in (A, B, C)]
    q = np.stack([s["quality"] for s in singles]).astype(np.float64)
    c = np.stack([s["color"] for s in singles]).astype(np.float64)
    all3 = (q > 0).all(0)
    assert all3.sum() > 100
    expect_q = np.sqrt((q ** 2).sum(0))
    assert np.allclose(abc["quality"][all3], expect_q[all3], rtol=1e-5)
    wts = q ** 2
    expect_c = (wts[..., None] * c).sum(0) / wts.sum(0)[..., None]
    assert np.allclose(abc["color"][all3], expect_c[all3], rtol=1e-5, atol=1e-6)


# ----------------------------------------------- seam triangles, prolongation

def brute_force_wrap(dep, uv, W, H, wrap):
    """Per pixel, every covering triangle of the reference traversal (quad
    gating of transform.cpp:161-167, seam triangles at j == 0 when wrapping),
    evaluated in float64: (seq, v, d, prol)."""
    h, w = dep.shape
    cands = {}
    seq = 0

    def tri(a, b, c):
        nonlocal seq
        seq += 1
        P = [uv[p].astype(np.float64) for p in (a, b, c)]
        D = [float(dep[p]) for p in (a, b, c)]
        A, B, C = P
        den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
        if den <= 0:
            return
        st = max(np.hypot(*(A - B)), np.hypot(*(A - C)), np.hypot(*(B - C)))
        v = min(10000.0, max(1.0, 10000 - 1000 * st))
        prol = any(x > 0 for x in D)
        for y in range(H):
            for x in range(W):
                l1 = ((B[1] - C[1]) * (x + .5 - C[0]) + (C[0] - B[0]) * (y + .5 - C[1])) / den
                l2 = ((C[1] - A[1]) * (x + .5 - C[0]) + (A[0] - C[0]) * (y + .5 - C[1])) / den
                l3 = 1 - l1 - l2
                if min(l1, l2, l3) >= -1e-6:
                    d = D[0] * l1 + D[1] * l2 + D[2] * l3
                    cands.setdefault((y, x), []).append((seq, v, d, prol))

    for i in range(h - 1):
        for j in range(w - 1):
            if dep[i, j + 1] > 0 and dep[i + 1, j] > 0 and uv[i, j + 1, 0] > 0 and uv[i + 1, j, 0] > 0:
                if dep[i, j] > 0 and uv[i, j, 0] > 0:
                    tri((i, j), (i, j + 1), (i + 1, j))
                if dep[i + 1, j + 1] > 0 and uv[i + 1, j + 1, 0] > 0:
                    tri((i + 1, j + 1), (i + 1, j), (i, j + 1))
                if wrap and j == 0:
                    tri((i, w - 1), (i, 0), (i + 1, w - 1))
                    tri((i + 1, 0), (i + 1, w - 1), (i, 0))
    return cands


def expected_winner(cs):
    """The winner by mathematics rather than by replaying the fold: if any
    covering triangle has prol false, the first of them is adopted whatever
    came before (the flag is still set), and afterwards only prol-false
    triangles with a strictly better score v/d^3 replace it; otherwise the
    largest v/d^3 wins, the earliest on ties.  Returns None when the margin to
    the runner-up is below 1e-5 (float rounding could decide)."""
    pf = [c for c in cs if not c[3]]
    pool = pf if pf else cs
    if pf:
        # prol-false candidates in these tests have a single member per pixel
        assert len(pf) == 1
        return pf[0]
    scored = sorted(pool, key=lambda c: (-c[1] / c[2] ** 3, c[0]))
    if len(scored) > 1:
        s0 = scored[0][1] / scored[0][2] ** 3
        s1 = scored[1][1] / scored[1][2] ** 3
        if s1 > s0 * (1 - 1e-5):
            return None
    return scored[0]


def wrap_mesh(rng, h, w, W, jitter=0.25):
    """A mesh laid out like an ERP source seen by a wrapping target: column j
    maps to u = (j + 1) * sx (j < w-1) and the last column wraps to the left
    edge at u = 0.3 * sx, so the seam triangles between columns w-1 and 0 cover
    the strip u in [0.3 sx, sx]."""
    sx = W / float(w)
    gy, gx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = (gx + 1.0) * sx
    u[:, w - 1] = 0.3 * sx
    vv = (gy + 0.5) * 1.7
    uv = np.stack([u, vv], -1) + rng.normal(0, jitter, (h, w, 2))
    uv[:, w - 1, 0] = np.maximum(uv[:, w - 1, 0], 0.05)
    return uv.astype(np.float32)


def test_raster_wrap_seam_triangles_match_brute_force():
    """wrap = 1 on a random mesh with all depths positive: the seam triangles
    (transform.cpp:170-175) cover the strip between the last and the first
    column, and every pixel's winner is the brute-force argmax of v/d^3."""
    rng = np.random.default_rng(16)
    h, w, W, H = 6, 7, 16, 11
    uv = wrap_mesh(rng, h, w, W)
    dep = rng.uniform(1.0, 4.0, (h, w)).astype(np.float32)
    col = rng.random((h, w, 3)).astype(np.float32)
    oc, od, osh = ob.transform_triangles(col, dep, uv, W, H, 1)
    oc0, od0, osh0 = ob.transform_triangles(col, dep, uv, W, H, 0)
    cands = brute_force_wrap(dep, uv, W, H, 1)
    checked = seam_px = 0
    for (y, x), cs in cands.items():
        win = expected_winner(cs)
        if win is None:
            continue
        assert abs(od[y, x] - win[2]) < 1e-5 * abs(win[2]) and abs(osh[y, x] - win[1]) < 1e-2, (y, x)
        checked += 1
        if np.isinf(od0[y, x]):
            seam_px += 1          # covered only by a seam triangle
    assert checked > 60 and seam_px > 5
    assert all(np.isinf(od[y, x]) for y in range(H) for x in range(W) if (y, x) not in cands)


def test_raster_prol_false_seam_triangle_overwrites_and_locks_the_pixel():
    """Seam triangle (w-1, i), (0, i), (w-1, i+1) with all three vertex depths
    <= 0 (prol false, transform.cpp:121-134).  It is drawn (only the quad's
    (i, 1) and (i+1, 0) vertices are gated), and on every pixel it covers it
    replaces whatever triangle was drawn before it (the flag is still set) and
    keeps the pixel against every triangle drawn after it (the flag is
    cleared), although its score v/d^3 (d <= 0) is below every other
    triangle's.  The mesh is random, so many large triangles overlap it on
    both sides of the traversal."""
    rng = np.random.default_rng(17)
    h, w, W, H = 5, 6, 16, 12
    uv = np.stack([rng.uniform(0.2, W, (h, w)), rng.uniform(0.0, H, (h, w))], -1).astype(np.float32)
    dep = rng.uniform(1.0, 2.0, (h, w)).astype(np.float32)
    i = 2
    dep[i, 0], dep[i, w - 1], dep[i + 1, w - 1] = 0.0, -1.0, -2.0
    uv[i, w - 1] = (2.0, 2.0)
    uv[i, 0] = (W - 3.0, 2.0)
    uv[i + 1, w - 1] = (2.0, H - 3.0)
    for a in ((i, 1), (i + 1, 0)):      # the gated vertices of quad (i, 0)
        assert dep[a] > 0 and uv[a][0] > 0
    col = rng.random((h, w, 3)).astype(np.float32)
    oc, od, osh = ob.transform_triangles(col, dep, uv, W, H, 1)
    cands = brute_force_wrap(dep, uv, W, H, 1)
    locked = [p for p, cs in cands.items() if any(not c[3] for c in cs)]
    before = after = 0
    for (y, x) in locked:
        cs = cands[(y, x)]
        pf = [c for c in cs if not c[3]]
        assert len(pf) == 1
        pf = pf[0]
        assert abs(od[y, x] - pf[2]) <= 1e-5 * max(1.0, abs(pf[2])), (y, x, od[y, x], pf)
        assert pf[2] <= 0 and abs(osh[y, x] - pf[1]) < 1e-2
        before += any(c[3] and c[0] < pf[0] for c in cs)
        after += any(c[3] and c[0] > pf[0] for c in cs)
    assert len(locked) > 20 and before > 5 and after > 5
    # every other pixel whose candidates all lie in front (d > 0: the second
    # seam triangle mixes the signs of its vertex depths) follows the argmax rule
    checked = 0
    for (y, x), cs in cands.items():
        if (y, x) in locked or any(c[2] <= 0 for c in cs):
            continue
        win = expected_winner(cs)
        if win is not None:
            assert abs(od[y, x] - win[2]) < 1e-5 * abs(win[2]), (y, x)
            checked += 1
    assert checked > 20
