"""GPU parity: every stage of the HIP path against the CPU oracle, through the C ABI.

The bar: bit-exact for decode, raster (on identical inputs), blend (exponent 1
and blend-by-max), inpaint, encode and for the fused path on perspective
targets; reprojection onto ERP targets within the stated bound (float atan2 /
asin of the device library vs the host's).
"""
import sys

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.uint32)
    b = np.ascontiguousarray(b, np.float32).view(np.uint32)
    an = np.isnan(a.view(np.float32))
    bn = np.isnan(b.view(np.float32))
    same = (a == b) | (an & bn)
    return bool(same.all()), int((~same).sum())


# ------------------------------------------------------------------ decode

@pytest.mark.parametrize("bd", [8, 10, 16])
def test_decode_color_depth(torch_cuda, bd):
    torch = torch_cuda
    say("case decode bd=%d" % bd)
    rng = np.random.default_rng(bd)
    h, w = 96, 160
    dt = R.plane_dtype(bd)
    mx = R.max_level(bd)
    y = rng.integers(0, mx + 1, (h, w)).astype(dt)
    u = rng.integers(0, mx + 1, (h // 2, w // 2)).astype(dt)
    v = rng.integers(0, mx + 1, (h // 2, w // 2)).astype(dt)
    d = rng.integers(0, mx + 1, (h, w)).astype(dt)
    cs = R.scale_of(bd)
    out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    R.decode_yuv420(to_dev(torch, y), to_dev(torch, u), to_dev(torch, v), bd, cs, out)
    ref = ob.decode_color(y, u, v, bd, cs)
    ok, n = bits_equal(out.cpu().numpy(), ref)
    assert ok, "decode colour: %d mismatches" % n
    for (zn, zf) in ((0.5, 25.0), (0.8, 1000.0), (1.0, 10.0)):
        outd = torch.empty((h, w), dtype=torch.float32, device="cuda")
        R.decode_depth(to_dev(torch, d), bd, cs, zn, zf, True, outd)
        refd = ob.decode_depth(d, bd, zn, zf, True, cs)
        ok, n = bits_equal(outd.cpu().numpy(), refd)
        assert ok, "decode depth (%g, %g): %d mismatches" % (zn, zf, n)


# ------------------------------------------------------------- reproject

def _reproject_case(torch, src_cam, dst_cam, src_pos, src_rot, dst_pos, dst_rot, depth):
    Rm, t = R.relative_pose(src_pos, src_rot, dst_pos, dst_rot)
    Ro, to = ob.relative_pose(src_pos, src_rot, dst_pos, dst_rot)
    assert bits_equal(Rm, Ro)[0] and bits_equal(t, to)[0], "host pose differs from the oracle"
    tab = to_dev(torch, R.erp_table(src_cam)) if src_cam.proj == 1 else None
    uv = torch.empty(depth.shape + (2,), dtype=torch.float32, device="cuda")
    vd = torch.empty(depth.shape, dtype=torch.float32, device="cuda")
    R.reproject(src_cam, dst_cam, Rm, t, to_dev(torch, depth), tab, 1.0, uv, vd)
    ruv, rvd, _ = ob.reproject(W.oracle_cam(src_cam, ob), W.oracle_cam(dst_cam, ob), Ro, to, depth)
    return uv.cpu().numpy(), vd.cpu().numpy(), ruv, rvd


def test_reproject_perspective_exact(torch_cuda):
    torch = torch_cuda
    say("case reproject persp->persp")
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    s = rg["sources"][0]
    d = ob.decode_depth(planes[0][3], 16, s["znear"], s["zfar"])
    uv, vd, ruv, rvd = _reproject_case(torch, s["cam"], rg["target"]["cam"], s["pos"], s["rot"], [0.01, 0.0, 0.02],
                                       [3.0, -2.0, 1.0], d)
    assert bits_equal(uv, ruv)[0] and bits_equal(vd, rvd)[0]


def test_reproject_erp_to_perspective_exact(torch_cuda):
    torch = torch_cuda
    say("case reproject erp->persp")
    rg = W.rig("erp4_1080p", scale=8)
    planes = W.render_sources(rg)
    s = rg["sources"][1]
    d = ob.decode_depth(planes[1][3], 16, s["znear"], s["zfar"])
    uv, vd, ruv, rvd = _reproject_case(torch, s["cam"], rg["target"]["cam"], s["pos"], s["rot"], [0.0, 0.05, 0.0],
                                       [20.0, 5.0, 0.0], d)
    ok1, n1 = bits_equal(uv, ruv)
    ok2, n2 = bits_equal(vd, rvd)
    assert ok1 and ok2, "uv mismatches %d, depth mismatches %d" % (n1, n2)


def test_reproject_to_erp_bound(torch_cuda):
    """ERP targets use the device atan2f/asinf: stated bound vs the oracle and an FP64 restatement."""
    torch = torch_cuda
    say("case reproject ->erp")
    rg = W.rig("erp4_1080p", scale=8)
    planes = W.render_sources(rg)
    s = rg["sources"][0]
    d = ob.decode_depth(planes[0][3], 16, s["znear"], s["zfar"])
    dst = R.cam(1, 1024, 512)
    uv, vd, ruv, rvd = _reproject_case(torch, s["cam"], dst, s["pos"], s["rot"], [0.0, 0.0, 0.0], [30.0, 0.0, 0.0], d)
    fin = np.isfinite(ruv).all(-1)
    err_uv = np.abs(uv - ruv)[fin]
    # u spans 1024 px over 2*pi: one float ulp of atan2 near pi is ~4e-7 rad -> 7e-5 px
    assert err_uv.max() < 2e-3, err_uv.max()
    ok, n = bits_equal(vd, rvd)       # the radius is the exact double-emulated norm
    assert ok, n


# ------------------------------------------------------------------ raster

def _raster_case(torch, color, depth, uv, out_w, out_h, wrap, tag):
    say("case raster %s" % tag)
    oc, od, osh = ob.transform_triangles(color, depth, uv, out_w, out_h, wrap)
    gc = torch.empty((out_h, out_w, 3), dtype=torch.float32, device="cuda")
    gd = torch.empty((out_h, out_w), dtype=torch.float32, device="cuda")
    gs = torch.empty((out_h, out_w), dtype=torch.float32, device="cuda")
    R.raster(to_dev(torch, color), to_dev(torch, depth), to_dev(torch, uv), out_w, out_h, wrap, 1.0, gc, gd, gs)
    for name, a, b in (("color", gc, oc), ("depth", gd, od), ("shape", gs, osh)):
        ok, n = bits_equal(a.cpu().numpy(), b)
        assert ok, "%s: %s %d mismatches" % (tag, name, n)


def _stage_inputs(rg, idx, planes, tgt_pos, tgt_rot, dst_cam=None):
    s = rg["sources"][idx]
    y, u, v, d = planes[idx]
    col = ob.decode_color(y, u, v, s["bd_color"])
    dep = ob.decode_depth(d, s["bd_depth"], s["znear"], s["zfar"])
    Ro, to = ob.relative_pose(s["pos"], s["rot"], tgt_pos, tgt_rot)
    dst = W.oracle_cam(dst_cam or rg["target"]["cam"], ob)
    uv, vd, wrap = ob.reproject(W.oracle_cam(s["cam"], ob), dst, Ro, to, dep)
    return col, vd, uv, wrap, dst


def test_raster_rects(torch_cuda):
    torch = torch_cuda
    rg = W.rig("rects_640")
    planes = W.render_sources(rg)
    for idx, (pos, rot) in enumerate((([0, 0, 0], [0, 0, 0]), ([0.02, -0.01, 0.03], [4.0, -3.0, 2.0]))):
        col, vd, uv, wrap, dst = _stage_inputs(rg, idx, planes, pos, rot)
        _raster_case(torch, col, vd, uv, dst.w, dst.h, wrap, "rects src%d" % idx)


def test_raster_erp_to_perspective(torch_cuda):
    torch = torch_cuda
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    col, vd, uv, wrap, dst = _stage_inputs(rg, 2, planes, [0.0, 0.0, 0.0], [10.0, 5.0, 0.0])
    _raster_case(torch, col, vd, uv, dst.w, dst.h, wrap, "erp->persp")


def test_raster_erp_to_erp_wrap(torch_cuda):
    torch = torch_cuda
    rg = W.rig("erp4_1080p", scale=8)
    planes = W.render_sources(rg)
    col, vd, uv, wrap, dst = _stage_inputs(rg, 0, planes, [0.0, 0.0, 0.0], [90.0, 0.0, 0.0], R.cam(1, 512, 256))
    assert wrap == 1
    _raster_case(torch, col, vd, uv, dst.w, dst.h, wrap, "erp->erp wrap")


def test_raster_identity_warp_all_ties(torch_cuda):
    """Source camera == target camera: every pixel centre is (up to float
    rounding of the reprojection) a vertex shared by six triangles, so the
    winners are near ties and every pixel goes through the in-order replay."""
    torch = torch_cuda
    rg = W.rig("rects_640", scale=4)
    planes = W.render_sources(rg)
    s = rg["sources"][0]
    col, vd, uv, wrap, dst = _stage_inputs(rg, 0, planes, s["pos"], s["rot"])
    _raster_case(torch, col, vd, uv, dst.w, dst.h, wrap, "identity")


def test_raster_exact_grid_identity(torch_cuda):
    """Closed form: vertices exactly on the pixel centres reproduce the source
    colour and depth at every pixel (barycentric weight exactly 1)."""
    torch = torch_cuda
    rng = np.random.default_rng(9)
    h, w = 48, 64
    gy, gx = np.mgrid[0:h, 0:w]
    uv = np.stack([gx + 0.5, gy + 0.5], -1).astype(np.float32)
    depth = rng.uniform(1.0, 5.0, (h, w)).astype(np.float32)
    col = rng.random((h, w, 3)).astype(np.float32)
    _raster_case(torch, col, depth, uv, w, h, 0, "exact grid")
    oc, od, osh = ob.transform_triangles(col, depth, uv, w, h, 0)
    assert np.array_equal(oc, col) and np.array_equal(od, depth)


def test_raster_random_triangles(torch_cuda):
    """Random positions, depths and NaN/inf vertices: exercises specials, giant
    triangles and near ties."""
    torch = torch_cuda
    rng = np.random.default_rng(7)
    for case in range(4):
        ih, iw = 40, 50
        uv = np.stack([rng.uniform(-10, 90, (ih, iw)), rng.uniform(-10, 70, (ih, iw))], -1).astype(np.float32)
        if case >= 1:  # smooth grid + jitter (mostly small triangles)
            gy, gx = np.mgrid[0:ih, 0:iw]
            uv = np.stack([gx * 1.6 + rng.normal(0, 0.3, gx.shape), gy * 1.5 + rng.normal(0, 0.3, gy.shape)],
                          -1).astype(np.float32)
        depth = rng.uniform(0.5, 20, (ih, iw)).astype(np.float32)
        if case == 2:
            depth[rng.random((ih, iw)) < 0.05] = np.nan
            uv[rng.random((ih, iw)) < 0.02] = np.inf
        if case == 3:
            depth[:] = 3.0   # equal depths: ties decided by validity and order
        col = rng.random((ih, iw, 3)).astype(np.float32)
        _raster_case(torch, col, depth, uv, 80, 64, 1 if case == 3 else 0, "random %d" % case)


def test_raster_dense_bin_beyond_the_replay_buffer(torch_cuda):
    """48,705 source quads whose vertices all land in one 16x16 target tile:
    the tile's bin holds every quad (far beyond the 8,192 entries the in-order
    replay sorts in LDS), most pixels have thousands of candidates (up to ~12k,
    more than the replay's per-pixel lists hold), and the far (2^21) triangles
    of the first two source rows cover most pixels with depths the raster's
    score cannot order, so those pixels take the replay's sorted in-order scan
    of the whole bin.  Bit-exact against the oracle's sequential fold."""
    torch = torch_cuda
    rng = np.random.default_rng(21)
    ih, iw = 192, 256
    uv = np.stack([rng.uniform(0.01, 16, (ih, iw)), rng.uniform(0, 16, (ih, iw))], -1).astype(np.float32)
    depth = rng.uniform(1, 4, (ih, iw)).astype(np.float32)
    depth[:2] = 2.0 ** 21
    col = rng.random((ih, iw, 3)).astype(np.float32)
    _raster_case(torch, col, depth, uv, 16, 16, 0, "dense bin, 48705 quads in one tile")


# ------------------------------------------------------------------- blend

@pytest.mark.parametrize("exp", [1.0, -1.0, 5.0])
def test_blend_steps(torch_cuda, exp):
    torch = torch_cuda
    say("case blend exp=%g" % exp)
    rng = np.random.default_rng(3)
    h, w = 64, 96
    ref = ob.Blender(w, h, exp)
    acc = {"color": torch.zeros((h, w, 3), device="cuda"), "quality": torch.zeros((h, w), device="cuda"),
           "depth": torch.zeros((h, w), device="cuda"), "validity": torch.zeros((h, w), device="cuda"),
           "mask": torch.zeros((h, w), dtype=torch.uint8, device="cuda")}
    for k in range(4):
        c = rng.random((h, w, 3)).astype(np.float32)
        d = rng.uniform(0.5, 10, (h, w)).astype(np.float32)
        sh = rng.uniform(1, 10000, (h, w)).astype(np.float32)
        unc = rng.random((h, w)) < 0.3
        d[unc] = np.inf
        sh[unc] = 0.0
        c[unc] = 0.0
        ref.blend(c, d, sh)
        R.blend_step(acc, to_dev(torch, c), to_dev(torch, d), to_dev(torch, sh), exp, k == 0)
        for name, a, b in (("color", acc["color"], ref.color), ("quality", acc["quality"], ref.quality),
                           ("depth", acc["depth"], ref.depth), ("validity", acc["validity"], ref.validity)):
            a = a.cpu().numpy()
            if exp in (1.0, -1.0):
                ok, n = bits_equal(a, b)
                assert ok, "step %d %s: %d mismatches" % (k, name, n)
            else:   # powf of the device library vs the host's: within a few ulps
                np.testing.assert_allclose(a, b, rtol=2e-6, atol=0)
        assert np.array_equal(acc["mask"].cpu().numpy() != 0, ref.depth_mask != 0)


# ----------------------------------------------------------------- inpaint

def _inpaint_case(torch, quality, tag):
    say("case inpaint %s" % tag)
    h, w = quality.shape
    rng = np.random.default_rng(11)
    col = rng.random((h, w, 3)).astype(np.float32)
    rcol, rmap, sweeps = ob.inpaint(col, quality, want_map=True)
    gcol = to_dev(torch, col)
    gmap = torch.empty((h, w, 3), dtype=torch.int32, device="cuda")
    R.inpaint(gcol, to_dev(torch, quality), map_out=gmap)
    m = gmap.cpu().numpy()
    assert np.array_equal(m, rmap), "%s: map mismatches %d" % (tag, int((m != rmap).any(-1).sum()))
    assert np.array_equal(gcol.cpu().numpy(), rcol)


def test_inpaint_patterns(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    h, w = 120, 160
    for frac in (0.01, 0.1, 0.4, 0.7):
        q = np.where(rng.random((h, w)) < frac, 0.0, 1.0).astype(np.float32)
        _inpaint_case(torch, q, "random %.2f" % frac)
    q = np.ones((h, w), np.float32)
    q[20:90, 30:140] = 0.0      # one big rectangle
    q[:, :7] = 0.0              # left border strip
    q[100:, :] = np.nan         # bottom strip (NaN quality is a hole)
    _inpaint_case(torch, q, "blocks")
    yy, xx = np.mgrid[0:h, 0:w]
    q = np.where(((xx - 80) ** 2 + (yy - 60) ** 2 < 50 ** 2) & ((xx - 80) ** 2 + (yy - 60) ** 2 > 20 ** 2), 0.0, 1.0)
    _inpaint_case(torch, q.astype(np.float32), "ring")
    _inpaint_case(torch, np.zeros((h, w), np.float32), "all holes")
    _inpaint_case(torch, np.ones((h, w), np.float32), "no holes")
    q = np.ones((h, w), np.float32)
    q[::2, ::3] = 0.0
    _inpaint_case(torch, q, "lattice")


def test_inpaint_large_components(torch_cuda):
    """Components far beyond the LDS box, solved in strips of 256 pixels per
    line on separate blocks: lines of 2586 pixels along rows (wide box) and
    columns (tall box), a wedge and a spiral corridor."""
    torch = torch_cuda
    rng = np.random.default_rng(8)
    for tag, (h, w) in (("wide", (70, 2600)), ("tall", (2600, 70))):
        q = np.ones((h, w), np.float32)
        q[5:h - 9, 3:w - 11] = 0.0
        q[rng.random((h, w)) < 0.002] = 1.0   # a few known islands inside
        _inpaint_case(torch, q, tag)
    h, w = 360, 640
    yy, xx = np.mgrid[0:h, 0:w]
    q = np.where(xx * 0.6 + yy > 300, 0.0, 1.0).astype(np.float32)   # an uncovered wedge
    q[rng.random((h, w)) < 0.01] = 0.0
    _inpaint_case(torch, q, "wedge")
    # a corridor winding between known walls (one long component)
    h, w = 300, 2400
    q = np.zeros((h, w), np.float32)
    y0, x0, y1, x1 = 1, 1, h - 2, w - 2
    while y0 <= y1 and x0 <= x1:
        q[y0, x0:x1 + 1] = 1.0
        q[y0:y1 + 1, x1] = 1.0
        if y1 - y0 >= 2:
            q[y1, x0 + 2:x1 + 1] = 1.0
        if x1 - x0 >= 2:
            q[y0 + 2:y1 + 1, x0 + 2] = 1.0
        y0, x0, y1, x1 = y0 + 6, x0 + 6, y1 - 6, x1 - 6
    _inpaint_case(torch, q, "spiral")


def test_inpaint_strip_capacity_overflow(torch_cuda):
    """190 nested L-shaped components whose boxes all span most of a 600 x 4000
    view: their strip carries exceed the workspace's share (npx / 16 slots),
    so the plan places a prefix of them on the strip solver and leaves the rest
    to the one-block solver; every map entry still equals the oracle's."""
    torch = torch_cuda
    h, w = 600, 4000
    q = np.ones((h, w), np.float32)
    for k in range(190):
        a = 3 * k + 1
        q[a, a:w - 2] = 0.0     # the L's row
        q[a:h - 2, a] = 0.0     # and its column
    _inpaint_case(torch, q, "nested L overflow")


def test_inpaint_component_beyond_line_length(torch_cuda):
    """A component whose box exceeds 4096 pixels in both directions (several
    block spans per line)."""
    torch = torch_cuda
    rng = np.random.default_rng(9)
    h = w = 4200
    q = np.ones((h, w), np.float32)
    q[2:h - 3, 4:w - 2] = 0.0
    q[rng.random((h, w)) < 0.0005] = 1.0   # known islands inside
    q[h // 2, : w // 3] = 1.0              # and a wall from the left border
    _inpaint_case(torch, q, "4200x4200")


# ------------------------------------------------------------------ encode

@pytest.mark.parametrize("bd", [8, 10])
def test_encode(torch_cuda, bd):
    torch = torch_cuda
    say("case encode bd=%d" % bd)
    rng = np.random.default_rng(bd)
    h, w = 64, 96
    col = rng.uniform(-0.01, 1.01, (h, w, 3)).astype(np.float32)
    col[0, 0] = [1.0, 1.0 / 1023 * 1023, 0.5]
    ry, ru, rv = ob.encode_yuv420(col, bd)
    dt = torch.uint8 if bd <= 8 else torch.int16
    gy = torch.empty((h, w), dtype=dt, device="cuda")
    gu = torch.empty((h // 2, w // 2), dtype=dt, device="cuda")
    gv = torch.empty((h // 2, w // 2), dtype=dt, device="cuda")
    R.encode_yuv420(to_dev(torch, col), bd, gy, gu, gv)
    npdt = np.uint8 if bd <= 8 else np.uint16
    for a, b in ((gy, ry), (gu, ru), (gv, rv)):
        assert np.array_equal(a.cpu().numpy().view(npdt), b)


# ------------------------------------------------------------ fused view

def _synth_case(torch, rg, pos, rot, tag, exact=True, tile=None):
    say("case synthesize %s, tile %s" % (tag, tile))
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgt = W.device_target(rg, pos, rot)
    syn = R.Synthesizer(srcs, tgt, tile=tile)
    out = syn.run()
    torch.cuda.synchronize()
    t = rg["target"]
    ref = ob.synthesize(W.oracle_sources(rg, planes, ob), W.oracle_cam(t["cam"], ob), pos, rot, t["bitdepth"],
                        t["blending_exp"])
    npdt = np.uint8 if t["bitdepth"] <= 8 else np.uint16
    for k in ("quality", "depth", "validity", "color"):
        ok, n = bits_equal(out[k].cpu().numpy(), ref[k])
        if exact:
            assert ok, "%s: %s %d mismatches" % (tag, k, n)
    for k in ("y", "u", "v"):
        a = out[k].cpu().numpy().view(npdt)
        if exact:
            assert np.array_equal(a, ref[k]), "%s plane %s differs" % (tag, k)
    return out, ref, syn


def test_synthesize_rects_640(torch_cuda):
    _synth_case(torch_cuda, W.rig("rects_640"), [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], "rects_640")


@pytest.mark.parametrize("tile", [16, 32])
def test_synthesize_rects_moved(torch_cuda, tile):
    _synth_case(torch_cuda, W.rig("rects_640", scale=2), [0.03, 0.01, -0.02], [6.0, 2.0, -1.0], "rects moved",
                tile=tile)


@pytest.mark.parametrize("tile", [16, 32])
def test_synthesize_erp4_quarter(torch_cuda, tile):
    _synth_case(torch_cuda, W.rig("erp4_1080p", scale=4), [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], "erp4 quarter",
                tile=tile)


def test_synthesize_repeatable(torch_cuda):
    torch = torch_cuda
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    syn = R.Synthesizer(W.device_sources(rg, planes), W.device_target(rg, [0.05, 0.0, 0.0], [15.0, 0.0, 0.0]))
    a = {k: v.clone() for k, v in syn.run().items() if v is not None}
    b = syn.run()
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k


@pytest.mark.parametrize("precision,tile", [(2.0, 16), (1.5, 16), (0.5, 16), (1.5, 32)])
def test_synthesize_precision(torch_cuda, precision, tile):
    """Precision != 1: rasterise, blend and inpaint at (int)(0.5 + size * p)
    (SynthesizedView.cpp:326-331; validity stretch divided by p), then resize
    to the view's size before encoding (image_writing.cpp:189-190).  Maps at
    the raster size and the encoded planes bit-exact against the oracle (the
    resize rule is the oracle's restatement of OpenCV's, pinned on the CPU);
    the per-stage path gives the same bits."""
    torch = torch_cuda
    say("case synthesize precision %g, tile %d" % (precision, tile))
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    tgt = W.device_target(rg, [0.02, -0.01, 0.01], [3.0, -1.0, 0.5])
    tgt.rescale = precision
    tgt.tile = tile
    srcs = W.device_sources(rg, planes)
    out = R.Synthesizer(srcs, tgt).run()
    torch.cuda.synchronize()
    t = rg["target"]
    ref = ob.synthesize(W.oracle_sources(rg, planes, ob), W.oracle_cam(t["cam"], ob), list(tgt.pos), list(tgt.rot),
                        t["bitdepth"], t["blending_exp"], rescale=precision)
    rw, rh = R.raster_size(tgt)
    assert out["color"].shape[:2] == (rh, rw) and ref["color"].shape[:2] == (rh, rw)
    for k in ("quality", "depth", "validity", "color"):
        ok, n = bits_equal(out[k].cpu().numpy(), ref[k])
        assert ok, "precision %g: %s %d mismatches" % (precision, k, n)
    for k in ("y", "u", "v"):
        assert np.array_equal(out[k].cpu().numpy(), ref[k]), "precision %g: plane %s" % (precision, k)
    sw = R.StagewiseSynthesizer(srcs, tgt).run()
    torch.cuda.synchronize()
    for k in ("y", "u", "v", "color", "depth"):
        assert torch.equal(sw[k].view(torch.int32) if sw[k].dtype == torch.float32 else sw[k],
                           out[k].view(torch.int32) if out[k].dtype == torch.float32 else out[k]), k

