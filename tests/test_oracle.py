"""CPU: the oracle pinned against the reference's own golden values and against
what the mathematics fixes (closed forms, brute force on tiny inputs, an FP64
restatement).  Each check is chosen so that a plausible restatement mistake (a
dropped term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

import oracle_bind as ob

# ------------------------------------------------------------ golden values
# reference rvs/src/unit_test.cpp:66-308 (cameras, expected values, tolerances)


def erp5():
    return ob.make_cam("erp", 5, 5, hor=(-180, 180), ver=(-90, 90))


def persp43():
    return ob.make_cam("persp", 4, 3, f=(3, 3), p=(2, 1))


def erp_reference_image_pos():                 # unit_test.cpp:90-101
    ref = np.zeros((5, 5, 2), np.float32)
    for j in range(5):
        ref[0, j] = (j + 0.5, 0.001)
        ref[1, j] = (j + 0.5, 1.5)
        ref[2, j] = (j + 0.5, 2.5)
        ref[3, j] = (j + 0.5, 3.5)
        ref[4, j] = (j + 0.5, 4.999)
    return ref


def erp_reference_world_pos(pos, depth):       # unit_test.cpp:110-123
    phi = np.pi - 2 * np.pi * (pos[..., 0].astype(np.float64) / 5.0)
    theta = np.pi / 2 - np.pi * (pos[..., 1].astype(np.float64) / 5.0)
    d = depth.astype(np.float64)
    return np.stack([d * np.cos(theta) * np.cos(phi), d * np.cos(theta) * np.sin(phi), d * np.sin(theta)],
                    -1).astype(np.float32)


def ref_depth(shape, i, j):                    # ones with one 2 (unit_test.cpp:103-108, 235-240)
    d = np.ones(shape, np.float32)
    d[i, j] = 2.0
    return d


def test_erp_generate_image_pos():             # unit_test.cpp:132-145
    pos = ob.generate_image_pos(erp5())
    assert np.abs(pos - erp_reference_image_pos()).max() < 1e-6


def test_erp_unproject():                      # unit_test.cpp:147-162
    pos = erp_reference_image_pos()
    d = ref_depth((5, 5), 2, 2)
    xyz = ob.unproject(erp5(), pos, d)
    assert np.abs(xyz - erp_reference_world_pos(pos, d)).max() < 1e-6


def test_erp_project():                        # unit_test.cpp:164-188
    pos = erp_reference_image_pos()
    d = ref_depth((5, 5), 2, 2)
    uv, depth, wrap = ob.project(erp5(), erp_reference_world_pos(pos, d))
    assert np.abs(uv - pos).max() < 1e-4
    assert np.abs(depth - d).max() < 1e-6
    assert wrap == 1                           # WrappingMethod::horizontal


def persp_reference_world_pos(pos, depth):     # unit_test.cpp:242-250
    s = np.stack([np.full(pos.shape[:2], 3.0), 2.0 - pos[..., 0].astype(np.float64),
                  1.0 - pos[..., 1].astype(np.float64)], -1)
    return (s * (depth[..., None].astype(np.float64) / 3.0)).astype(np.float32)


def test_persp_generate_image_pos():           # unit_test.cpp:252-265
    pos = ob.generate_image_pos(persp43())
    gy, gx = np.mgrid[0:3, 0:4]
    ref = np.stack([gx + 0.5, gy + 0.5], -1).astype(np.float32)
    assert np.abs(pos - ref).max() < 1e-6


def test_persp_unproject():                    # unit_test.cpp:267-282
    pos = ob.generate_image_pos(persp43())
    d = ref_depth((3, 4), 2, 2)
    xyz = ob.unproject(persp43(), pos, d)
    assert np.abs(xyz - persp_reference_world_pos(pos, d)).max() < 1e-6


def test_persp_project():                      # unit_test.cpp:284-308
    pos = ob.generate_image_pos(persp43())
    d = ref_depth((3, 4), 2, 2)
    uv, depth, wrap = ob.project(persp43(), persp_reference_world_pos(pos, d))
    assert np.abs(uv - pos).max() < 1e-6
    assert np.abs(depth - d).max() < 1e-6
    assert wrap == 0


def test_persp_unproject_invalid_depth_is_nan():
    pos = ob.generate_image_pos(persp43())
    d = ref_depth((3, 4), 2, 2)
    d[0, 0] = 0.0
    d[1, 1] = np.nan
    xyz = ob.unproject(persp43(), pos, d)
    assert np.isnan(xyz[0, 0]).all() and np.isnan(xyz[1, 1]).all() and np.isfinite(xyz[2, 2]).all()


def test_persp_project_behind_is_nan():
    xyz = np.array([[[-1.0, 0.2, 0.1], [0.0, 0.0, 0.0], [2.0, 0.0, 0.0]]], np.float32)
    uv, depth, _ = ob.project(persp43(), xyz)
    assert np.isnan(uv[0, 0]).all() and np.isnan(uv[0, 1]).all() and np.isnan(depth[0, :2]).all()
    assert uv[0, 2].tolist() == [2.0, 1.0] and depth[0, 2] == 2.0


# ------------------------------------------------------------------- geometry

def fp64_R(rot):
    """Rz(yaw) Ry(pitch) Rx(roll) in double (Parameters.cpp:67-97)."""
    a, b, c = np.radians(np.asarray(rot, np.float64))
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, np.cos(c), -np.sin(c)], [0, np.sin(c), np.cos(c)]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("rot", [(0, 0, 0), (90, 0, 0), (10, 5, -3), (-68.9187, -3.20608, 16.5), (180, 45, 30)])
def test_euler_matches_fp64_and_is_orthonormal(rot):
    R = ob.euler_to_R(rot).astype(np.float64)
    assert np.abs(R - fp64_R(rot)).max() < 2e-6
    assert np.abs(R @ R.T - np.eye(3)).max() < 1e-6
    assert abs(np.linalg.det(R) - 1) < 1e-6


def test_euler_closed_forms():
    assert np.array_equal(ob.euler_to_R((0, 0, 0)), np.eye(3, dtype=np.float32))
    R = ob.euler_to_R((90, 0, 0)).astype(np.float64)   # yaw 90: x -> y (left)
    assert np.abs(R @ [1, 0, 0] - [0, 1, 0]).max() < 1e-6
    R = ob.euler_to_R((0, 90, 0)).astype(np.float64)   # pitch 90: x -> -z
    assert np.abs(R @ [1, 0, 0] - [0, 0, -1]).max() < 1e-6
    R = ob.euler_to_R((0, 0, 90)).astype(np.float64)   # roll 90: y -> z
    assert np.abs(R @ [0, 1, 0] - [0, 0, 1]).max() < 1e-6


def test_relative_pose_maps_points_between_cameras():
    """x_to = R x_from + t must equal R_to^T (R_from x_from + p_from - p_to)."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        pf, pt = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        rf, rt = rng.uniform(-180, 180, 3), rng.uniform(-90, 90, 3)
        R, t = ob.relative_pose(pf, rf, pt, rt)
        x = rng.uniform(-5, 5, 3)
        want = fp64_R(rt).T @ (fp64_R(rf) @ x + pf - pt)
        got = R.astype(np.float64) @ x + t
        assert np.abs(got - want).max() < 5e-5


def test_relative_pose_identity():
    R, t = ob.relative_pose([1, 2, 3], [10, 20, 30], [1, 2, 3], [10, 20, 30])
    assert np.abs(R - np.eye(3)).max() < 1e-6 and np.all(t == 0)


def fp64_reproject(src, dst, rot_src, pos_src, rot_dst, pos_dst, depth):
    """Stage (a) in double precision: the reference within which the float path is bounded."""
    h, w = depth.shape
    gy, gx = np.mgrid[0:h, 0:w].astype(np.float64)
    u, v = gx + 0.5, gy + 0.5
    d = depth.astype(np.float64)
    if src.proj == 1:
        v[0, :], v[-1, :] = 1e-3, h - 1e-3
        phi = np.radians(src.hor[1]) - np.radians(src.hor[1] - src.hor[0]) * u / w
        th = np.radians(src.ver[1]) - np.radians(src.ver[1] - src.ver[0]) * v / h
        xyz = np.stack([d * np.cos(th) * np.cos(phi), d * np.cos(th) * np.sin(phi), d * np.sin(th)], -1)
    else:
        xyz = np.stack([d, -(d / src.f[0]) * (u - src.p[0]), -(d / src.f[1]) * (v - src.p[1])], -1)
    world = xyz @ fp64_R(rot_src).T + np.asarray(pos_src)
    x = (world - np.asarray(pos_dst)) @ fp64_R(rot_dst)
    if dst.proj == 1:
        r = np.linalg.norm(x, axis=-1)
        phi = np.arctan2(x[..., 1], x[..., 0])
        th = np.arcsin(x[..., 2] / r)
        uu = dst.w * dst.hor[1] / (dst.hor[1] - dst.hor[0]) - np.degrees(phi) * dst.w / (dst.hor[1] - dst.hor[0])
        vv = dst.h * dst.ver[1] / (dst.ver[1] - dst.ver[0]) - np.degrees(th) * dst.h / (dst.ver[1] - dst.ver[0])
        return np.stack([uu, vv], -1), r
    ok = x[..., 0] > 0
    uu = np.where(ok, -dst.f[0] * x[..., 1] / x[..., 0] + dst.p[0], np.nan)
    vv = np.where(ok, -dst.f[1] * x[..., 2] / x[..., 0] + dst.p[1], np.nan)
    return np.stack([uu, vv], -1), np.where(ok, x[..., 0], np.nan)


@pytest.mark.parametrize("case", ["erp->persp", "persp->persp", "erp->erp", "persp->erp"])
def test_reproject_within_bound_of_fp64(case):
    """Stated bound of stage (a): max |uv_float32 - uv_fp64| below 0.02 px and
    relative depth error below 2e-6 on these rigs."""
    rng = np.random.default_rng(1)
    s_erp, d_erp = case.startswith("erp"), case.endswith("erp")
    src = ob.make_cam("erp", 256, 128) if s_erp else ob.make_cam("persp", 160, 120, f=(120, 120), p=(80, 60))
    dst = ob.make_cam("erp", 320, 160) if d_erp else ob.make_cam("persp", 200, 150, f=(150, 150), p=(100, 75))
    depth = rng.uniform(0.5, 20, (src.h, src.w)).astype(np.float32)
    ps, rs, pd, rd = [0.1, -0.05, 0.02], [5, -3, 1], [-0.02, 0.03, 0.0], [-4, 2, 0.5]
    R, t = ob.relative_pose(ps, rs, pd, rd)
    uv, vd, _ = ob.reproject(src, dst, R, t, depth)
    ruv, rvd = fp64_reproject(src, dst, rs, ps, rd, pd, depth)
    fin = np.isfinite(ruv).all(-1) & np.isfinite(uv).all(-1)
    # the bound is stated for points that land in the target image (grazing
    # points far outside it are culled by the raster's bounding-box clamp)
    fin &= (ruv[..., 0] > -1) & (ruv[..., 0] < dst.w + 1) & (ruv[..., 1] > -1) & (ruv[..., 1] < dst.h + 1)
    if d_erp:   # exclude the longitude seam, where u jumps by the image width
        fin &= np.abs(uv[..., 0] - ruv[..., 0]) < dst.w / 2
    assert fin.sum() > 1000
    assert np.abs(uv - ruv)[fin].max() < 0.02
    assert (np.abs(vd - rvd) / rvd)[fin].max() < 2e-6


# --------------------------------------------------------------------- decode

def test_decode_color_closed_form():
    """Even sizes: U,V upsampling weights are 0, 1/2, 1, so each output is an
    exact mean of at most four samples; last column/row repeat the edge."""
    bd = 10
    y = np.arange(16, dtype=np.uint16).reshape(4, 4) * 60
    u = np.array([[0, 100], [200, 1023]], np.uint16)
    v = np.array([[7, 9], [11, 13]], np.uint16)
    out = ob.decode_color(y, u, v, bd)
    s = np.float32(1.0) / np.float32(1023)
    U = np.array([[0, 50, 100, 100], [100, 330.75, 561.5, 561.5], [200, 611.5, 1023, 1023],
                  [200, 611.5, 1023, 1023]], np.float32)
    assert np.array_equal(out[..., 1], U * s)
    assert np.array_equal(out[..., 0], y.astype(np.float32) * s)
    assert np.array_equal(out[0, 1, 2], np.float32(8.0) * s)


def test_decode_depth_end_points():
    raw = np.array([[0, 65535, 32768]], np.uint16)
    d = ob.decode_depth(raw, 16, 0.5, 25.0)
    assert d[0, 0] == np.float32(25.0) or abs(d[0, 0] - 25.0) < 4e-6      # raw 0 -> far
    assert abs(d[0, 1] - 0.5) < 1e-7                                         # raw max -> near
    v = np.float64(32768) / 65535
    assert abs(d[0, 2] - 25 * 0.5 / (0.5 + v * 24.5)) < 1e-6
    d = ob.decode_depth(raw, 16, 0.8, 1000.0)                                # far >= 1000: near / v
    assert np.isinf(d[0, 0]) and abs(d[0, 1] - 0.8) < 1e-7


def test_decode_depth_invalid_is_nan_only_at_zero():
    raw = np.array([[0, 1, 2]], np.uint16)
    d = ob.decode_depth(raw, 16, 0.0, 10.0, has_invalid=True)   # near = 0 -> depth 0 -> NaN
    assert np.isnan(d).all()
    d = ob.decode_depth(raw, 16, 0.0, 10.0, has_invalid=False)
    assert np.isnan(d[0, 0]) and (d[0, 1:] == 0).all()          # 0/0 at raw 0, else exactly 0


# --------------------------------------------------------------------- raster

def grid_uv(h, w, sx=1.0, sy=1.0, ox=0.0, oy=0.0):
    gy, gx = np.mgrid[0:h, 0:w]
    return np.stack([(gx + 0.5) * sx + ox, (gy + 0.5) * sy + oy], -1).astype(np.float32)


def test_raster_identity_grid_reproduces_source():
    rng = np.random.default_rng(2)
    h, w = 12, 16
    col = rng.random((h, w, 3)).astype(np.float32)
    dep = rng.uniform(1, 3, (h, w)).astype(np.float32)
    oc, od, osh = ob.transform_triangles(col, dep, grid_uv(h, w), w, h, 0)
    assert np.array_equal(oc, col) and np.array_equal(od, dep)
    # every triangle has unit legs: stretch sqrt(2) -> validity 10000 - 1000*sqrt(2)
    assert np.allclose(osh, 10000 - 1000 * np.sqrt(2), rtol=0, atol=1e-3)


def test_raster_mirrored_grid_is_culled():
    h, w = 8, 8
    uv = grid_uv(h, w)
    uv[..., 0] = w - uv[..., 0]           # mirror -> negative orientation -> den <= 0
    oc, od, osh = ob.transform_triangles(np.ones((h, w, 3), np.float32), np.ones((h, w), np.float32), uv, w, h, 0)
    assert np.isinf(od).all() and (osh == 0).all() and (oc == 0).all()


def test_raster_shift_by_integer():
    """A uniform integer shift of the mesh moves the image (closed form)."""
    rng = np.random.default_rng(3)
    h, w = 10, 12
    col = rng.random((h, w, 3)).astype(np.float32)
    dep = np.full((h, w), 2.0, np.float32)
    oc, od, osh = ob.transform_triangles(col, dep, grid_uv(h, w, ox=3.0, oy=2.0), w + 3, h + 2, 0)
    assert np.array_equal(oc[2:, 3:], col) and np.isinf(od[:2]).all() and np.isinf(od[:, :3]).all()


def test_raster_nearer_surface_wins_and_stretch_validity():
    """Two overlapping meshes: the nearer one wins (same validity); a stretched
    mesh has validity 10000 - 1000*stretch (closed form)."""
    h, w = 2, 2
    uv = np.array([[[1.0, 1.0], [5.0, 1.0]], [[1.0, 5.0], [5.0, 5.0]]], np.float32)   # legs 4, diagonal 4*sqrt2
    col = np.ones((h, w, 3), np.float32)
    oc, od, osh = ob.transform_triangles(col, np.full((h, w), 3.0, np.float32), uv, 8, 8, 0)
    cov = np.isfinite(od)
    assert cov.sum() > 10
    assert np.allclose(osh[cov], 10000 - 1000 * 4 * np.sqrt(2), atol=1e-2)
    # stretch beyond 10 -> clamped to validity 1
    oc, od, osh = ob.transform_triangles(col, np.full((h, w), 3.0, np.float32), uv * 3, 20, 20, 0)
    assert (osh[np.isfinite(od)] == 1.0).all()


def brute_force_raster(col, dep, uv, W, H):
    """Independent evaluation by math: per pixel, the covering triangle (same
    coverage predicate) with the largest v/d^3, lowest traversal index on ties;
    valid when no two candidates are within 1e-5 of each other."""
    h, w = dep.shape
    cands = {}
    seq = 0
    for i in range(h - 1):
        for j in range(w - 1):
            for tri in (((i, j), (i, j + 1), (i + 1, j)), ((i + 1, j + 1), (i + 1, j), (i, j + 1))):
                P = [uv[a].astype(np.float64) for a in tri]
                D = [float(dep[a]) for a in tri]
                A, B, C = P
                den = (B[1] - C[1]) * (A[0] - C[0]) + (C[0] - B[0]) * (A[1] - C[1])
                seq += 1
                if den <= 0:
                    continue
                st = max(np.hypot(*(A - B)), np.hypot(*(A - C)), np.hypot(*(B - C)))
                v = min(10000.0, max(1.0, 10000 - 1000 * st))
                for y in range(H):
                    for x in range(W):
                        l1 = ((B[1] - C[1]) * (x + .5 - C[0]) + (C[0] - B[0]) * (y + .5 - C[1])) / den
                        l2 = ((C[1] - A[1]) * (x + .5 - C[0]) + (A[0] - C[0]) * (y + .5 - C[1])) / den
                        l3 = 1 - l1 - l2
                        if min(l1, l2, l3) >= -1e-6:
                            d = D[0] * l1 + D[1] * l2 + D[2] * l3
                            cands.setdefault((y, x), []).append((v / d ** 3, -seq, d, v))
    return cands


def test_raster_matches_brute_force_argmax():
    rng = np.random.default_rng(4)
    h, w, W, H = 6, 7, 14, 12
    uv = (grid_uv(h, w, 2.0, 1.8) + rng.normal(0, 0.4, (h, w, 2))).astype(np.float32)
    dep = rng.uniform(1.0, 4.0, (h, w)).astype(np.float32)
    col = rng.random((h, w, 3)).astype(np.float32)
    oc, od, osh = ob.transform_triangles(col, dep, uv, W, H, 0)
    cands = brute_force_raster(col, dep, uv, W, H)
    checked = 0
    for (y, x), cs in cands.items():
        cs.sort(reverse=True)
        if len(cs) > 1 and cs[1][0] > cs[0][0] * (1 - 1e-5):
            continue
        assert abs(od[y, x] - cs[0][2]) < 1e-5 * cs[0][2] and abs(osh[y, x] - cs[0][3]) < 1e-2
        checked += 1
    assert checked > 50
    assert all(np.isinf(od[y, x]) for y in range(H) for x in range(W) if (y, x) not in cands)


# ---------------------------------------------------------------------- blend

def test_blend_single_view_is_identity():
    rng = np.random.default_rng(5)
    h, w = 8, 9
    c = rng.random((h, w, 3)).astype(np.float32)
    d = rng.uniform(1, 5, (h, w)).astype(np.float32)
    s = rng.uniform(1, 10000, (h, w)).astype(np.float32)
    b = ob.Blender(w, h, 1.0)
    b.blend(c, d, s)
    assert np.array_equal(b.color, c) and np.array_equal(b.quality, s / d) and np.array_equal(b.validity, s)


def test_blend_two_equal_quality_views_is_mean_and_sum_quality():
    h, w = 4, 5
    c1 = np.full((h, w, 3), 0.2, np.float32)
    c2 = np.full((h, w, 3), 0.6, np.float32)
    d = np.full((h, w), 2.0, np.float32)
    s = np.full((h, w), 4000.0, np.float32)
    b = ob.Blender(w, h, 1.0)
    b.blend(c1, d, s)
    b.blend(c2, d, s)
    assert np.allclose(b.color, 0.4, atol=1e-6)
    assert np.allclose(b.quality, 4000.0, atol=1e-3)   # (q1 + q2)^(1/a) with a = 1
    assert (b.validity == 4000.0).all()


def test_blend_empty_colour_where_uncovered():
    h, w = 3, 3
    inf = np.full((h, w), np.inf, np.float32)
    z = np.zeros((h, w), np.float32)
    b = ob.Blender(w, h, 1.0)
    b.blend(np.zeros((h, w, 3), np.float32), inf, z)
    b.blend(np.zeros((h, w, 3), np.float32), inf, z)
    assert (b.color == np.array([0, 1, 0], np.float32)).all() and (b.quality == 0).all()


def test_blend_weights_follow_exponent():
    h, w = 1, 1
    c1 = np.zeros((h, w, 3), np.float32)
    c2 = np.ones((h, w, 3), np.float32)
    d = np.ones((h, w), np.float32)
    b = ob.Blender(w, h, 2.0)
    b.blend(c1, d, np.full((h, w), 1.0, np.float32))
    b.blend(c2, d, np.full((h, w), 3.0, np.float32))
    assert abs(b.color[0, 0, 0] - 9.0 / 10.0) < 1e-6         # weights 1^2, 3^2
    assert abs(b.quality[0, 0] - np.sqrt(10.0)) < 1e-5
    bm = ob.Blender(w, h, -1.0)                               # blend-by-max
    bm.blend(c1, d, np.full((h, w), 1.0, np.float32))
    bm.blend(c2, d, np.full((h, w), 3.0, np.float32))
    assert bm.color[0, 0, 0] == 1.0 and bm.quality[0, 0] == 3.0


# -------------------------------------------------------------------- inpaint

def bfs_dist(hole):
    """Exact L1 distance to the nearest known pixel (brute force)."""
    ky, kx = np.nonzero(~hole)
    h, w = hole.shape
    out = np.zeros((h, w), np.int64)
    for y in range(h):
        for x in range(w):
            if hole[y, x]:
                out[y, x] = (np.abs(ky - y) + np.abs(kx - x)).min() if len(ky) else w + h
    return out


def test_inpaint_distance_is_exact_l1_and_source_is_known():
    rng = np.random.default_rng(6)
    for frac in (0.1, 0.5, 0.85):
        h, w = 17, 23
        hole = rng.random((h, w)) < frac
        hole[0, 0] = False
        q = np.where(hole, 0.0, 1.0).astype(np.float32)
        col = rng.random((h, w, 3)).astype(np.float32)
        out, m, sweeps = ob.inpaint(col, q, want_map=True)
        assert np.array_equal(m[..., 2], np.where(hole, bfs_dist(hole), 0))
        sx, sy = m[..., 0], m[..., 1]
        assert (~hole[sy, sx]).all()
        assert (np.abs(sx - np.arange(w)[None]) + np.abs(sy - np.arange(h)[:, None]) == m[..., 2]).all()
        assert np.array_equal(out, col[sy, sx])
        assert sweeps <= 6


def test_inpaint_tie_choices_by_hand():
    """1x5 row K H H H K: the forward sweep (x ascending) gives the holes the
    left source at distance 1, 2, 3; the backward sweep improves x=3 (distance 1
    from the right) and keeps x=2 (tie at 2 keeps the current value)."""
    q = np.array([[1, 0, 0, 0, 1]], np.float32)
    col = np.array([[[1, 1, 1], [0, 0, 0], [0, 0, 0], [0, 0, 0], [2, 2, 2]]], np.float32)
    out, m, sweeps = ob.inpaint(col, q, want_map=True)
    assert m[0, :, 0].tolist() == [0, 0, 0, 4, 4] and m[0, :, 2].tolist() == [0, 1, 2, 1, 0]
    assert out[0, :, 0].tolist() == [1, 1, 1, 2, 2]
    # column 5x1: K H H H K -> the forward sweep runs y ascending inside each x
    out, m, _ = ob.inpaint(col.transpose(1, 0, 2).copy(), q.T.copy(), want_map=True)
    assert m[:, 0, 1].tolist() == [0, 0, 0, 4, 4]
    # a hole with a known pixel above (up, rank 2) and to the left (left, rank 1)
    # at the same distance takes the left one (first in the neighbour order)
    q = np.array([[1, 1], [1, 0]], np.float32)
    out, m, _ = ob.inpaint(np.zeros((2, 2, 3), np.float32), q, want_map=True)
    assert m[1, 1].tolist() == [0, 1, 1]


def _spiral(h, w):
    """A one-pixel-wide known wall winding inwards: holes form a long corridor."""
    hole = np.ones((h, w), bool)
    y0, x0, y1, x1 = 1, 1, h - 2, w - 2
    while y0 <= y1 and x0 <= x1:
        hole[y0, x0:x1 + 1] = False
        hole[y0:y1 + 1, x1] = False
        if y1 - y0 >= 2:
            hole[y1, x0 + 2:x1 + 1] = False
        if x1 - x0 >= 2:
            hole[y0 + 2:y1 + 1, x0 + 2] = False
        y0, x0, y1, x1 = y0 + 4, x0 + 4, y1 - 4, x1 - 4
    return hole


def test_inpaint_two_sweeps_reach_the_fixed_point():
    """The alternating sweeps of compute_nearest stop after at most three
    sweeps, the third changing nothing: forward + backward already give every
    hole its exact L1 distance (two-pass city-block distance transform).  The
    device solvers stop after two sweeps on that ground.  Shapes that would
    need more sweeps if paths had to bend through holes: spirals, combs, thin
    diagonals, sparse and dense random holes, single known pixels."""
    rng = np.random.default_rng(11)
    shapes = []
    for h, w in ((23, 31), (40, 17), (64, 64)):
        shapes.append(_spiral(h, w))
        comb = np.ones((h, w), bool)
        comb[::3, : w - 2] = False            # teeth from the left
        comb[1::6, 2:] = False                # and from the right
        shapes.append(comb)
        diag = np.ones((h, w), bool)
        for k in range(0, min(h, w), 2):
            diag[k, k] = False
        shapes.append(diag)
        one = np.ones((h, w), bool)
        one[rng.integers(h), rng.integers(w)] = False
        shapes.append(one)
        for frac in (0.3, 0.7, 0.95, 0.995):
            shapes.append(rng.random((h, w)) < frac)
        blob = np.ones((h, w), bool)
        blob[rng.random((h, w)) < 0.02] = False
        shapes.append(blob)
    for hole in shapes:
        q = np.where(hole, 0.0, 1.0).astype(np.float32)
        col = rng.random(hole.shape + (3,)).astype(np.float32)
        out, m, sweeps = ob.inpaint(col, q, want_map=True)
        assert sweeps <= 3, sweeps
        if (~hole).any():
            assert np.array_equal(m[..., 2], np.where(hole, bfs_dist(hole), 0))


def test_inpaint_no_known_pixel_keeps_colour():
    col = np.random.default_rng(7).random((4, 5, 3)).astype(np.float32)
    out, m, sweeps = ob.inpaint(col, np.zeros((4, 5), np.float32), want_map=True)
    assert np.array_equal(out, col) and (m[..., 2] == 9).all() and sweeps == 1


# --------------------------------------------------------------------- encode

def test_encode_closed_form():
    h, w = 4, 4
    col = np.zeros((h, w, 3), np.float32)
    col[..., 0] = np.linspace(0, 1, 16, dtype=np.float32).reshape(4, 4)
    col[0:2, 0:2, 1] = np.array([[1, 2], [3, 5]], np.float32) / 1023   # sum 11 -> (11 + 2) >> 2 = 3
    col[..., 2] = 0.5
    y, u, v = ob.encode_yuv420(col, 10)
    assert np.array_equal(y, np.trunc(col[..., 0] * np.float32(1023)).astype(np.uint16))
    q = np.trunc(col[0:2, 0:2, 1] * np.float32(1023)).astype(np.int64)
    assert u[0, 0] == (q.sum() + 2) >> 2
    assert (v == 511).all()
    # fl(1/1023)*1023 can land just below an integer: truncation loses one code
    s = np.float32(1.0) / np.float32(1023)
    vals = (np.arange(1024, dtype=np.float32) * s)
    col = np.zeros((2, 1024, 3), np.float32)
    col[:, :, 0] = vals
    y, _, _ = ob.encode_yuv420(col, 10)
    assert (y[0].astype(np.int64) - np.arange(1024) <= 0).all()
    assert (y[0].astype(np.int64) - np.arange(1024) >= -1).all()


def test_encode_negative_and_over_range_follow_32bit_truncation():
    col = np.zeros((2, 2, 3), np.float32)
    col[0, 0, 0] = -0.0001     # -> 0
    col[0, 1, 0] = 1.2         # 1227.6 -> 1227 (no saturation, as static_cast<ushort>)
    y, _, _ = ob.encode_yuv420(col, 10)
    assert y[0, 0] == 0 and y[0, 1] == 1227
