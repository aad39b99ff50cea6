"""Camera rigs of the measured workloads, with synthetic sources.

rig_rects_640:  two perspective 640x480 8-bit sources at lateral offsets +-0.05 m
                (f = 500, principal point at the centre, 16-bit inverse depth in
                [1, 10] m) and one 640x480 target between them; scene: rectangles
                before a background plane (seed 1).
rig_erp4_1080p: four equirectangular 4096x2048 10-bit sources on a 0.3 m square
                (16-bit inverse depth in [0.5, 25] m) and a 1920x1080 perspective
                target with a 90 degree horizontal field of view (f = 960) at the
                square's centre; scene: spheres in a box (seed 2).
"""
import numpy as np

import rvs_amd as R


def rig(name, scale=1):
    """Returns dict(sources=[...], target=dict(...)).  `scale` divides every
    resolution (and focal length) for small parity cases."""
    if name == "rects_640":
        w, h, f = 640 // scale, 480 // scale, 500.0 / scale
        srcs = []
        for y in (0.05, -0.05):
            srcs.append(dict(cam=R.cam(0, w, h, f=(f, f), p=(w / 2, h / 2)), pos=[0.0, y, 0.0], rot=[0.0, 0.0, 0.0],
                             bd_color=8, bd_depth=16, znear=1.0, zfar=10.0))
        tgt = dict(cam=R.cam(0, w, h, f=(f, f), p=(w / 2, h / 2)), pos=[0.0, 0.0, 0.0], rot=[0.0, 0.0, 0.0],
                   bitdepth=8, blending_exp=1.0)
        return dict(kind=0, seed=1, sources=srcs, target=tgt)
    if name == "erp4_1080p":
        sw, sh = 4096 // scale, 2048 // scale
        tw, th, f = 1920 // scale, 1080 // scale, 960.0 / scale
        srcs = []
        for (x, y) in ((0.15, 0.15), (0.15, -0.15), (-0.15, 0.15), (-0.15, -0.15)):
            srcs.append(dict(cam=R.cam(1, sw, sh), pos=[x, y, 0.0], rot=[0.0, 0.0, 0.0], bd_color=10, bd_depth=16,
                             znear=0.5, zfar=25.0))
        tgt = dict(cam=R.cam(0, tw, th, f=(f, f), p=(tw / 2, th / 2)), pos=[0.0, 0.0, 0.0], rot=[0.0, 0.0, 0.0],
                   bitdepth=10, blending_exp=1.0)
        return dict(kind=1, seed=2, sources=srcs, target=tgt)
    if name == "persp11_2048":
        # 11 perspective 2048x2048 10-bit views on a 3 x 4 grid (one slot unused),
        # 0.1 m pitch, f = 1200 px, principal point at the centre, facing +x
        sz, f = 2048 // scale, 1200.0 / scale
        srcs = []
        for r in range(3):
            for c in range(4):
                if r == 2 and c == 3:
                    continue
                srcs.append(dict(cam=R.cam(0, sz, sz, f=(f, f), p=(sz / 2, sz / 2)),
                                 pos=[0.0, (c - 1.5) * 0.1, (1 - r) * 0.1], rot=[0.0, 0.0, 0.0], bd_color=10,
                                 bd_depth=16, znear=0.5, zfar=25.0))
        tgt = dict(cam=perspective_cam(1920 // scale, 1080 // scale), pos=[0.0, 0.0, 0.0], rot=[0.0, 0.0, 0.0],
                   bitdepth=10, blending_exp=1.0)
        return dict(kind=1, seed=4, sources=srcs, target=tgt)
    if name.startswith("erp11_4096"):
        # 11 ERP 4096x2048 10-bit views with 16-bit depth on a 3 x 4 grid (one
        # slot unused, 0.1 m pitch); "erp11_4096:<seed>" picks the scene
        seed = int(name.split(":")[1]) if ":" in name else 5
        sw, sh = 4096 // scale, 2048 // scale
        srcs = []
        for r in range(3):
            for c in range(4):
                if r == 2 and c == 3:
                    continue
                srcs.append(dict(cam=R.cam(1, sw, sh), pos=[0.0, (c - 1.5) * 0.1, (1 - r) * 0.1],
                                 rot=[0.0, 0.0, 0.0], bd_color=10, bd_depth=16, znear=0.5, zfar=25.0))
        tgt = dict(cam=R.cam(1, 8192 // scale, 4096 // scale), pos=[0.0, 0.0, 0.0], rot=[0.0, 0.0, 0.0],
                   bitdepth=10, blending_exp=1.0)
        return dict(kind=1, seed=seed, sources=srcs, target=tgt)
    raise KeyError(name)


def render_sources(rg, threads=0):
    """Raw planes (y, u, v, depth) for every source of a rig."""
    out = []
    for s in rg["sources"]:
        out.append(R.render_scene(rg["kind"], rg["seed"], s["cam"], s["pos"], s["rot"], s["bd_color"],
                                  s["bd_depth"], s["znear"], s["zfar"], threads=threads))
    return out


def oracle_sources(rg, planes, ob):
    """The same sources in the oracle's description (ob = tests.oracle_bind)."""
    res = []
    for s, (y, u, v, d) in zip(rg["sources"], planes):
        c = s["cam"]
        oc = ob.make_cam(c.proj, c.w, c.h, f=tuple(c.f), p=tuple(c.p), hor=tuple(c.hor), ver=tuple(c.ver),
                         full_hor=c.full_hor)
        res.append(dict(cam=oc, pos=s["pos"], rot=s["rot"], bd_color=s["bd_color"], bd_depth=s["bd_depth"],
                        znear=s["znear"], zfar=s["zfar"], has_invalid=True, y=y, u=u, v=v, depth=d))
    return res


def oracle_cam(c, ob):
    return ob.make_cam(c.proj, c.w, c.h, f=tuple(c.f), p=tuple(c.p), hor=tuple(c.hor), ver=tuple(c.ver),
                       full_hor=c.full_hor)


def device_sources(rg, planes, device="cuda", prep=True):
    """Resident sources; prep=True also builds their culling tables for the
    rig's depth scale (fl(1/max_level) of the first source's depth bit depth)."""
    srcs = [R.DeviceSource(s["cam"], s["pos"], s["rot"], s["bd_color"], s["bd_depth"], s["znear"], s["zfar"],
                           p, device=device) for s, p in zip(rg["sources"], planes)]
    if prep:
        ds = R.scale_of(rg["sources"][0]["bd_depth"])
        for s in srcs:
            s.prepare(ds)
    return srcs


def device_target(rg, pos=None, rot=None):
    t = rg["target"]
    s0 = rg["sources"][0]
    return R.make_target(t["cam"], pos if pos is not None else t["pos"], rot if rot is not None else t["rot"],
                         t["bitdepth"], t["blending_exp"], R.scale_of(s0["bd_color"]), R.scale_of(s0["bd_depth"]))


def psnr(a, b, maxv):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mse = np.mean((a - b) ** 2)
    return float("inf") if mse == 0 else 10.0 * np.log10(maxv * maxv / mse)


def perspective_cam(w, h, hfov_deg=90.0):
    """Perspective camera of w x h pixels, principal point at the centre."""
    f = (w / 2.0) / float(np.tan(np.radians(hfov_deg) / 2.0))
    return R.cam(0, w, h, f=(f, f), p=(w / 2.0, h / 2.0))


# 1280x720, 1920x1080 and 3840x2160 perspective targets (90 degree HFOV)
BATCH_SIZES = ((1280, 720), (1920, 1080), (3840, 2160))


def random_walk(n, seed, pos_lim=0.3, yaw_lim=30.0, pitch_lim=10.0):
    """n poses (pos[3], rot[3] = yaw, pitch, roll in degrees) of a seeded smooth
    random walk clamped to +-pos_lim metres, +-yaw_lim and +-pitch_lim degrees,
    roll 0."""
    rng = np.random.default_rng(seed)
    p = np.zeros(3)
    a = np.zeros(2)
    vp = np.zeros(3)
    va = np.zeros(2)
    out = []
    for _ in range(n):
        vp = 0.8 * vp + rng.normal(0.0, 0.02, 3)
        va = 0.8 * va + rng.normal(0.0, 2.0, 2)
        p = np.clip(p + vp, -pos_lim, pos_lim)
        a = np.clip(a + va, [-yaw_lim, -pitch_lim], [yaw_lim, pitch_lim])
        out.append(([float(x) for x in np.float32(p)], [float(np.float32(a[0])), float(np.float32(a[1])), 0.0]))
    return out


def batch_targets(rg, n, seed=3, scale=1, sizes=BATCH_SIZES):
    """n targets for the sources of rig rg: poses from random_walk(seed), sizes
    cycling through `sizes` divided by `scale`."""
    t = rg["target"]
    s0 = rg["sources"][0]
    res = []
    for k, (pos, rot) in enumerate(random_walk(n, seed)):
        w, h = sizes[k % len(sizes)]
        c = perspective_cam(w // scale, h // scale)
        res.append(R.make_target(c, pos, rot, t["bitdepth"], t["blending_exp"], R.scale_of(s0["bd_color"]),
                                 R.scale_of(s0["bd_depth"])))
    return res


# Serving mix: 720p, 1080p, 1440p and 4K perspective (90 degree HFOV) and
# 2048x1024 equirectangular targets.
SERVE_SIZES = ((1280, 720, 0), (1920, 1080, 0), (2560, 1440, 0), (3840, 2160, 0), (2048, 1024, 1))


def serve_requests(rigs, n, seed=4, scale=1, sizes=SERVE_SIZES):
    """n requests (scene index, RvsTarget) over the given rigs: scene, size and
    pose drawn from a seeded generator; poses stay inside each rig's bounds
    (+-0.1 m and yaw +-20 degrees, pitch +-8 degrees around the rig centre)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        k = int(rng.integers(len(rigs)))
        rg = rigs[k]
        t = rg["target"]
        s0 = rg["sources"][0]
        w, h, proj = sizes[int(rng.integers(len(sizes)))]
        c = R.cam(1, w // scale, h // scale) if proj == 1 else perspective_cam(w // scale, h // scale)
        pos = [float(np.float32(x)) for x in rng.uniform(-0.1, 0.1, 3)]
        rot = [float(np.float32(rng.uniform(-20.0, 20.0))), float(np.float32(rng.uniform(-8.0, 8.0))), 0.0]
        out.append((k, R.make_target(c, pos, rot, t["bitdepth"], t["blending_exp"], R.scale_of(s0["bd_color"]),
                                     R.scale_of(s0["bd_depth"]))))
    return out
