"""GPU: batched synthesis (many target views of uneven size from the same
resident sources in one launch set), frustum culling of the binning work, and
HIP-graph capture of a batch.

The bar: every view of a batch is bit-identical (all maps and YUV planes) to
the same view synthesized alone, and to the CPU oracle; culled poses (targets
looking away from most of the source data) match the oracle exactly; graph
replays reproduce the eager result on every replay.
"""
import sys

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W

pytestmark = pytest.mark.gpu

MAPS = ("color", "y", "u", "v")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _alone(srcs, tgt):
    syn = R.Synthesizer(srcs, tgt, want_maps=False)
    return {k: v.clone() for k, v in syn.run().items() if v is not None}


def _oracle(rg, planes, tgt):
    t = rg["target"]
    c = tgt.cam
    oc = ob.make_cam(c.proj, c.w, c.h, f=tuple(c.f), p=tuple(c.p), hor=tuple(c.hor), ver=tuple(c.ver),
                     full_hor=c.full_hor)
    return ob.synthesize(W.oracle_sources(rg, planes, ob), oc, list(tgt.pos), list(tgt.rot), tgt.bitdepth,
                         t["blending_exp"])


def _check_oracle(out, ref, bitdepth, tag):
    npdt = np.uint8 if bitdepth <= 8 else np.uint16
    for k in ("y", "u", "v"):
        a = out[k].cpu().numpy().view(npdt)
        assert np.array_equal(a, ref[k]), "%s: plane %s differs in %d samples" % (tag, k, int((a != ref[k]).sum()))


@pytest.mark.parametrize("tile", [16, 32])
def test_batch_equals_alone_mixed_sizes(torch_cuda, tile):
    """The batch at either raster tile size equals each view alone at the
    default size (16) and the oracle."""
    torch = torch_cuda
    say("case batch erp4/4 mixed sizes x6, tile %d" % tile)
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgts = W.batch_targets(rg, 6, seed=3, scale=4)
    bs = R.BatchSynthesizer(srcs, tgts, want_maps=True, tile=tile)
    outs = bs.run()
    torch.cuda.synchronize()
    for k, t in enumerate(tgts):
        say("  view %d %dx%d" % (k, t.cam.w, t.cam.h))
        syn = R.Synthesizer(srcs, t, want_maps=True)
        alone = syn.run()
        torch.cuda.synchronize()
        for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
            assert torch.equal(outs[k][key], alone[key]), "view %d map %s" % (k, key)
    for k in (0, 1):
        _check_oracle(outs[k], _oracle(rg, planes, tgts[k]), tgts[k].bitdepth, "batch view %d" % k)


@pytest.mark.parametrize("tile", [16, 32])
def test_batch_rects_many_views(torch_cuda, tile):
    torch = torch_cuda
    say("case batch rects/2 x12, tile %d" % tile)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgts = []
    for k, (pos, rot) in enumerate(W.random_walk(12, seed=7, pos_lim=0.05, yaw_lim=8.0, pitch_lim=4.0)):
        tgts.append(W.device_target(rg, pos, rot))
    bs = R.BatchSynthesizer(srcs, tgts, tile=tile)
    outs = bs.run()
    torch.cuda.synchronize()
    for k in range(len(tgts)):
        alone = _alone(srcs, tgts[k])
        for key in MAPS:
            assert torch.equal(outs[k][key], alone[key]), "view %d map %s" % (k, key)
    for k in (0, 5, 11):
        _check_oracle(outs[k], _oracle(rg, planes, tgts[k]), tgts[k].bitdepth, "rects view %d" % k)


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("bd", [8, 10])
def test_planes_only_batch_equals_full(torch_cuda, bd, tile):
    """Views that want only the encoded planes: tiles without holes or
    deferred pixels are encoded by the raster pass itself and the colour map
    (in the workspace) is written only on their borders; the planes equal those
    of the same batch with full colour maps and the oracle's, at 8 and 10 bits,
    with holes (the rects scene's uncovered border) and deferred pixels."""
    torch = torch_cuda
    say("case planes-only batch, %d bit, tile %d" % (bd, tile))
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgts = []
    for pos, rot in W.random_walk(6, seed=12, pos_lim=0.05, yaw_lim=8.0, pitch_lim=4.0):
        t = W.device_target(rg, pos, rot)
        t.bitdepth = bd
        tgts.append(t)
    full = R.BatchSynthesizer(srcs, tgts).run()
    full = [{k: o[k].clone() for k in ("y", "u", "v")} for o in full]
    part = R.BatchSynthesizer(srcs, tgts, want_color=False, tile=tile).run()
    torch.cuda.synchronize()
    for k in range(len(tgts)):
        assert part[k]["color"] is None
        for key in ("y", "u", "v"):
            assert torch.equal(part[k][key], full[k][key]), "view %d plane %s" % (k, key)
    for k in (0, 3):
        _check_oracle(part[k], _oracle(rg, planes, tgts[k]), bd, "planes-only view %d" % k)


@pytest.mark.parametrize("tile", [16, 32])
def test_internal_maps_ignore_stale_workspace(torch_cuda, tile):
    """Internal maps (quality, colour of planes-only views) are stored only
    where the later passes read them: a workspace full of NaN (every unread
    quality a hole) gives the same colour and planes as a batch that outputs
    every map."""
    torch = torch_cuda
    say("case stale workspace, tile %d" % tile)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgts = [W.device_target(rg, pos, rot)
            for pos, rot in W.random_walk(4, seed=21, pos_lim=0.05, yaw_lim=8.0, pitch_lim=4.0)]
    full = [{k: v.clone() for k, v in o.items() if v is not None}
            for o in R.BatchSynthesizer(srcs, tgts, want_maps=True, tile=tile).run()]
    for want_color in (True, False):
        bs = R.BatchSynthesizer(srcs, tgts, want_color=want_color, tile=tile)
        bs.ws.fill_(0xFF)
        outs = bs.run()
        torch.cuda.synchronize()
        for k in range(len(tgts)):
            for key in (MAPS if want_color else ("y", "u", "v")):
                assert torch.equal(outs[k][key], full[k][key]), "colour %s view %d map %s" % (want_color, k, key)


@pytest.mark.parametrize("rot", [[90.0, 0.0, 0.0], [180.0, 0.0, 0.0], [-60.0, 40.0, 0.0], [30.0, -80.0, 10.0]])
def test_culled_poses_match_oracle_erp(torch_cuda, rot):
    """Targets that see little of each ERP source: binning work items are culled."""
    torch = torch_cuda
    say("case culled erp rot=%s" % rot)
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes, prep=rot[0] != 180.0)   # one case builds its tables per call
    tgt = W.device_target(rg, [0.1, -0.05, 0.02], rot)
    R.profile_enable(True)
    try:
        out = R.Synthesizer(srcs, tgt, want_maps=False).run()
        torch.cuda.synchronize()
        cnt = R.profile_counters()
    finally:
        R.profile_enable(False)
    items = 4 * ((1024 - 1 + 63) // 64) * ((512 - 1 + 3) // 4)
    say("  work items kept %d of %d" % (cnt["bin_work_items"], items))
    assert cnt["bin_work_items"] < items
    _check_oracle(out, _oracle(rg, planes, tgt), tgt.bitdepth, "culled %s" % rot)


@pytest.mark.parametrize("yaw", [25.0, 70.0, 180.0])
def test_culled_poses_match_oracle_perspective(torch_cuda, yaw):
    torch = torch_cuda
    say("case culled rects yaw=%s" % yaw)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgt = W.device_target(rg, [0.0, 0.0, 0.0], [yaw, 3.0, 0.0])
    out = R.Synthesizer(srcs, tgt, want_maps=False).run()
    torch.cuda.synchronize()
    _check_oracle(out, _oracle(rg, planes, tgt), tgt.bitdepth, "rects yaw %s" % yaw)


@pytest.mark.parametrize("tile", [16, 32])
def test_batch_graph_replay(torch_cuda, tile):
    """A batch captured into a HIP graph replays to the eager result, replay
    after replay, with unrelated eager work in between."""
    torch = torch_cuda
    say("case graph replay, tile %d" % tile)
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgts = W.batch_targets(rg, 4, seed=5, scale=4)
    bs = R.BatchSynthesizer(srcs, tgts, tile=tile)
    eager = [{k: v.clone() for k, v in o.items() if v is not None} for o in bs.run()]
    torch.cuda.synchronize()
    for o in bs.outs:
        for k in MAPS:
            o[k].zero_()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bs.enqueue()   # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bs.enqueue()
    for rep in range(3):
        for o in bs.outs:
            for k in MAPS:
                o[k].zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()   # unrelated eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert bs.check() == 0
        for v, o in enumerate(bs.outs):
            for k in MAPS:
                assert torch.equal(o[k], eager[v][k]), "replay %d view %d map %s" % (rep, v, k)
        del junk


@pytest.mark.parametrize("name,scale", [("erp4_1080p", 4), ("rects_640", 2)])
def test_stagewise_equals_fused(torch_cuda, name, scale):
    """The per-stage call structure (reproject, raster, blend per source, then
    inpaint, encode) gives the fused path's bits."""
    torch = torch_cuda
    say("case stagewise vs fused %s/%d" % (name, scale))
    rg = W.rig(name, scale=scale)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgt = W.device_target(rg, [0.02, -0.01, 0.0], [4.0, -2.0, 0.0])
    fused = R.Synthesizer(srcs, tgt, want_maps=True).run()
    torch.cuda.synchronize()
    fused = {k: v.clone() for k, v in fused.items() if v is not None}
    stage = R.StagewiseSynthesizer(srcs, tgt).run()
    torch.cuda.synchronize()
    for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
        assert torch.equal(stage[key], fused[key]), key


def _planes_within(out, ref, bitdepth, tag, frac=0.995, db=48.0):
    """North-star colour tolerance: within 1 code on >= 99.5% of samples and
    >= 48 dB PSNR, per plane."""
    npdt = np.uint8 if bitdepth <= 8 else np.uint16
    maxv = (1 << bitdepth) - 1
    for k in ("y", "u", "v"):
        a = out[k].cpu().numpy().view(npdt).astype(np.int64)
        b = ref[k].astype(np.int64)
        close = float(np.mean(np.abs(a - b) <= 1))
        p = W.psnr(a, b, maxv)
        say("  %s plane %s: within 1 code %.5f, PSNR %.1f dB" % (tag, k, close, p))
        assert close >= frac and p >= db, "%s plane %s" % (tag, k)


def test_erp_target_within_tolerance(torch_cuda):
    """ERP targets project with the device atan2f/asinf (the host oracle uses
    libm): colour planes within the stated tolerance, and a batch mixing ERP
    and perspective targets equals each view synthesized alone."""
    torch = torch_cuda
    say("case erp target")
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    erp = R.make_target(R.cam(1, 512, 256), [0.05, 0.02, 0.0], [20.0, 0.0, 0.0], 10, 1.0,
                        R.scale_of(10), R.scale_of(16))
    persp = W.batch_targets(rg, 2, seed=9, scale=4)
    outs = R.BatchSynthesizer(srcs, [erp] + persp).run()
    torch.cuda.synchronize()
    ref = _oracle(rg, planes, erp)
    _planes_within(outs[0], ref, 10, "erp 512x256")
    for k, t in enumerate([erp] + persp):
        alone = _alone(srcs, t)
        for key in MAPS:
            assert torch.equal(outs[k][key], alone[key]), "view %d map %s" % (k, key)


def test_perspective_grid_sources(torch_cuda):
    """Eleven perspective sources (the serving rig's second scene) -> one view, exact."""
    torch = torch_cuda
    say("case persp11 / 4")
    rg = W.rig("persp11_2048", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    tgt = W.device_target(rg, [0.01, 0.03, -0.02], [5.0, 2.0, 0.0])
    out = R.Synthesizer(srcs, tgt, want_maps=False).run()
    torch.cuda.synchronize()
    _check_oracle(out, _oracle(rg, planes, tgt), tgt.bitdepth, "persp11")


@pytest.mark.parametrize("tile", [16, 32])
@pytest.mark.parametrize("size", [(64, 36), (32, 18)])
def test_dense_sources_compact_bins(torch_cuda, tile, size):
    """Sources far denser than the target (eleven 512x512 views into 64x36 or
    32x18): thousands of quads per (tile, source) bin, at 32x18 one bin alone
    beyond the raster's LDS cache of a phase (1536 quads per 16x16 tile area),
    exact against the oracle; the entry list grows on demand from a
    deliberately small capacity."""
    torch = torch_cuda
    say("case dense bins, tile %d, %dx%d" % (tile, size[0], size[1]))
    rg = W.rig("persp11_2048", scale=4)
    planes = W.render_sources(rg)
    srcs = W.device_sources(rg, planes)
    t = rg["target"]
    tgt = R.make_target(W.perspective_cam(*size), [0.01, 0.02, 0.0], [3.0, 1.0, 0.0], t["bitdepth"],
                        t["blending_exp"], R.scale_of(10), R.scale_of(16))
    R.profile_enable(True)
    try:
        syn = R.Synthesizer(srcs, tgt, entry_cap=4096, want_maps=False, tile=tile)
        out = syn.run()
        torch.cuda.synchronize()
        cnt = R.profile_counters()
    finally:
        R.profile_enable(False)
    say("  largest bin %d, entries %d, capacity grown to %d" % (cnt["max_bin"], cnt["bin_entries"], syn.entry_cap))
    assert cnt["max_bin"] > 1024 and syn.entry_cap >= cnt["bin_entries"]
    if size == (32, 18):
        assert cnt["max_bin"] > 1536 * (tile // 16) ** 2
    _check_oracle(out, _oracle(rg, planes, tgt), tgt.bitdepth, "dense")


@pytest.mark.parametrize("tile", [16, 32])
def test_binning_beyond_the_kept_range(torch_cuda, monkeypatch, tile):
    """Work items beyond the kept range (here all but one: RVS_KEEP_ITEMS=1)
    are counted and filled by the re-evaluating binning passes, wide tile
    ranges included: every map equals the default run's (all items kept,
    16-pixel tiles)."""
    torch = torch_cuda
    say("case binning beyond the kept range, tile %d" % tile)
    cases = []
    rg = W.rig("erp4_1080p", scale=4)
    srcs = W.device_sources(rg, W.render_sources(rg))
    erp = R.make_target(R.cam(1, 512, 256), [0.05, 0.02, 0.0], [20.0, 0.0, 0.0], 10, 1.0,
                        R.scale_of(10), R.scale_of(16))
    cases.append(("erp4/4", srcs, [erp] + W.batch_targets(rg, 3, seed=3, scale=4)))
    rg_b = W.rig("persp11_2048", scale=4)
    srcs_b = W.device_sources(rg_b, W.render_sources(rg_b))
    cases.append(("persp11/4", srcs_b, [W.device_target(rg_b, [0.01, 0.03, -0.02], [5.0, 2.0, 0.0])]))
    for tag, s, tg in cases:
        say("  %s, %d views" % (tag, len(tg)))
        monkeypatch.delenv("RVS_KEEP_ITEMS", raising=False)
        ref = [{k: v.clone() for k, v in o.items() if v is not None}
               for o in R.BatchSynthesizer(s, tg, want_maps=True).run()]
        torch.cuda.synchronize()
        monkeypatch.setenv("RVS_KEEP_ITEMS", "1")
        R.profile_enable(True)
        try:
            outs = R.BatchSynthesizer(s, tg, want_maps=True, tile=tile).run()
            torch.cuda.synchronize()
            cnt = R.profile_counters()
        finally:
            R.profile_enable(False)
            monkeypatch.delenv("RVS_KEEP_ITEMS", raising=False)
        assert cnt["bin_work_items"] > 1
        for k in range(len(tg)):
            for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
                assert torch.equal(outs[k][key], ref[k][key]), "%s view %d map %s" % (tag, k, key)


@pytest.mark.parametrize("tile", [16, 32])
def test_replay_split_equals_in_order(torch_cuda, monkeypatch, tile):
    """Deferred tiles replay sources in parallel (a block per tile and source,
    then a blend) up to a workspace limit and one block per tile beyond it:
    RVS_REPLAY_TILES=0 sends every tile the in-order way, =1 all but the first;
    every map equals the default run's."""
    torch = torch_cuda
    say("case replay split, tile %d" % tile)
    cases = []
    rg = W.rig("erp4_1080p", scale=4)
    srcs = W.device_sources(rg, W.render_sources(rg))
    cases.append(("erp4/4", srcs, W.batch_targets(rg, 4, seed=5, scale=4)))
    rg_b = W.rig("persp11_2048", scale=4)
    srcs_b = W.device_sources(rg_b, W.render_sources(rg_b))
    cases.append(("persp11/2", srcs_b, [W.device_target(rg_b, [0.01, 0.03, -0.02], [5.0, 2.0, 0.0]),
                                         W.device_target(rg_b, [-0.02, 0.0, 0.03], [-4.0, 1.0, 0.0])]))
    for tag, s, tg in cases:
        monkeypatch.delenv("RVS_REPLAY_TILES", raising=False)
        R.profile_enable(True)
        try:
            ref = [{k: v.clone() for k, v in o.items() if v is not None}
                   for o in R.BatchSynthesizer(s, tg, want_maps=True, tile=tile).run()]
            torch.cuda.synchronize()
            cnt = R.profile_counters()
        finally:
            R.profile_enable(False)
        say("  %s, %d views, %d deferred tiles" % (tag, len(tg), cnt["replay_tiles"]))
        assert cnt["replay_tiles"] > 1
        for lim in ("0", "1"):
            monkeypatch.setenv("RVS_REPLAY_TILES", lim)
            try:
                outs = R.BatchSynthesizer(s, tg, want_maps=True, tile=tile).run()
                torch.cuda.synchronize()
            finally:
                monkeypatch.delenv("RVS_REPLAY_TILES", raising=False)
            for k in range(len(tg)):
                for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
                    assert torch.equal(outs[k][key], ref[k][key]), "%s limit %s view %d map %s" % (tag, lim, k, key)


def test_without_kept_geometry(torch_cuda, monkeypatch):
    """RVS_KEPT_GEOM=0: kept work items store only their tile ranges (no
    vertex windows or validities), and the raster, the fill pass of wide
    ranges and the replay evaluate vertices themselves: every map equals the
    default run's."""
    torch = torch_cuda
    say("case without kept geometry")
    rg = W.rig("persp11_2048", scale=4)
    s = W.device_sources(rg, W.render_sources(rg))
    tg = [W.device_target(rg, [0.01, 0.03, -0.02], [5.0, 2.0, 0.0]),
          R.make_target(R.cam(1, 512, 256), [0.05, 0.02, 0.0], [20.0, 0.0, 0.0], 10, 1.0,
                        R.scale_of(10), R.scale_of(16))]
    monkeypatch.delenv("RVS_KEPT_GEOM", raising=False)
    ref = [{k: v.clone() for k, v in o.items() if v is not None}
           for o in R.BatchSynthesizer(s, tg, want_maps=True).run()]
    torch.cuda.synchronize()
    monkeypatch.setenv("RVS_KEPT_GEOM", "0")
    try:
        outs = R.BatchSynthesizer(s, tg, want_maps=True).run()
        torch.cuda.synchronize()
    finally:
        monkeypatch.delenv("RVS_KEPT_GEOM", raising=False)
    for k in range(len(tg)):
        for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
            assert torch.equal(outs[k][key], ref[k][key]), "view %d map %s" % (k, key)
