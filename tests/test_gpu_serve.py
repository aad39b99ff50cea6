"""GPU: the serving loop (rvs::Server through librvs_host.so).

Requests for two resident scenes at mixed target sizes arrive while others are
on the device; each result must be bit-identical to the same view synthesized
alone.  Memory-budget admission rejects a request that cannot fit.  A
large-serving case (config 5's scene shape at reduced resolution: scenes of
11 ERP sources, ERP and perspective targets, closed loop at full concurrency)
holds to the same rule.
"""
import sys

import numpy as np
import pytest

import rvs_amd as R
import serving as S
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _yuv_alone(srcs, tgt):
    out = R.Synthesizer(srcs, tgt, want_maps=False).run()
    import torch
    torch.cuda.synchronize()
    return b"".join(out[k].cpu().numpy().tobytes() for k in ("y", "u", "v"))


def _scenes():
    rg_a = W.rig("erp4_1080p", scale=4)
    rg_b = W.rig("rects_640", scale=2)
    src_a = W.device_sources(rg_a, W.render_sources(rg_a))
    src_b = W.device_sources(rg_b, W.render_sources(rg_b))
    return rg_a, src_a, rg_b, src_b


def _requests(rg_a, rg_b, n):
    ta = W.batch_targets(rg_a, n, seed=11, scale=4)
    tb = [W.device_target(rg_b, p, r) for p, r in W.random_walk(n, seed=12, pos_lim=0.04, yaw_lim=6.0,
                                                                 pitch_lim=3.0)]
    return [(0, ta[k]) if k % 3 else (1, tb[k]) for k in range(n)]


@pytest.mark.parametrize("split", [False, True])
def test_served_results_equal_alone(torch_cuda, split):
    say("case serve open loop, 2 scenes, 18 requests")
    rg_a, src_a, rg_b, src_b = _scenes()
    reqs = _requests(rg_a, rg_b, 18)
    srv = S.Server([src_a, src_b], memory_budget=6 << 30, max_batch=8, slots=2, keep_outputs=True,
                   split_finish=split)
    try:
        st = srv.run_open(reqs, rate=2000.0, seed=4)
        say("  summary", S.summarize(st))
        assert (st["rc"] == 0).all()
        assert st["batch_views"].max() > 1          # requests were batched together
        for k, (scene, tgt) in enumerate(reqs):
            want = _yuv_alone(src_a if scene == 0 else src_b, tgt)
            assert srv.output(k) == want, "request %d (scene %d)" % (k, scene)
        # closed loop over the same requests gives the same bytes
        st2 = srv.run_closed(reqs, concurrency=6)
        assert (st2["rc"] == 0).all()
        for k, (scene, tgt) in enumerate(reqs):
            assert srv.output(k) == _yuv_alone(src_a if scene == 0 else src_b, tgt), "closed loop request %d" % k
    finally:
        srv.close()


def test_budget_rejects_request_that_cannot_fit(torch_cuda):
    say("case serve budget rejection")
    rg_a, src_a, rg_b, src_b = _scenes()
    big = W.batch_targets(rg_a, 1, seed=2, scale=1, sizes=((3840, 2160),))[0]
    small = W.device_target(rg_b)
    srv = S.Server([src_a, src_b], memory_budget=64 << 20, max_batch=4, slots=1, keep_outputs=True)
    try:
        with pytest.raises(R.RvsError):
            srv.run_closed([(1, small), (0, big), (1, small)], concurrency=3)
    finally:
        srv.close()


def test_closed_loop_keeps_its_window_after_rejections(torch_cuda):
    """Concurrency 1 with an oversized request and an odd-sized one in the
    middle: each fails with its own code, and every later request is still
    submitted and served (a rejected request must not shrink the closed loop's
    window)."""
    say("case serve closed loop at concurrency 1 with rejected requests")
    rg_a, src_a, rg_b, src_b = _scenes()
    big = W.batch_targets(rg_a, 1, seed=2, scale=1, sizes=((3840, 2160),))[0]
    odd = W.device_target(rg_b)
    odd.cam.w -= 1
    small = [W.device_target(rg_b, p, r) for p, r in W.random_walk(3, seed=13, pos_lim=0.04, yaw_lim=6.0,
                                                                   pitch_lim=3.0)]
    reqs = [(1, small[0]), (0, big), (1, small[1]), (1, odd), (1, small[2])]
    srv = S.Server([src_a, src_b], memory_budget=64 << 20, max_batch=4, slots=1, keep_outputs=True)
    try:
        with pytest.raises(R.RvsError) as ei:
            srv.run_closed(reqs, concurrency=1)
        rc = list(ei.value.stats["rc"])
        say("  rc per request", rc)
        assert rc == [0, R.RVS_E_CAPACITY, 0, R.RVS_E_UNSUPPORTED, 0]
        for k in (0, 2, 4):
            assert srv.output(k) == _yuv_alone(src_b, reqs[k][1]), "request %d" % k
        # the server stays usable
        st = srv.run_closed([reqs[0], reqs[2]], concurrency=1)
        assert (st["rc"] == 0).all()
    finally:
        srv.close()


def test_pool_admits_a_request_larger_than_a_slot_share(torch_cuda):
    """Slots share one memory arena: a request that needs most of the budget is
    admitted once the device is free, even with several slots."""
    say("case serve pooled budget, 3 slots, one large request")
    rg_a, src_a, rg_b, src_b = _scenes()
    big = W.batch_targets(rg_a, 1, seed=2, scale=2, sizes=((1920, 1080),))[0]
    small = W.device_target(rg_b)
    srv = S.Server([src_a, src_b], memory_budget=160 << 20, max_batch=4, slots=3, keep_outputs=True,
                   host_staging_bytes=8 << 20)
    try:
        st = srv.run_closed([(1, small), (0, big), (1, small)], concurrency=3)
        assert (st["rc"] == 0).all()
        assert srv.output(1) == _yuv_alone(src_a, big)
    finally:
        srv.close()


@pytest.mark.parametrize("split", [False, True])
def test_served_erp11_scenes_equal_alone(torch_cuda, split):
    say("case serve 2 scenes x 11 ERP sources, ERP + perspective targets, closed loop")
    scale = 16
    rigs = [W.rig("erp11_4096:%d" % seed, scale=scale) for seed in (5, 6)]
    srcs = [W.device_sources(rg, W.render_sources(rg)) for rg in rigs]
    rng = np.random.default_rng(9)
    reqs = []
    for k in range(12):
        sc = int(rng.integers(2))
        rg = rigs[sc]
        pos = [float(np.float32(x)) for x in rng.uniform(-0.1, 0.1, 3)]
        rot = [float(np.float32(rng.uniform(-20, 20))), float(np.float32(rng.uniform(-8, 8))), 0.0]
        # ERP 8192x4096 / 16, 4K perspective / 8 (an even height)
        cam = rg["target"]["cam"] if k % 2 == 0 else W.perspective_cam(3840 // 8, 2160 // 8)
        s0 = rg["sources"][0]
        reqs.append((sc, R.make_target(cam, pos, rot, 10, 1.0, R.scale_of(s0["bd_color"]),
                                       R.scale_of(s0["bd_depth"]))))
    srv = S.Server(srcs, memory_budget=8 << 30, max_batch=4, slots=2, keep_outputs=True, split_finish=split)
    try:
        st = srv.run_closed(reqs, concurrency=len(reqs))
        say("  summary", S.summarize(st))
        assert (st["rc"] == 0).all()
        assert st["batch_views"].max() > 1
        for k, (scene, tgt) in enumerate(reqs):
            assert srv.output(k) == _yuv_alone(srcs[scene], tgt), "request %d (scene %d)" % (k, scene)
    finally:
        srv.close()
