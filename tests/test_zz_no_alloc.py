"""GPU: rvs_synthesize_noalloc (workspace = the library's static device
scratch, no allocation) against rvs_synthesize with a caller workspace, bit
for bit on every map and YUV plane, on two calls in a row, the second with a
smaller view from smaller sources; and a view too large for the scratch is
refused with RVS_E_CAPACITY."""
import sys

import numpy as np
import pytest

import rvs_amd as R
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _outputs(torch, t):
    w, h = t.cam.w, t.cam.h
    dt = torch.uint8 if t.bitdepth <= 8 else torch.int16
    f = dict(dtype=torch.float32, device="cuda")
    return {"color": torch.empty((h, w, 3), **f), "quality": torch.empty((h, w), **f),
            "depth": torch.empty((h, w), **f), "validity": torch.empty((h, w), **f),
            "y": torch.empty((h, w), dtype=dt, device="cuda"),
            "u": torch.empty((h // 2, w // 2), dtype=dt, device="cuda"),
            "v": torch.empty((h // 2, w // 2), dtype=dt, device="cuda")}


def _same(a, b):
    a = a.cpu().numpy()
    b = b.cpu().numpy()
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    return np.array_equal(a, b)


def test_noalloc_equals_synthesize_on_two_calls(torch_cuda):
    torch = torch_cuda
    cases = [("erp4_1080p", 2, [0.05, -0.03, 0.01], [12.0, -4.0, 0.0]),   # 960x540 from 4 ERP 2048x1024
             ("rects_640", 2, [0.02, 0.01, -0.01], [3.0, 1.0, 0.0])]      # then 320x240 from 2 x 320x240
    for name, scale, pos, rot in cases:
        say("case noalloc %s / %d" % (name, scale))
        rg = W.rig(name, scale=scale)
        srcs = W.device_sources(rg, W.render_sources(rg))
        tgt = W.device_target(rg, pos, rot)
        ref = R.Synthesizer(srcs, tgt).run()
        got = _outputs(torch, tgt)
        assert R.synthesize_noalloc(srcs, tgt, got) == 0
        torch.cuda.synchronize()
        for k in ("color", "quality", "depth", "validity", "y", "u", "v"):
            assert _same(got[k], ref[k]), "%s: %s differs from rvs_synthesize" % (name, k)


def test_noalloc_refuses_a_view_beyond_the_scratch(torch_cuda):
    torch = torch_cuda
    say("case noalloc refusal")
    rg = W.rig("erp4_1080p", scale=4)
    srcs = W.device_sources(rg, W.render_sources(rg))
    big = W.batch_targets(rg, 1, seed=2, sizes=((7680, 4320),))[0]
    out = {"color": torch.empty((16, 16, 3), dtype=torch.float32, device="cuda")}   # never written
    with pytest.raises(R.RvsError) as ei:
        R.synthesize_noalloc(srcs, big, out)
    assert ei.value.code == R.RVS_E_CAPACITY
    assert "static scratch" in str(ei.value)
