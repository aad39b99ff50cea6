"""GPU: multispectral blending (BlendingMethod "Multispectral",
BlendedView.cpp:62-88; split_frequencies / calcBlurring, blending.cpp:62-159,
371-394) against the oracle.

The device sums each blur window exactly in fixed point (2^-40 units) where the
reference sums a double integral image, so the split is bit-exact whenever
the double sums are exact (inputs on a 2^-6 grid) and otherwise may differ
where a window sum lies within rounding distance of an integer (ColorSpace
YUV truncates it) or of a float midpoint (ColorSpace RGB rounds it): the
tolerance is at most 0.1 % of the pixels, each by one step of the window's
count (YUV) or two float ulps (RGB).  Whole frames: depth, quality and validity
(the low band's, independent of colour) bit-exact; colour within the parity
tolerance (1 code on >= 99.5 % of the samples, PSNR >= 48 dB per plane).
"""
import json
import sys

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W
from test_gpu_app import _write_case
from test_gpu_configs import within_tolerance

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _maps(rng, h, w, grid):
    col = rng.random((h, w, 3)).astype(np.float32)
    if grid:
        col = (np.floor(col * 64) / 64).astype(np.float32)
    depth = rng.uniform(1, 8, (h, w)).astype(np.float32)
    depth[rng.random((h, w)) < 0.1] = np.inf                   # uncovered pixels: quality 0
    shape = np.where(rng.random((h, w)) < 0.9, rng.uniform(0.5, 4000, (h, w)), 0).astype(np.float32)
    return col, depth, shape


def _window_counts(mask, rad):
    h, w = mask.shape
    I = np.zeros((h + 1, w + 1), np.int64)
    I[1:, 1:] = np.cumsum(np.cumsum(mask.astype(np.int64), 0), 1)
    ys = np.arange(h)
    xs = np.arange(w)
    y0, y1 = np.maximum(0, ys - rad), np.minimum(h, ys + rad + 1)
    x0, x1 = np.maximum(0, xs - rad), np.minimum(w, xs + rad + 1)
    return I[y1][:, x1] - I[y0][:, x1] - I[y1][:, x0] + I[y0][:, x0]


def _split_gpu(torch, col, depth, shape, rgb):
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in dict(c=col, d=depth, s=shape).items()}
    low = torch.empty_like(t["c"])
    high = torch.empty_like(t["c"])
    R.split_frequencies(t["c"], t["d"], t["s"], low, high, rgb=rgb)
    torch.cuda.synchronize()
    return low.cpu().numpy(), high.cpu().numpy()


@pytest.mark.parametrize("rgb", [False, True])
@pytest.mark.parametrize("h,w", [(61, 97), (360, 640), (1, 300)])
def test_split_exact_on_grid_values(torch_cuda, rgb, h, w):
    say("case split grid %dx%d rgb=%s" % (w, h, rgb))
    col, depth, shape = _maps(np.random.default_rng(h * 7 + w), h, w, grid=True)
    low, high = _split_gpu(torch_cuda, col, depth, shape, rgb)
    rl, rh = ob.split_frequencies(col, depth, shape, rgb=rgb)
    assert np.array_equal(low.view(np.uint32), rl.view(np.uint32))
    assert np.array_equal(high.view(np.uint32), rh.view(np.uint32))


@pytest.mark.parametrize("rgb", [False, True])
def test_split_within_rounding_on_real_values(torch_cuda, rgb):
    say("case split real values rgb=%s" % rgb)
    h, w = 270, 480
    col, depth, shape = _maps(np.random.default_rng(3), h, w, grid=False)
    low, high = _split_gpu(torch_cuda, col, depth, shape, rgb)
    rl, rh = ob.split_frequencies(col, depth, shape, rgb=rgb)
    diff = low != rl
    assert diff.mean() <= 1e-3, "%d of %d samples differ" % (diff.sum(), diff.size)
    if rgb:
        tol = 2 * np.spacing(np.abs(rl)) + 1e-30
        assert np.all(np.abs(low - rl) <= tol)
    else:
        assert not diff[..., 1:].any()                         # U, V stay in the low band unchanged
        mask = (shape / depth > 0)
        cnt = _window_counts(mask, max(h, w) // 20)
        d = np.abs(low[..., 0].astype(np.float64) - rl[..., 0])
        bad = diff[..., 0]
        assert np.all(np.abs(d[bad] * cnt[bad] - 1.0) < 1e-3), "a difference other than one truncation step"
    assert np.array_equal((col - low).view(np.uint32), high.view(np.uint32))


def test_add_color(torch_cuda):
    torch = torch_cuda
    a = torch.rand((33, 17, 3), device="cuda")
    b = torch.rand((33, 17, 3), device="cuda")
    out = torch.empty_like(a)
    R.add_color(a, b, out)
    torch.cuda.synchronize()
    assert torch.equal(out, a + b)


def _oracle(rg, planes, pos, rot, exps, rgb):
    t = rg["target"]
    return ob.synthesize(W.oracle_sources(rg, planes, ob), W.oracle_cam(t["cam"], ob), pos, rot, t["bitdepth"], 0.0,
                         multispectral=exps, rgb=rgb)


def _check_frame(out, ref, t, tag):
    for k in ("quality", "depth", "validity"):
        a = np.ascontiguousarray(out[k], np.float32)
        assert np.array_equal(a.view(np.uint32), ref[k].view(np.uint32)) or \
            np.array_equal(np.nan_to_num(a, nan=-7.0), np.nan_to_num(ref[k], nan=-7.0)), "%s: %s" % (tag, k)
    same = float(np.mean(np.ascontiguousarray(out["color"], np.float32) == ref["color"]))
    say("  %s: colour samples bit-exact %.5f" % (tag, same))
    maxv = (1 << t["bitdepth"]) - 1
    for k in ("y", "u", "v"):
        within_tolerance(out[k], ref[k], maxv, "%s %s" % (tag, k))


@pytest.mark.parametrize("rgb", [False, True])
def test_stagewise_multispectral_matches_oracle(torch_cuda, rgb):
    torch = torch_cuda
    say("case stagewise multispectral rgb=%s" % rgb)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    pos, rot = [0.02, -0.01, 0.01], [3.0, -1.0, 0.5]
    tgt = W.device_target(rg, pos, rot)
    exps = (1.0, 4.0)
    sw = R.StagewiseSynthesizer(W.device_sources(rg, planes), tgt, multispectral=exps, rgb=rgb).run()
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in sw.items()}
    ref = _oracle(rg, planes, list(tgt.pos), list(tgt.rot), exps, rgb)
    _check_frame(out, ref, rg["target"], "stagewise rgb=%s" % rgb)


@pytest.mark.parametrize("rgb", [False, True])
def test_application_multispectral(tmp_path, torch_cuda, rgb):
    say("case application multispectral rgb=%s" % rgb)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    trace = [[0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [0.01, -0.02, 0.005, 3.0, -1.5, 0.5]]
    cfg = _write_case(tmp_path, rg, planes, trace,
                      extra={"BlendingMethod": "Multispectral", "BlendingLowFreqFactor": 2.0,
                             "BlendingHighFreqFactor": 6.0, "ColorSpace": "RGB" if rgb else "YUV"})
    assert R.app_run(cfg, "") == len(trace)
    t = rg["target"]
    w, h = t["cam"].w, t["cam"].h
    got = (tmp_path / "out.yuv").read_bytes()
    dt = np.uint16 if t["bitdepth"] > 8 else np.uint8
    fb = w * h * 3 // 2 * np.dtype(dt).itemsize
    assert len(got) == len(trace) * fb
    for k, row in enumerate(trace):
        pos = list(np.asarray([t["pos"][c] + np.float32(row[c]) for c in range(3)], np.float32))
        ref = _oracle(rg, planes, pos, list(row[3:]), (2.0, 6.0), rgb)
        buf = np.frombuffer(got[k * fb:(k + 1) * fb], dt)
        y = buf[:w * h].reshape(h, w)
        u = buf[w * h:w * h * 5 // 4].reshape(h // 2, w // 2)
        v = buf[w * h * 5 // 4:].reshape(h // 2, w // 2)
        maxv = (1 << t["bitdepth"]) - 1
        for name, a in (("y", y), ("u", u), ("v", v)):
            within_tolerance(a, ref[name], maxv, "app frame %d %s" % (k, name))
    cfg_d = json.loads(open(cfg).read())
    assert cfg_d["BlendingMethod"] == "Multispectral"
