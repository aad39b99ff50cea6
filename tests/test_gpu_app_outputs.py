"""GPU: the configuration-driven Application's multi-frame inputs and its
mask, masked-colour, depth and masked-depth outputs, against the oracle.

Source files hold two frames (two different scenes); NumberOfFrames = 3 and
six output frames, so the input frame index runs 0 1 2 2 1 0 (the reference's
getExtendedIndex ping-pong, Pipeline.cpp:157-167) and frame 2 lies past the end
of the files, where the reader wraps to the first frame (image_loading.cpp:95-100).
Each output frame k is the oracle's view of the loaded frame at pose-trace row k.
Outputs (Pipeline.cpp:303-336, image_writing.cpp:104-139,227-302):
  mask        255 where the blended validity is not > ValidityThreshold (NaN
              included), else 0; chroma 128;
  masked      the inpainted colour with masked pixels set to 0.5, encoded;
  depth       the blended depth as the file's inverse-depth code: checked to
              decode back to the oracle's depth within one code, the nearest
              depth to the top code, holes (infinite depth) to 0;
  masked depth  the depth file's codes with masked pixels 0.
The same case run through the command-line tool gives the same bytes.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W
from test_gpu_app import _cam_json

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(tmp_path, threshold):
    rg = W.rig("rects_640", scale=2)
    frames = [W.render_sources(rg), W.render_sources(dict(rg, seed=11))]   # two scenes = two frames
    cams, tex, dep = [], [], []
    for k, s in enumerate(rg["sources"]):
        cams.append(_cam_json("v%d" % k, s["cam"], s["pos"], s["rot"], s["bd_color"], s["bd_depth"], s["znear"],
                              s["zfar"]))
        tb, db = b"", b""
        for fr in frames:
            y, u, v, d = fr[k]
            tb += y.tobytes() + u.tobytes() + v.tobytes()
            db += d.tobytes() + np.full(d.size // 2, 32768, np.uint16).tobytes()
        (tmp_path / ("v%d_tex.yuv" % k)).write_bytes(tb)
        (tmp_path / ("v%d_dep.yuv" % k)).write_bytes(db)
        tex.append(str(tmp_path / ("v%d_tex.yuv" % k)))
        dep.append(str(tmp_path / ("v%d_dep.yuv" % k)))
    t = rg["target"]
    cams.append(_cam_json("viewport", t["cam"], t["pos"], t["rot"], t["bitdepth"], 16, 1.0, 10.0))
    (tmp_path / "cams.json").write_text(json.dumps({"Version": "3.0", "cameras": cams}))
    trace = [[0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [0.01, -0.01, 0.0, 2.0, -1.0, 0.0], [-0.01, 0.02, 0.005, -3.0, 1.0, 0.0],
             [0.0, 0.0, 0.01, 1.0, 0.0, 0.0], [0.02, 0.0, -0.01, 0.0, 2.0, 0.0], [-0.02, -0.01, 0.0, 3.0, -2.0, 0.0]]
    (tmp_path / "trace.csv").write_text("X,Y,Z,Yaw,Pitch,Roll\n" +
                                        "".join(",".join(repr(float(x)) for x in row) + "\n" for row in trace))
    out = {k: str(tmp_path / (k + ".yuv")) for k in ("out", "mask", "masked", "depth", "mdepth")}
    cfg = {"Version": "3.0", "InputCameraParameterFile": str(tmp_path / "cams.json"),
           "VirtualCameraParameterFile": str(tmp_path / "cams.json"),
           "InputCameraNames": ["v0", "v1"], "VirtualCameraNames": ["viewport"],
           "ViewImageNames": tex, "DepthMapNames": dep,
           "OutputFiles": [out["out"]], "OutputMasks": [out["mask"]], "MaskedOutputFiles": [out["masked"]],
           "DepthOutputFiles": [out["depth"]], "MaskedDepthOutputFiles": [out["mdepth"]],
           "ValidityThreshold": threshold, "VirtualPoseTraceName": str(tmp_path / "trace.csv"),
           "NumberOfFrames": 3, "NumberOfOutputFrames": len(trace), "BlendingFactor": 1.0, "Precision": 1.0}
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    return rg, frames, trace, str(tmp_path / "cfg.json"), out


def _extended(k, n):
    g, r = divmod(k, n)
    return n - r - 1 if g % 2 else r


def _expected(rg, frames, trace, threshold):
    t = rg["target"]
    w, h = t["cam"].w, t["cam"].h
    res = []
    for k, row in enumerate(trace):
        fin = _extended(k, 3)
        loaded = fin if fin < len(frames) else 0          # past the end of the file: frame 0
        pos = list(np.asarray([t["pos"][c] + np.float32(row[c]) for c in range(3)], np.float32))
        r = ob.synthesize(W.oracle_sources(rg, frames[loaded], ob), W.oracle_cam(t["cam"], ob), pos, list(row[3:]),
                          t["bitdepth"], t["blending_exp"])
        mask = ~(r["validity"] > threshold)              # NaN validity is masked
        col = r["color"].copy()
        col[mask] = 0.5
        my, mu, mv = ob.encode_yuv420(col, t["bitdepth"])
        res.append(dict(r=r, mask=mask, loaded=loaded, masked=my.tobytes() + mu.tobytes() + mv.tobytes()))
    assert [e["loaded"] for e in res] == [0, 1, 0, 0, 1, 0]
    return res, w, h


def _frames(path, n, frame_bytes):
    b = open(path, "rb").read()
    assert len(b) == n * frame_bytes, (path, len(b), n * frame_bytes)
    return [b[k * frame_bytes:(k + 1) * frame_bytes] for k in range(n)]


@pytest.mark.parametrize("threshold", [5000.0, 9000.0])
def test_multiframe_outputs_match_the_oracle(tmp_path, torch_cuda, threshold):
    say("case app: 2-frame files, NumberOfFrames 3, 6 output frames, mask/masked/depth outputs, threshold %g"
        % threshold)
    rg, frames, trace, cfg, out = _case(tmp_path, threshold)
    assert R.app_run(cfg, "") == len(trace)
    exp, w, h = _expected(rg, frames, trace, threshold)
    n = len(trace)
    colour = _frames(out["out"], n, w * h * 3 // 2)
    masks = _frames(out["mask"], n, w * h * 3 // 2)
    masked = _frames(out["masked"], n, w * h * 3 // 2)
    depth = _frames(out["depth"], n, w * h * 2 * 3 // 2)
    mdepth = _frames(out["mdepth"], n, w * h * 2 * 3 // 2)
    zn, zf = 1.0, 10.0
    for k, e in enumerate(exp):
        r = e["r"]
        assert colour[k] == r["y"].tobytes() + r["u"].tobytes() + r["v"].tobytes(), "colour frame %d" % k
        m = np.frombuffer(masks[k], np.uint8)
        assert np.array_equal(m[:w * h].reshape(h, w), np.where(e["mask"], 255, 0).astype(np.uint8)), "mask %d" % k
        assert (m[w * h:] == 128).all()
        assert masked[k] == e["masked"], "masked colour frame %d" % k
        dcode = np.frombuffer(depth[k], np.uint16)
        dy = dcode[:w * h].reshape(h, w).astype(np.float64)
        assert (dcode[w * h:] == 32768).all()                      # neutral chroma of a 16-bit YUV420 depth file
        d = r["depth"].astype(np.float64)
        fin = np.isfinite(d) & (d > 0)
        inr = fin & (d >= zn) & (d <= zf)
        # decode (image_loading.cpp + process_input_frame.cu:358-381) of the written code gives the depth back
        v = dy / 65535.0
        back = zf * zn / (zn + v * (zf - zn))
        inv_step = (1 / zn - 1 / zf) / 65535.0                     # one code in inverse depth
        assert np.all(np.abs(1 / back[inr] - 1 / d[inr]) <= 1.01 * inv_step), "depth frame %d" % k
        assert np.all(dy[~fin] == 0), "holes (no depth) are written as 0"
        assert np.all(dy[fin & (d <= zn)] == 65535), "nearer than the range: the top code"
        assert np.all(dy[fin & (d > zf)] == 0), "beyond the range: code 0"
        md = np.frombuffer(mdepth[k], np.uint16)[:w * h].reshape(h, w)
        assert np.array_equal(md, np.where(e["mask"], 0, dy).astype(np.uint16)), "masked depth frame %d" % k
    # the command-line tool writes the same bytes
    before = {key: open(p, "rb").read() for key, p in out.items()}
    for p in out.values():
        os.remove(p)
    exe = os.path.join(ROOT, "rvs-cuda_amd", "lib", "rvs")
    r = subprocess.run([exe, cfg, "--cuda"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total time" in r.stdout
    for key, p in out.items():
        assert open(p, "rb").read() == before[key], "CLI output %s differs" % key


def test_precision_2_outputs(tmp_path, torch_cuda):
    """Precision 2: the Application rasterises at twice the view's size; the
    colour is resized to the view's size before encoding, the mask with
    nearest sampling (the top-left pixel of each 2x2 block) and the depth map
    with the linear rule (2x2 block means) before they are written
    (Pipeline.cpp:305-335)."""
    say("case app precision 2")
    rg, frames, trace, cfg, out = _case(tmp_path, 5000.0)
    c = json.loads(open(cfg).read())
    c.update(Precision=2.0, NumberOfFrames=1, NumberOfOutputFrames=1)
    del c["MaskedOutputFiles"]
    del c["MaskedDepthOutputFiles"]
    open(cfg, "w").write(json.dumps(c))
    assert R.app_run(cfg, "") == 1
    t = rg["target"]
    w, h = t["cam"].w, t["cam"].h
    r = ob.synthesize(W.oracle_sources(rg, frames[0], ob), W.oracle_cam(t["cam"], ob), list(t["pos"]), [0.0] * 3,
                      t["bitdepth"], t["blending_exp"], rescale=2.0)
    assert r["validity"].shape == (2 * h, 2 * w)
    assert open(out["out"], "rb").read() == r["y"].tobytes() + r["u"].tobytes() + r["v"].tobytes()
    m = np.frombuffer(open(out["mask"], "rb").read(), np.uint8)[:w * h].reshape(h, w)
    rmask = ~(r["validity"] > 5000.0)
    assert np.array_equal(m, np.where(rmask[0::2, 0::2], 255, 0).astype(np.uint8))
    d2 = ob.resize_color(np.repeat(r["depth"][:, :, None], 3, axis=2), w, h)[:, :, 0].astype(np.float64)
    dy = np.frombuffer(open(out["depth"], "rb").read(), np.uint16)[:w * h].reshape(h, w).astype(np.float64)
    zn, zf = 1.0, 10.0
    inr = np.isfinite(d2) & (d2 >= zn) & (d2 <= zf)
    back = zf * zn / (zn + dy / 65535.0 * (zf - zn))
    assert np.all(np.abs(1 / back[inr] - 1 / d2[inr]) <= 1.01 * (1 / zn - 1 / zf) / 65535.0)
    assert np.all(dy[~np.isfinite(d2)] == 0)

