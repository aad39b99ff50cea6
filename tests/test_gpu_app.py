"""GPU: the configuration-file driven Application (drop-in for the reference's
RVS executable) on raw YUV files, fused path and view-by-view plugin path,
against the oracle frame by frame.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W

pytestmark = pytest.mark.gpu


def _cam_json(name, c, pos, rot, bd, bdd, zn, zf):
    d = {"Name": name, "Position": list(map(float, pos)), "Rotation": list(map(float, rot)), "Depthmap": 1,
         "Background": 0, "Depth_range": [zn, zf], "Resolution": [c.w, c.h], "BitDepthColor": bd,
         "BitDepthDepth": bdd, "ColorSpace": "YUV420", "DepthColorSpace": "YUV420", "HasInvalidDepth": True}
    if c.proj == 1:
        d.update(Projection="Equirectangular", Hor_range=[c.hor[0], c.hor[1]], Ver_range=[c.ver[0], c.ver[1]])
    else:
        d.update(Projection="Perspective", Focal=[c.f[0], c.f[1]], Principle_point=[c.p[0], c.p[1]])
    return d


def _write_case(tmp_path, rg, planes, trace, extra=None):
    cams, tex, dep = [], [], []
    for k, (s, (y, u, v, d)) in enumerate(zip(rg["sources"], planes)):
        cams.append(_cam_json("v%d" % k, s["cam"], s["pos"], s["rot"], s["bd_color"], s["bd_depth"], s["znear"],
                              s["zfar"]))
        (tmp_path / ("v%d_tex.yuv" % k)).write_bytes(y.tobytes() + u.tobytes() + v.tobytes())
        chroma = np.full(d.size // 2, 32768, np.uint16)   # YUV420 depth file: Y plane + neutral chroma
        (tmp_path / ("v%d_dep.yuv" % k)).write_bytes(d.tobytes() + chroma.tobytes())
        tex.append("v%d_tex.yuv" % k)
        dep.append("v%d_dep.yuv" % k)
    t = rg["target"]
    cams.append(_cam_json("viewport", t["cam"], t["pos"], t["rot"], t["bitdepth"], 16, 1.0, 10.0))
    (tmp_path / "cams.json").write_text(json.dumps({"Version": "3.0", "cameras": cams}))
    (tmp_path / "trace.csv").write_text("X,Y,Z,Yaw,Pitch,Roll\n" +
                                        "".join(",".join(repr(float(x)) for x in row) + "\n" for row in trace))
    cfg = {"Version": "3.0", "InputCameraParameterFile": str(tmp_path / "cams.json"),
           "VirtualCameraParameterFile": str(tmp_path / "cams.json"),
           "InputCameraNames": ["v%d" % k for k in range(len(planes))], "VirtualCameraNames": ["viewport"],
           "ViewImageNames": [str(tmp_path / x) for x in tex], "DepthMapNames": [str(tmp_path / x) for x in dep],
           "OutputFiles": [str(tmp_path / "out.yuv")], "OutputMasks": [str(tmp_path / "mask.yuv")],
           "DepthOutputFiles": [str(tmp_path / "depth.yuv")],
           "VirtualPoseTraceName": str(tmp_path / "trace.csv"), "NumberOfFrames": 1,
           "NumberOfOutputFrames": len(trace), "BlendingFactor": t["blending_exp"], "Precision": 1.0}
    cfg.update(extra or {})
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    return str(tmp_path / "cfg.json")


def _oracle_frames(rg, planes, trace):
    t = rg["target"]
    out = b""
    for row in trace:
        pos = [t["pos"][k] + np.float32(row[k]) for k in range(3)]
        pos = list(np.asarray(pos, np.float32))
        rot = list(row[3:])
        r = ob.synthesize(W.oracle_sources(rg, planes, ob), W.oracle_cam(t["cam"], ob), pos, rot, t["bitdepth"],
                          t["blending_exp"])
        out += r["y"].tobytes() + r["u"].tobytes() + r["v"].tobytes()
    return out


@pytest.mark.parametrize("stagewise", [False, True])
def test_application_matches_oracle(tmp_path, torch_cuda, stagewise):
    print("case application stagewise=%s" % stagewise, flush=True)
    rg = W.rig("rects_640", scale=2)
    planes = W.render_sources(rg)
    trace = [[0.0, 0.0, 0.0, 0.0, 0.0, 0.0], [0.01, -0.02, 0.005, 3.0, -1.5, 0.5], [-0.02, 0.01, 0.0, -2.0, 1.0, 0.0]]
    cfg = _write_case(tmp_path, rg, planes, trace)
    n = R.app_run(cfg, "", stagewise=stagewise)
    assert n == len(trace)
    got = (tmp_path / "out.yuv").read_bytes()
    want = _oracle_frames(rg, planes, trace)
    assert len(got) == len(want)
    assert got == want
    w, h = rg["target"]["cam"].w, rg["target"]["cam"].h
    assert (tmp_path / "mask.yuv").stat().st_size == len(trace) * w * h * 3 // 2
    assert (tmp_path / "depth.yuv").stat().st_size == len(trace) * w * h * 3


def test_application_erp_sources(tmp_path, torch_cuda):
    print("case application erp", flush=True)
    rg = W.rig("erp4_1080p", scale=4)
    planes = W.render_sources(rg)
    trace = [[0.0, 0.0, 0.0, 10.0, 0.0, 0.0], [0.05, 0.05, 0.0, -20.0, 5.0, 0.0]]
    cfg = _write_case(tmp_path, rg, planes, trace)
    assert R.app_run(cfg, "") == 2
    assert (tmp_path / "out.yuv").read_bytes() == _oracle_frames(rg, planes, trace)


def test_application_reports_missing_file(tmp_path, torch_cuda):
    rg = W.rig("rects_640", scale=4)
    planes = W.render_sources(rg)
    cfg = _write_case(tmp_path, rg, planes, [[0, 0, 0, 0, 0, 0]])
    os.remove(tmp_path / "v0_tex.yuv")
    with pytest.raises(RuntimeError, match="Failed to read raw YUV color file"):
        R.app_run(cfg, "")
