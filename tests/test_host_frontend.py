"""CPU: the reference-format front end (JSON reader, camera parameters,
configuration, pose traces) and the C ABI library exports.

The JSON reader is compared with the reference's own JsonParser.cpp, compiled
from the reference tree into oracle/_ref by oracle/Makefile (skipped where that
tree is absent).  Expected values of the reference's unit tests
(rvs/src/unit_test.cpp:310-388) are restated here; the reference's Config/
files are fixtures under tests/golden/reference_config.
"""
import ctypes as C
import json
import os
import re
import shutil

import numpy as np
import pytest

import oracle_bind as ob
import rvs_amd as R
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_config")


# ------------------------------------------------------------------- JSON

UNIT_DOC = """
{
	"Content_name": "ClassroomVideo",
	"BoundingBox_center" : [0.0, 42.0, -1e3],
	"Fps" : 30,
	"cameras" :
	[
		{
			"Name": "v0",
			"Background" : 0
		},
		{
		}
	]
}"""


def test_json_reference_unit_case():            # unit_test.cpp:310-342
    q = R.json_query
    assert q(UNIT_DOC, "Content_name", "string") == (0, "ClassroomVideo")
    assert q(UNIT_DOC, "BoundingBox_center", "size") == (0, "3")
    assert [q(UNIT_DOC, "BoundingBox_center/%d" % i, "double")[1] for i in range(3)] == ["0", "42", "-1000"]
    assert q(UNIT_DOC, "Fps", "double") == (0, "30")
    assert q(UNIT_DOC, "cameras", "size") == (0, "2")
    assert q(UNIT_DOC, "cameras/0/Name", "string") == (0, "v0")
    assert q(UNIT_DOC, "cameras/0/Background", "double") == (0, "0")
    assert q(UNIT_DOC, "cameras/1", "size") == (0, "0")


def test_json_overrides():                      # unit_test.cpp:344-368
    doc = '{"Alpha": "RVS", "Beta" : [1.0, 2.0, 3.0], "Gamma" : 30}'
    assert R.json_query(doc, "Gamma", "int") == (0, "30")
    assert R.json_query(doc, "Gamma", "int", overrides='{"Gamma" : 40}') == (0, "40")
    assert R.json_query(doc, "Alpha", "string", overrides='{"Gamma" : 40}') == (0, "RVS")


QUERIES = [
    ('{"a": 1.5}', "a", "int"),            # not an integer
    ('{"a": 2.5}', "a", "int"),            # lround -> 3, error -0.5 is accepted by the reference rule
    ('{"a": -2.5}', "a", "int"),
    ('{"a": 1e3}', "a", "int"),
    ('{"a": false}', "a", "bool"),
    ('{"a": false}', "?a", "truthy"),
    ('{"a": true}', "a", "bool"),
    ('{"a": null}', "?a", "truthy"),
    ('{"a": "x\\"y"}', "a", "string"),     # escapes are rejected
    ('{"a": [1, 2,]}', "a", "size"),
    ('{"a": 1} x', "a", "int"),            # stray character
    ('{"a": 1, "a": 2}', "a", "int"),      # the last duplicate wins
    ('{"a": [1,2]}', "a/2", "int"),        # out of range
    ('{"a": {"b": 1}}', "a/b", "double"),
    ('{"a": {"b": 1}}', "b", "double"),    # missing required key
    ('{"a": -}', "a", "double"),
    ('[1, 2, 3]', "1", "double"),
    ('  {  }  ', "", "size"),
    ('{"a": tru}', "a", "bool"),
    ('{"a": 0.1}', "a", "float"),
    ('{"a": 123456789.123456789}', "a", "double"),
    ('{"a": "s"}', "a", "double"),         # type error
    ('{"a": 3}', "a", "type"),
]


@pytest.mark.parametrize("doc,path,conv", QUERIES)
def test_json_matches_reference_parser(doc, path, conv):
    if ob.ref_json() is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    mine = R.json_query(doc, path, conv)
    ref = ob.ref_json_query(doc, path, conv)
    assert mine[0] == ref[0], (mine, ref)
    if ref[0] == 0:
        assert mine[1] == ref[1]


def test_json_random_documents_match_reference_parser():
    if ob.ref_json() is None:
        pytest.skip("reference tree absent: oracle/_ref not built")
    rng = np.random.default_rng(0)

    def gen(depth):
        k = rng.integers(0, 6 if depth < 3 else 3)
        if k == 0:
            return repr(float(rng.normal() * 10.0 ** int(rng.integers(-3, 6))))
        if k == 1:
            return '"%s"' % "".join(rng.choice(list("abc xyz_01"), rng.integers(0, 6)))
        if k == 2:
            return ["true", "false", "null"][rng.integers(0, 3)]
        if k == 3:
            return "[" + " , ".join(gen(depth + 1) for _ in range(rng.integers(0, 4))) + "]"
        keys = ["k%d" % i for i in range(rng.integers(0, 4))]
        return "{" + ",\n ".join('"%s": %s' % (kk, gen(depth + 1)) for kk in keys) + "}"

    for _ in range(200):
        doc = '{"root": %s, "n": %s}' % (gen(0), int(rng.integers(-5, 5)))
        for path, conv in (("root", "type"), ("n", "int"), ("root", "size"), ("root/k0", "type"),
                           ("root/0", "type"), ("?root", "truthy")):
            assert R.json_query(doc, path, conv) == ob.ref_json_query(doc, path, conv), (doc, path, conv)


# ------------------------------------------------------------- pose trace

POSE_DOC = """X,Y,Z,Yaw,Pitch,Roll
0, 0, 0, 7.744e-06, 1.10991e-05, 2.98821e-06
-0.000289001, -4.49e-05, -2.50004e-05, -1337, 0.0172351, 0.00275806
-0.00062, -8.20999e-05, 1.0, 13.1, 42, 0.00589616
-0.047839, 0.0798379, -0.030835, -68.9187, -3.20608, 16.5

"""


def test_pose_trace_reference_unit_case():       # unit_test.cpp:370-388
    t = R.pose_trace_parse(POSE_DOC)
    assert t.shape == (4, 6)
    assert t[0, 0] == 0.0
    assert abs(t[1, 1] + 4.49e-05) < 1e-12
    assert t[2, 2] == 1.0
    assert t[1, 3] == -1337.0
    assert t[2, 4] == 42.0
    assert t[3, 5] == np.float32(16.5)


@pytest.mark.parametrize("bad", ["X,Y,Z,Yaw,Pitch\n1,2,3,4,5\n", "X,Y,Z,Yaw,Pitch,Roll\n1,2,3,4,5\n",
                                 "X,Y,Z,Yaw,Pitch,Roll\n1,2,3,4,5,6\n\n1,2,3,4,5,6\n",
                                 "X,Y,Z,Yaw,Pitch,Roll\na,2,3,4,5,6\n"])
def test_pose_trace_rejects(bad):
    with pytest.raises(ValueError):
        R.pose_trace_parse(bad)


def test_pose_trace_crlf_and_spaces():
    t = R.pose_trace_parse(" X , Y,Z,Yaw,Pitch,Roll \r\n1,2,3,4,5,6\r\n  \r\n")
    assert t.tolist() == [[1, 2, 3, 4, 5, 6]]


def test_reference_pose_trace_file():
    with open(os.path.join(GOLDEN, "PoseTrace.csv")) as f:
        t = R.pose_trace_parse(f.read())
    assert t.shape == (10, 6) and (t == 0).all()


# ----------------------------------------------------------------- config

@pytest.fixture
def ref_config_dir(tmp_path):
    cfg = tmp_path / "Config"
    cfg.mkdir()
    for f in os.listdir(GOLDEN):
        shutil.copy(os.path.join(GOLDEN, f), cfg / f)
    work = tmp_path / "work"
    work.mkdir()
    old = os.getcwd()
    os.chdir(work)      # the reference configuration names "../Config/..." relative to the working directory
    yield tmp_path
    os.chdir(old)


def test_reference_configuration_loads(ref_config_dir):
    d = R.config_describe("../Config/RVS-A01.json")
    assert d["version"] == "3.0" and d["inputs"] == 4 and d["virtual"] == 1
    assert d["blending_factor"] == 1.0 and d["precision"] == 1.0
    assert d["number_of_frames"] == 10 and d["number_of_output_frames"] == 10 and d["pose_trace"] == 10
    inputs, viewport = d["cameras"][:4], d["cameras"][4]
    for c in inputs:
        assert c["proj"] == "Equirectangular" and c["w"] == 4096 and c["h"] == 2048 and c["full_hor"] == 1
        assert c["bd"] == 10 and c["bdd"] == 16
    assert viewport["proj"] == "Perspective" and viewport["w"] == 2048 and viewport["far"] == 1000.0


def _write_config(tmp_path, cams, extra=None, camfile_version="3.0"):
    cam = {"Version": camfile_version, "cameras": cams}
    (tmp_path / "cams.json").write_text(json.dumps(cam))
    cfg = {"Version": "3.0", "InputCameraParameterFile": str(tmp_path / "cams.json"),
           "VirtualCameraParameterFile": str(tmp_path / "cams.json"), "InputCameraNames": ["a"],
           "VirtualCameraNames": ["b"], "ViewImageNames": ["a.yuv"], "DepthMapNames": ["ad.yuv"]}
    cfg.update(extra or {})
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    return str(tmp_path / "cfg.json")


def _cam(name, **kw):
    c = {"Name": name, "Projection": "Perspective", "Position": [0, 0, 0], "Rotation": [0, 0, 0], "Depthmap": 1,
         "Background": 0, "Depth_range": [1, 10], "Resolution": [64, 48], "Focal": [50, 50],
         "Principle_point": [32, 24], "BitDepthColor": 8, "BitDepthDepth": 16, "ColorSpace": "YUV420",
         "DepthColorSpace": "YUV420"}
    c.update(kw)
    return c


def test_config_defaults(tmp_path):
    d = R.config_describe(_write_config(tmp_path, [_cam("a"), _cam("b")]))
    assert d["blending_factor"] == 5.0 and d["validity_threshold"] == 5000.0 and d["precision"] == 1.0
    assert d["number_of_frames"] == 1 and d["cameras"][0]["invalid"] == 1


@pytest.mark.parametrize("cams,extra,msg", [
    ([_cam("a", Depth_range=[10, 1]), _cam("b")], None, "Inverted depth range"),
    ([_cam("a", Depth_range=[1, 1001]), _cam("b")], None, "far > 1000"),
    ([_cam("a"), _cam("a")], None, "duplicate camera"),
    ([_cam("a")], None, "does not have camera 'b'"),
    ([_cam("a", ColorSpace="YUV444"), _cam("b")], None, "YUV420 color space"),
    ([_cam("a", Depthmap=2), _cam("b")], None, "Depthmap 1"),
    ([_cam("a"), _cam("b")], {"Version": "4.0"}, "does not match the RVS version"),
    ([_cam("a"), _cam("b")], {"DepthMapNames": []}, "Length of DepthMapNames"),
    ([_cam("a"), _cam("b")], {"OutputFiles": ["x", "y"]}, "Length of OutputFiles"),
    ([_cam("a", HasInvalidDepth=False), _cam("b")], {"ColorSpace": "XYZ"}, "Unknown color space"),
])
def test_config_rejects(tmp_path, cams, extra, msg):
    with pytest.raises(ValueError, match=msg):
        R.config_describe(_write_config(tmp_path, cams, extra))


def test_config_pose_trace_too_short(tmp_path):
    (tmp_path / "pt.csv").write_text("X,Y,Z,Yaw,Pitch,Roll\n0,0,0,0,0,0\n")
    with pytest.raises(ValueError, match="pose trace"):
        R.config_describe(_write_config(tmp_path, [_cam("a"), _cam("b")],
                                        {"VirtualPoseTraceName": str(tmp_path / "pt.csv"), "NumberOfFrames": 2}))


def test_full_hor_range_needs_integer_bounds(tmp_path):
    erp = dict(Projection="Equirectangular", Hor_range=[-180, 180], Ver_range=[-90, 90])
    d = R.config_describe(_write_config(tmp_path, [_cam("a", **erp), _cam("b")]))
    assert d["cameras"][0]["full_hor"] == 1
    erp["Hor_range"] = [-179.5, 180]
    d = R.config_describe(_write_config(tmp_path, [_cam("a", **erp), _cam("b")]))
    assert d["cameras"][0]["full_hor"] == 0


# --------------------------------------------------------------- C ABI

def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rvs_[a-z0-9_]+)\s*\(", text)) - {"rvs_stream_t"})


@pytest.mark.parametrize("header,lib", [("rvs_hip.h", R.HIP_LIB), ("rvs_host.h", R.HOST_LIB)])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) > 5
    L = C.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_functions_of_the_abi_match_the_oracle():
    """Pure host entry points (no GPU needed): relative pose and ERP tables."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        pf, pt = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        rf, rt = rng.uniform(-180, 180, 3), rng.uniform(-60, 60, 3)
        Ra, ta = R.relative_pose(pf, rf, pt, rt)
        Ro, to = ob.relative_pose(pf, rf, pt, rt)
        assert np.array_equal(Ra.view(np.uint32), Ro.view(np.uint32)) and np.array_equal(ta, to)
    cam = R.cam(1, 64, 32)
    tab = R.erp_table(cam)
    # unprojecting unit depth with the oracle gives the same products of the table
    ocam = ob.make_cam("erp", 64, 32)
    xyz = ob.unproject(ocam, ob.generate_image_pos(ocam), np.ones((32, 64), np.float32))
    cp, sp, ct, st = tab[:64], tab[64:128], tab[128:160], tab[160:]
    assert np.array_equal(xyz[..., 0], (ct[:, None] * cp[None, :]).astype(np.float32))
    assert np.array_equal(xyz[..., 1], (ct[:, None] * sp[None, :]).astype(np.float32))
    assert np.array_equal(xyz[..., 2], np.broadcast_to(st[:, None], (32, 64)))


def test_workspace_planning_limits():
    """Workspace planning is host code (no GPU): a valid batch has a size, and
    views with a side above 65536 (tile coordinates are 16-bit) or batches
    beyond 2^31 pixels are refused with a size of 0 and a message."""
    rg = W.rig("erp4_1080p")
    srcs = []
    for s in rg["sources"]:
        st = R.RvsSource()
        st.cam = s["cam"]
        srcs.append(st)
    arr = (R.RvsSource * len(srcs))(*srcs)
    ok = W.batch_targets(rg, 2, seed=4, sizes=((1920, 1080),))
    assert R.hip().rvs_synth_batch_workspace_size(arr, len(srcs), (R.RvsTarget * 2)(*ok), 2, 0) > 0
    wide = R.make_target(R.cam(0, 70000, 64), [0.0] * 3, [0.0] * 3, 10, 1.0, R.scale_of(10), R.scale_of(16))
    assert R.hip().rvs_synth_batch_workspace_size(arr, len(srcs), (R.RvsTarget * 1)(wide), 1, 0) == 0
    assert b"side above" in R.hip().rvs_last_error()
    big = R.make_target(R.cam(0, 65536, 40000), [0.0] * 3, [0.0] * 3, 10, 1.0, R.scale_of(10), R.scale_of(16))
    assert R.hip().rvs_synth_batch_workspace_size(arr, len(srcs), (R.RvsTarget * 1)(big), 1, 0) == 0
    assert b"too large" in R.hip().rvs_last_error()


def test_serving_autotune_choice():
    """The serving autotuner's choice: within the p99 bound, rates within 2 %
    of the best count as equal and the lowest p99 wins; beyond the bound only
    if nothing meets it (then the lowest p99)."""
    import autotune as A
    t = lambda b, s, v, p: {"max_batch": b, "slots": s, "views_s": v, "p99_ms": p}
    trials = [t(16, 2, 3071.0, 12.9), t(32, 2, 3101.0, 25.1), t(32, 3, 3110.0, 49.0), t(64, 2, 3300.0, 60.0)]
    assert A.pick_serving(trials)["max_batch"] == 16            # 3071 >= 0.98 * 3110, lowest p99
    assert A.pick_serving(trials, tie=0.0)["slots"] == 3        # strict: the best rate within the bound
    assert A.pick_serving(trials, p99_bound_ms=20.0)["max_batch"] == 16
    assert A.pick_serving([t(8, 1, 100.0, 80.0), t(4, 1, 90.0, 70.0)])["max_batch"] == 4
