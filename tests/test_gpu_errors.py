"""GPU: error paths of the C ABI.

Bad arguments are refused with RVS_E_INVALID / RVS_E_UNSUPPORTED and a
message (the reference throws std::runtime_error at the same places, e.g.
odd sizes for YUV420 and a Precision that is not a positive factor), an entry list that is too
small reports the exact capacity it needed, and a failed call leaves the
library usable: the next valid call gives the same bits as before.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import rvs_amd as R
import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scene():
    rg = W.rig("rects_640", scale=4)
    return rg, W.device_sources(rg, W.render_sources(rg))


def _yuv(out):
    return b"".join(out[k].cpu().numpy().tobytes() for k in ("y", "u", "v"))


def _target(rg, w, h, rescale=1.0):
    t = rg["target"]
    s0 = rg["sources"][0]
    c = R.cam(0, w, h, f=tuple(t["cam"].f), p=(w / 2, h / 2))
    return R.make_target(c, t["pos"], t["rot"], t["bitdepth"], t["blending_exp"], R.scale_of(s0["bd_color"]),
                         R.scale_of(s0["bd_depth"]), rescale=rescale)


def test_bad_arguments_are_refused_and_leave_the_library_usable(torch_cuda):
    torch = torch_cuda
    say("case error paths")
    rg, srcs = _scene()
    good = W.device_target(rg)
    ref = _yuv(R.Synthesizer(srcs, good, want_maps=False).run())
    torch.cuda.synchronize()

    # odd target size: YUV420 output cannot be encoded
    with pytest.raises(R.RvsError) as e:
        R.Synthesizer(srcs, _target(rg, 161, 120), want_maps=False).run()
    assert "odd" in str(e.value)

    # a Precision that is not a positive factor
    for bad in (0.0, -1.0, float("nan")):
        with pytest.raises(R.RvsError) as e:
            R.Synthesizer(srcs, _target(rg, 160, 120, rescale=bad), want_maps=False)
        assert e.value.code == R.RVS_E_INVALID and "Precision" in str(e.value)

    # an empty batch
    with pytest.raises(R.RvsError):
        R.BatchSynthesizer(srcs, [])

    # a workspace smaller than the plan
    syn = R.Synthesizer(srcs, good, want_maps=False)
    vo = R.RvsViewOut()
    for k in ("color", "y", "u", "v"):
        setattr(vo, k, syn.out[k].data_ptr())
    needed = C.c_uint32(0)
    rc = R.hip().rvs_synthesize(syn.arr, len(srcs), C.byref(syn.target), syn.ws.data_ptr(), 16, 0,
                                C.byref(vo), C.byref(needed), None)
    assert rc == R.RVS_E_INVALID
    assert b"workspace" in R.hip().rvs_last_error()

    # a source preparation buffer that is too small
    s0 = srcs[0].struct()
    buf = torch.empty(64, dtype=torch.uint8, device="cuda")
    rc = R.hip().rvs_source_prep(C.byref(s0), R.scale_of(rg["sources"][0]["bd_depth"]), buf.data_ptr(), 64, None)
    assert rc == R.RVS_E_INVALID

    # the library still gives the same bits
    again = _yuv(R.Synthesizer(srcs, good, want_maps=False).run())
    torch.cuda.synchronize()
    assert again == ref


def test_too_small_entry_list_reports_the_needed_capacity(torch_cuda):
    torch = torch_cuda
    say("case entry capacity")
    rg, srcs = _scene()
    tgt = W.device_target(rg)
    syn = R.Synthesizer(srcs, tgt, want_maps=False, entry_cap=256)
    vo = R.RvsViewOut()
    for k in ("color", "y", "u", "v"):
        setattr(vo, k, syn.out[k].data_ptr())
    needed = C.c_uint32(0)
    rc = R.hip().rvs_synthesize(syn.arr, len(srcs), C.byref(syn.target), syn.ws.data_ptr(), syn.ws_bytes, 256,
                                C.byref(vo), C.byref(needed), None)
    assert rc == R.RVS_E_CAPACITY and needed.value > 256
    # exactly the needed capacity suffices, and gives the default run's bits
    exact = R.Synthesizer(srcs, tgt, want_maps=False, entry_cap=needed.value)
    rc = exact.run()
    torch.cuda.synchronize()
    assert _yuv(rc) == _yuv(R.Synthesizer(srcs, tgt, want_maps=False).run())
    torch.cuda.synchronize()
