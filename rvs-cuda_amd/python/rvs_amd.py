"""Python binding of the MI355X view-synthesis path (ctypes over the C ABI of
include/rvs_hip.h and include/rvs_host.h).

Device buffers are torch tensors on the GPU (torch is plumbing here: memory and
streams); every computation happens in librvs_hip.so.  Loading fails loudly when
the built libraries are missing: there is no fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIB_DIR, "librvs_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "librvs_host.so")

RVS_OK = 0
RVS_E_INVALID = -1
RVS_E_HIP = -2
RVS_E_CAPACITY = -3
RVS_E_UNSUPPORTED = -4


class RvsCam(C.Structure):
    _fields_ = [("proj", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("f", C.c_float * 2), ("p", C.c_float * 2),
                ("hor", C.c_float * 2), ("ver", C.c_float * 2), ("full_hor", C.c_int)]


class RvsXform(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("t", C.c_float * 3)]


class RvsSource(C.Structure):
    _fields_ = [("cam", RvsCam), ("bitdepth_color", C.c_int), ("bitdepth_depth", C.c_int),
                ("znear", C.c_float), ("zfar", C.c_float), ("has_invalid", C.c_int),
                ("pos", C.c_float * 3), ("rot", C.c_float * 3),
                ("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("depth", C.c_void_p),
                ("erp_table", C.c_void_p), ("prep", C.c_void_p), ("prep_depth_scale", C.c_float)]


class RvsTarget(C.Structure):
    _fields_ = [("cam", RvsCam), ("pos", C.c_float * 3), ("rot", C.c_float * 3),
                ("bitdepth", C.c_int), ("blending_exp", C.c_float), ("rescale", C.c_float),
                ("color_scale", C.c_float), ("depth_scale", C.c_float), ("tile", C.c_int)]


class RvsViewOut(C.Structure):
    _fields_ = [("color", C.c_void_p), ("quality", C.c_void_p), ("depth", C.c_void_p),
                ("validity", C.c_void_p), ("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p)]


class RvsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rvs error %d: %s" % (code, msg))
        self.code = code


_hip = None
_host = None


def _need(path):
    if not os.path.exists(path):
        raise RuntimeError("%s is not built; run __graft_entry__.build() (make -C rvs-cuda_amd/csrc)" % path)


def hip():
    """librvs_hip.so (loaded once)."""
    global _hip
    if _hip is None:
        _need(HIP_LIB)
        L = C.CDLL(HIP_LIB)
        vp, i, f, sz, u32 = C.c_void_p, C.c_int, C.c_float, C.c_size_t, C.c_uint32
        P = C.POINTER
        sig = {
            "rvs_last_error": ([], C.c_char_p),
            "rvs_version": ([], C.c_char_p),
            "rvs_relative_pose": ([vp, vp, vp, vp, P(RvsXform)], i),
            "rvs_erp_table_len": ([P(RvsCam)], sz),
            "rvs_erp_table_fill": ([P(RvsCam), vp], i),
            "rvs_source_prep_size": ([P(RvsCam)], sz),
            "rvs_source_prep": ([P(RvsSource), f, vp, sz, vp], i),
            "rvs_decode_yuv420": ([vp, vp, vp, i, i, i, f, vp, vp], i),
            "rvs_decode_depth": ([vp, i, i, i, f, f, f, i, vp, vp], i),
            "rvs_reproject": ([P(RvsCam), P(RvsCam), P(RvsXform), vp, vp, f, vp, vp, vp], i),
            "rvs_raster_workspace_size": ([i, i, i, i, u32], sz),
            "rvs_raster": ([vp, vp, vp, i, i, i, i, i, f, vp, sz, u32, vp, vp, vp, P(u32), vp], i),
            "rvs_blend_step": ([vp, vp, vp, vp, vp, vp, vp, vp, i, i, f, i, vp], i),
            "rvs_inpaint_workspace_size": ([i, i], sz),
            "rvs_inpaint": ([vp, vp, i, i, vp, sz, vp, vp], i),
            "rvs_encode_yuv420": ([vp, i, i, i, vp, vp, vp, vp], i),
            "rvs_resize_color_workspace_size": ([i, i, i, i], sz),
            "rvs_resize_color": ([vp, i, i, vp, i, i, vp, sz, vp], i),
            "rvs_split_frequencies_workspace_size": ([i, i, i], sz),
            "rvs_split_frequencies": ([vp, vp, vp, i, i, i, vp, sz, vp, vp, vp], i),
            "rvs_add_color": ([vp, vp, vp, i, i, vp], i),
            "rvs_synth_workspace_size": ([P(RvsSource), i, P(RvsTarget), u32], sz),
            "rvs_synthesize": ([P(RvsSource), i, P(RvsTarget), vp, sz, u32, P(RvsViewOut), P(u32), vp], i),
            "rvs_synth_batch_workspace_size": ([P(RvsSource), i, P(RvsTarget), i, u32], sz),
            "rvs_synthesize_batch": ([P(RvsSource), i, P(RvsTarget), i, vp, sz, u32, P(RvsViewOut), P(u32), vp],
                                     i),
            "rvs_synthesize_batch_async": ([P(RvsSource), i, P(RvsTarget), i, vp, sz, u32, P(RvsViewOut), vp], i),
            "rvs_synth_check": ([P(RvsSource), i, P(RvsTarget), i, vp, u32, P(u32), vp], i),
            "rvs_synthesize_noalloc": ([P(RvsSource), i, P(RvsTarget), u32, P(RvsViewOut), P(u32), vp], i),
            "rvs_profile_enable": ([i], i),
            "rvs_profile_read": ([vp], i),
            "rvs_profile_counters": ([vp], i),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _hip = L
    return _hip


def host():
    """librvs_host.so (loaded once)."""
    global _host
    if _host is None:
        _need(HOST_LIB)
        L = C.CDLL(HOST_LIB)
        L.rvs_scene_render.argtypes = [C.c_int, C.c_uint64, C.POINTER(RvsCam), C.c_void_p, C.c_void_p,
                                       C.c_int, C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int]
        L.rvs_scene_render.restype = C.c_int
        L.rvs_host_last_error.restype = C.c_char_p
        L.rvs_json_query.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.rvs_json_query.restype = C.c_int
        L.rvs_pose_trace_parse.argtypes = [C.c_char_p, C.c_void_p, C.c_int]
        L.rvs_pose_trace_parse.restype = C.c_int
        L.rvs_config_describe.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.rvs_config_describe.restype = C.c_int
        L.rvs_app_run.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.rvs_app_run.restype = C.c_int
        _host = L
    return _host


def check(rc):
    if rc != RVS_OK:
        raise RvsError(rc, hip().rvs_last_error().decode())
    return rc


# ------------------------------------------------------------------ helpers

def cam(proj, w, h, f=(0.0, 0.0), p=(0.0, 0.0), hor=(-180.0, 180.0), ver=(-90.0, 90.0), full_hor=None):
    c = RvsCam()
    c.proj = 1 if proj in (1, "erp", "Equirectangular") else 0
    c.w, c.h = int(w), int(h)
    c.f[0], c.f[1] = f
    c.p[0], c.p[1] = p
    c.hor[0], c.hor[1] = hor
    c.ver[0], c.ver[1] = ver
    if full_hor is None:
        full_hor = c.proj == 1 and float(hor[0]) == -180.0 and float(hor[1]) == 180.0
    c.full_hor = int(bool(full_hor))
    return c


def plane_dtype(bitdepth):
    return np.uint8 if bitdepth <= 8 else np.uint16


def max_level(bitdepth):
    return (1 << bitdepth) - 1


def scale_of(bitdepth):
    """fl(1.0f / max_level(bitdepth)) as image_loading.cpp:134 computes it."""
    return float(np.float32(1.0) / np.float32(max_level(bitdepth)))


def relative_pose(pos_from, rot_from, pos_to, rot_to):
    x = RvsXform()
    a = [np.ascontiguousarray(v, np.float32) for v in (pos_from, rot_from, pos_to, rot_to)]
    check(hip().rvs_relative_pose(*(v.ctypes.data for v in a), C.byref(x)))
    return np.array(x.R, np.float32).reshape(3, 3), np.array(x.t, np.float32)


def erp_table(c):
    n = hip().rvs_erp_table_len(C.byref(c))
    out = np.zeros(n, np.float32)
    check(hip().rvs_erp_table_fill(C.byref(c), out.ctypes.data))
    return out


def render_scene(kind, seed, c, pos, rot, bd_color, bd_depth, znear, zfar, threads=0):
    """Synthetic source view: returns raw planes (y, u, v, depth) as numpy arrays."""
    dt_c, dt_d = plane_dtype(bd_color), plane_dtype(bd_depth)
    y = np.zeros((c.h, c.w), dt_c)
    u = np.zeros((c.h // 2, c.w // 2), dt_c)
    v = np.zeros((c.h // 2, c.w // 2), dt_c)
    d = np.zeros((c.h, c.w), dt_d)
    p = np.ascontiguousarray(pos, np.float32)
    r = np.ascontiguousarray(rot, np.float32)
    rc = host().rvs_scene_render(kind, seed, C.byref(c), p.ctypes.data, r.ctypes.data, bd_color, bd_depth,
                                 znear, zfar, y.ctypes.data, u.ctypes.data, v.ctypes.data, d.ctypes.data,
                                 threads)
    if rc != 0:
        raise RuntimeError("rvs_scene_render failed")
    return y, u, v, d


def json_query(text, path, conv, overrides=None):
    buf = C.create_string_buffer(4096)
    rc = host().rvs_json_query(text.encode(), (overrides or "").encode(), path.encode(), conv.encode(), buf, 4096)
    return rc, buf.value.decode()


def pose_trace_parse(text, max_rows=4096):
    out = np.zeros((max_rows, 6), np.float32)
    n = host().rvs_pose_trace_parse(text.encode(), out.ctypes.data, max_rows)
    if n < 0:
        raise ValueError(host().rvs_host_last_error().decode())
    return out[:n]


def config_describe(path):
    import json
    buf = C.create_string_buffer(1 << 16)
    rc = host().rvs_config_describe(path.encode(), buf, 1 << 16)
    if rc != 0:
        raise ValueError(buf.value.decode())
    return json.loads(buf.value.decode())


def app_run(config_path, source_path="", stagewise=False):
    n = host().rvs_app_run(config_path.encode(), source_path.encode(), int(stagewise))
    if n < 0:
        raise RuntimeError(host().rvs_host_last_error().decode())
    return n


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


# ---------------------------------------------------------- stage wrappers
# (all tensors are torch CUDA tensors, contiguous)

def decode_yuv420(y, u, v, bitdepth, color_scale, out, stream=None):
    h, w = y.shape
    check(hip().rvs_decode_yuv420(_ptr(y), _ptr(u), _ptr(v), bitdepth, w, h, color_scale, _ptr(out),
                                  _stream(stream)))


def decode_depth(raw, bitdepth, depth_scale, znear, zfar, has_invalid, out, stream=None):
    h, w = raw.shape
    check(hip().rvs_decode_depth(_ptr(raw), bitdepth, w, h, depth_scale, znear, zfar, int(has_invalid),
                                 _ptr(out), _stream(stream)))


def reproject(src_cam, dst_cam, R, t, depth, erp_tab, rescale, uv, vdepth, stream=None):
    x = RvsXform()
    for k in range(9):
        x.R[k] = float(np.asarray(R, np.float32).ravel()[k])
    for k in range(3):
        x.t[k] = float(np.asarray(t, np.float32)[k])
    check(hip().rvs_reproject(C.byref(src_cam), C.byref(dst_cam), C.byref(x), _ptr(depth), _ptr(erp_tab),
                              rescale, _ptr(uv), _ptr(vdepth), _stream(stream)))


def raster(color, depth, uv, out_w, out_h, wrap, rescale, out_color, out_depth, out_shape, workspace=None,
           entry_cap=0, stream=None):
    import torch
    in_h, in_w = depth.shape
    while True:
        need = hip().rvs_raster_workspace_size(in_w, in_h, out_w, out_h, entry_cap)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=depth.device)
        needed = C.c_uint32(0)
        rc = hip().rvs_raster(_ptr(color), _ptr(depth), _ptr(uv), in_w, in_h, out_w, out_h, int(wrap), rescale,
                              _ptr(workspace), need, entry_cap, _ptr(out_color), _ptr(out_depth), _ptr(out_shape),
                              C.byref(needed), _stream(stream))
        if rc == RVS_E_CAPACITY and needed.value > 0:
            entry_cap = grow_cap(needed.value)
            continue
        check(rc)
        return workspace


def blend_step(acc, color, depth, shape, exp, first, stream=None):
    h, w = depth.shape
    check(hip().rvs_blend_step(_ptr(acc["color"]), _ptr(acc["quality"]), _ptr(acc["depth"]),
                               _ptr(acc["validity"]), _ptr(acc["mask"]), _ptr(color), _ptr(depth), _ptr(shape),
                               w, h, exp, int(first), _stream(stream)))


def inpaint(color, quality, map_out=None, workspace=None, stream=None):
    import torch
    h, w = quality.shape
    need = hip().rvs_inpaint_workspace_size(w, h)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=quality.device)
    check(hip().rvs_inpaint(_ptr(color), _ptr(quality), w, h, _ptr(workspace), need, _ptr(map_out),
                            _stream(stream)))
    return workspace


def split_frequencies(color, depth, shape, low, high, rgb=False, workspace=None, stream=None):
    """Multispectral blending's frequency split of one synthesized view
    (BlendedView.cpp:70-80): low / high bands into `low`, `high` (h x w x 3)."""
    import torch
    h, w = depth.shape
    need = hip().rvs_split_frequencies_workspace_size(w, h, int(bool(rgb)))
    if need == 0:
        raise RvsError(RVS_E_INVALID, "split_frequencies: bad image size %dx%d" % (w, h))
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=depth.device)
    check(hip().rvs_split_frequencies(_ptr(color), _ptr(depth), _ptr(shape), w, h, int(bool(rgb)), _ptr(workspace),
                                      need, _ptr(low), _ptr(high), _stream(stream)))
    return workspace


def add_color(a, b, out, stream=None):
    h, w = a.shape[:2]
    check(hip().rvs_add_color(_ptr(a), _ptr(b), _ptr(out), w, h, _stream(stream)))


def encode_yuv420(color, bitdepth, y, u, v, stream=None):
    h, w = y.shape
    check(hip().rvs_encode_yuv420(_ptr(color), w, h, bitdepth, _ptr(y), _ptr(u), _ptr(v), _stream(stream)))


PROF_STAGES = ("bin", "raster", "replay", "inpaint", "encode")


def profile_enable(on=True):
    check(hip().rvs_profile_enable(int(on)))


def profile_read():
    ms = np.zeros(len(PROF_STAGES), np.float32)
    check(hip().rvs_profile_read(ms.ctypes.data))
    return dict(zip(PROF_STAGES, (float(x) for x in ms)))


def profile_counters():
    c = np.zeros(5, np.uint32)
    check(hip().rvs_profile_counters(c.ctypes.data))
    return {"replay_tiles": int(c[0]), "bin_overflow": int(c[1]), "max_bin": int(c[2]), "bin_entries": int(c[3]),
            "bin_work_items": int(c[4])}


# ------------------------------------------------------------- fused path

class DeviceSource:
    """An input view resident in HBM: raw planes + ERP table (InputView)."""

    def __init__(self, c, pos, rot, bd_color, bd_depth, znear, zfar, planes, has_invalid=True, device="cuda"):
        import torch
        self.cam = c
        self.pos, self.rot = list(pos), list(rot)
        self.bd_color, self.bd_depth = bd_color, bd_depth
        self.znear, self.zfar, self.has_invalid = znear, zfar, has_invalid
        self.planes = [torch.from_numpy(np.ascontiguousarray(p)).to(device) for p in planes]
        self.table = torch.from_numpy(erp_table(c)).to(device) if c.proj == 1 else None
        self.prep = None
        self.prep_scale = 0.0
        self.device = device

    def prepare(self, depth_scale, stream=None):
        """Builds the resident culling table (rvs_source_prep) for targets with this depth_scale."""
        import torch
        n = hip().rvs_source_prep_size(C.byref(self.cam))
        self.prep = torch.empty(n, dtype=torch.uint8, device=self.device)
        s = self.struct()
        check(hip().rvs_source_prep(C.byref(s), depth_scale, _ptr(self.prep), n, _stream(stream)))
        self.prep_scale = s.prep_depth_scale
        return self

    def struct(self):
        s = RvsSource()
        s.cam = self.cam
        s.bitdepth_color, s.bitdepth_depth = self.bd_color, self.bd_depth
        s.znear, s.zfar, s.has_invalid = self.znear, self.zfar, int(self.has_invalid)
        for k in range(3):
            s.pos[k], s.rot[k] = self.pos[k], self.rot[k]
        s.y, s.u, s.v, s.depth = (p.data_ptr() for p in self.planes)
        s.erp_table = self.table.data_ptr() if self.table is not None else None
        s.prep = self.prep.data_ptr() if self.prep is not None else None
        s.prep_depth_scale = self.prep_scale
        return s


def raster_size(target):
    """(w, h) at which a target is rasterised, blended and inpainted: its size
    scaled by its Precision, (int)(0.5f + w * rescale) (SynthesizedView.cpp:326-329)."""
    r = np.float32(target.rescale)
    if not (np.isfinite(r) and r > 0):   # refused by the library's plan; size the buffers as the view
        return (target.cam.w, target.cam.h)
    return (int(np.float32(0.5) + np.float32(target.cam.w) * r), int(np.float32(0.5) + np.float32(target.cam.h) * r))


def resize_color(src, out, stream=None):
    import torch
    sh, sw = src.shape[:2]
    dh, dw = out.shape[:2]
    nb = hip().rvs_resize_color_workspace_size(sw, sh, dw, dh)
    if nb == 0:
        raise RvsError(RVS_E_INVALID, "resize_color: bad sizes")
    ws = torch.empty(nb, dtype=torch.uint8, device=src.device)   # stream-ordered reuse only after the kernel
    check(hip().rvs_resize_color(_ptr(src), sw, sh, _ptr(out), dw, dh, _ptr(ws), nb, _stream(stream)))


def make_target(c, pos, rot, bitdepth, blending_exp, color_scale, depth_scale, rescale=1.0, tile=0):
    """tile: raster tile edge (0 / 16 or 32; a launch set uses its views' largest)."""
    t = RvsTarget()
    t.cam = c
    for k in range(3):
        t.pos[k], t.rot[k] = pos[k], rot[k]
    t.bitdepth, t.blending_exp, t.rescale = bitdepth, blending_exp, rescale
    t.color_scale, t.depth_scale = color_scale, depth_scale
    t.tile = tile
    return t


class Synthesizer:
    """One-view-at-a-time synthesis (the plain per-view path): owns a workspace
    and output buffers for one target size."""

    def __init__(self, sources, target, entry_cap=0, device="cuda", want_maps=True, want_color=True, tile=None):
        """want_color False: only the encoded planes (see BatchSynthesizer);
        tile: the raster tile edge (None: as the target says)."""
        import torch
        self.sources = sources
        if tile is not None:
            target = RvsTarget.from_buffer_copy(target)
            target.tile = tile
        self.target = target
        self.entry_cap = entry_cap
        self.device = device
        self.arr = (RvsSource * len(sources))(*[s.struct() for s in sources])
        w, h = target.cam.w, target.cam.h
        rw, rh = raster_size(target)   # float maps at the raster size
        dt = torch.uint8 if target.bitdepth <= 8 else torch.int16
        self.out = {
            "color": torch.empty((rh, rw, 3), dtype=torch.float32, device=device) if want_color else None,
            "quality": torch.empty((rh, rw), dtype=torch.float32, device=device) if want_maps else None,
            "depth": torch.empty((rh, rw), dtype=torch.float32, device=device) if want_maps else None,
            "validity": torch.empty((rh, rw), dtype=torch.float32, device=device) if want_maps else None,
            "y": torch.empty((h, w), dtype=dt, device=device),
            "u": torch.empty((h // 2, w // 2), dtype=dt, device=device),
            "v": torch.empty((h // 2, w // 2), dtype=dt, device=device),
        }
        self._alloc_ws()

    def _alloc_ws(self):
        import torch
        self.ws_bytes = hip().rvs_synth_workspace_size(self.arr, len(self.sources), C.byref(self.target),
                                                       self.entry_cap)
        if self.ws_bytes == 0:
            raise RvsError(-1, hip().rvs_last_error().decode())
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)

    def set_pose(self, pos, rot):
        for k in range(3):
            self.target.pos[k], self.target.rot[k] = pos[k], rot[k]

    def run(self, stream=None):
        vo = RvsViewOut()
        for k in ("color", "quality", "depth", "validity", "y", "u", "v"):
            setattr(vo, k, self.out[k].data_ptr() if self.out[k] is not None else None)
        needed = C.c_uint32(0)
        rc = hip().rvs_synthesize(self.arr, len(self.sources), C.byref(self.target), _ptr(self.ws),
                                  self.ws_bytes, self.entry_cap, C.byref(vo), C.byref(needed), _stream(stream))
        if rc == RVS_E_CAPACITY and needed.value > 0:
            self.entry_cap = grow_cap(needed.value)
            self._alloc_ws()
            return self.run(stream)
        check(rc)
        return self.out


def synthesize_noalloc(sources, target, out, entry_cap=0, stream=None):
    """rvs_synthesize_noalloc: one view with the library's static device
    scratch as workspace (no allocation).  out: dict of output tensors as
    Synthesizer.out (None entries allowed except colour).  Returns the needed
    entry capacity on RVS_E_CAPACITY (list overflow), else 0."""
    arr = (RvsSource * len(sources))(*[s.struct() for s in sources])
    vo = RvsViewOut()
    for k in ("color", "quality", "depth", "validity", "y", "u", "v"):
        setattr(vo, k, out[k].data_ptr() if out.get(k) is not None else None)
    needed = C.c_uint32(0)
    rc = hip().rvs_synthesize_noalloc(arr, len(sources), C.byref(target), entry_cap, C.byref(vo), C.byref(needed),
                                      _stream(stream))
    if rc == RVS_E_CAPACITY and needed.value > 0:
        return int(needed.value)
    check(rc)
    return 0


def grow_cap(needed):
    """List capacity for a re-run after RVS_E_CAPACITY: needed entries + 1/8, as uint32."""
    return min(int(needed) + int(needed) // 8 + 1024, 0xFFFFFFFF)


def _next_pow2(n):
    cap = 1
    while cap < n:
        cap <<= 1
    return cap


class BatchSynthesizer:
    """Many target views from the same resident sources in one launch set
    (rvs_synthesize_batch).  targets: list of RvsTarget (sizes may differ)."""

    def __init__(self, sources, targets, entry_cap=0, device="cuda", want_maps=False, want_color=True, tile=None):
        """want_color False: only the encoded planes (the colour map stays in
        the workspace, written only where the inpainting needs it); tile: the
        raster tile edge for every view (None: as the targets say)."""
        import torch
        self.sources = sources
        self.arr = (RvsSource * len(sources))(*[s.struct() for s in sources])
        self.entry_cap = entry_cap
        self.device = device
        self.want_color = want_color
        self.tile = tile
        self.set_targets(targets, want_maps)

    def set_targets(self, targets, want_maps=False):
        import torch
        self.n = len(targets)
        self.targets = (RvsTarget * self.n)(*targets)
        if self.tile is not None:
            for t in self.targets:
                t.tile = self.tile
        self.outs = []
        vo = (RvsViewOut * self.n)()
        for k, t in enumerate(targets):
            w, h = t.cam.w, t.cam.h
            rw, rh = raster_size(t)
            dt = torch.uint8 if t.bitdepth <= 8 else torch.int16
            o = {"color": torch.empty((rh, rw, 3), dtype=torch.float32, device=self.device) if self.want_color
                 else None,
                 "y": torch.empty((h, w), dtype=dt, device=self.device),
                 "u": torch.empty((h // 2, w // 2), dtype=dt, device=self.device),
                 "v": torch.empty((h // 2, w // 2), dtype=dt, device=self.device)}
            for key in ("quality", "depth", "validity"):
                o[key] = torch.empty((rh, rw), dtype=torch.float32, device=self.device) if want_maps else None
            self.outs.append(o)
            for key in ("color", "quality", "depth", "validity", "y", "u", "v"):
                setattr(vo[k], key, o[key].data_ptr() if o[key] is not None else None)
        self.vo = vo
        self._alloc_ws()

    def _alloc_ws(self):
        import torch
        self.ws_bytes = hip().rvs_synth_batch_workspace_size(self.arr, len(self.sources), self.targets, self.n,
                                                             self.entry_cap)
        if self.ws_bytes == 0:
            raise RvsError(-1, hip().rvs_last_error().decode())
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)

    def set_pose(self, k, pos, rot):
        for i in range(3):
            self.targets[k].pos[i], self.targets[k].rot[i] = pos[i], rot[i]

    def run(self, stream=None):
        needed = C.c_uint32(0)
        rc = hip().rvs_synthesize_batch(self.arr, len(self.sources), self.targets, self.n, _ptr(self.ws),
                                        self.ws_bytes, self.entry_cap, self.vo, C.byref(needed), _stream(stream))
        if rc == RVS_E_CAPACITY and needed.value > 0:
            self.entry_cap = grow_cap(needed.value)
            self._alloc_ws()
            return self.run(stream)
        check(rc)
        return self.outs

    def enqueue(self, stream=None):
        """Enqueue only (capturable in a HIP graph); call check() afterwards."""
        check(hip().rvs_synthesize_batch_async(self.arr, len(self.sources), self.targets, self.n, _ptr(self.ws),
                                               self.ws_bytes, self.entry_cap, self.vo, _stream(stream)))

    def check(self, stream=None):
        """Returns 0, or the entry capacity needed when the binned entries overflowed."""
        needed = C.c_uint32(0)
        rc = hip().rvs_synth_check(self.arr, len(self.sources), self.targets, self.n, _ptr(self.ws), self.entry_cap,
                                   C.byref(needed), _stream(stream))
        if rc == RVS_E_CAPACITY and needed.value > 0:
            return int(needed.value)
        check(rc)
        return 0


class StagewiseSynthesizer:
    """The plain per-view path with the reference's per-stage call structure
    (Pipeline::computeView over rvs_cuda_lib's entry points, Pipeline.cpp:169-301):
    input views decoded once at load; per view and source: reproject, raster
    (synchronous status check), blend; then inpaint and encode.  Same results
    as Synthesizer."""

    def __init__(self, sources, target, device="cuda", multispectral=None, rgb=False):
        """multispectral: None (BlendingMethod Simple with target.blending_exp)
        or (low exponent, high exponent): BlendedViewMultiSpec, ColorSpace RGB
        when rgb."""
        import torch
        self.sources = sources
        self.target = target
        self.device = device
        self.multispectral = multispectral
        self.rgb = rgb
        self.dec = []
        for s in sources:
            c = s.cam
            col = torch.empty((c.h, c.w, 3), dtype=torch.float32, device=device)
            dep = torch.empty((c.h, c.w), dtype=torch.float32, device=device)
            y, u, v, d = s.planes
            decode_yuv420(y, u, v, s.bd_color, target.color_scale, col)
            decode_depth(d, s.bd_depth, target.depth_scale, s.znear, s.zfar, s.has_invalid, dep)
            uv = torch.empty((c.h, c.w, 2), dtype=torch.float32, device=device)
            vd = torch.empty((c.h, c.w), dtype=torch.float32, device=device)
            self.dec.append((col, dep, uv, vd))
        w, h = target.cam.w, target.cam.h
        rw, rh = raster_size(target)
        f32 = dict(dtype=torch.float32, device=device)
        self.rc = torch.empty((rh, rw, 3), **f32)
        self.rd = torch.empty((rh, rw), **f32)
        self.rs = torch.empty((rh, rw), **f32)
        self.acc = {"color": torch.empty((rh, rw, 3), **f32), "quality": torch.empty((rh, rw), **f32),
                    "depth": torch.empty((rh, rw), **f32), "validity": torch.empty((rh, rw), **f32),
                    "mask": torch.empty((rh, rw), dtype=torch.uint8, device=device)}
        if multispectral is not None:
            # the two bands' accumulators and the split of the current view
            self.bands = [{k: torch.empty_like(v) for k, v in self.acc.items()} for _ in range(2)]
            self.split = (torch.empty((rh, rw, 3), **f32), torch.empty((rh, rw, 3), **f32))
            self.sws = None
        # write_color resizes an oversampled colour to the view's size before quantising
        self.resized = torch.empty((h, w, 3), **f32) if (rw, rh) != (w, h) else None
        dt = torch.uint8 if target.bitdepth <= 8 else torch.int16
        self.out = {"y": torch.empty((h, w), dtype=dt, device=device),
                    "u": torch.empty((h // 2, w // 2), dtype=dt, device=device),
                    "v": torch.empty((h // 2, w // 2), dtype=dt, device=device)}
        self.rws = None
        self.iws = None

    def run(self, stream=None):
        t = self.target
        w, h = raster_size(t)
        wrap = t.cam.proj == 1 and t.cam.full_hor
        for k, s in enumerate(self.sources):
            col, dep, uv, vd = self.dec[k]
            Rm, tv = relative_pose(s.pos, s.rot, list(t.pos), list(t.rot))
            reproject(s.cam, t.cam, Rm, tv, dep, s.table, t.rescale, uv, vd, stream)
            self.rws = raster(col, vd, uv, w, h, wrap, t.rescale, self.rc, self.rd, self.rs, workspace=self.rws,
                              stream=stream)
            if self.multispectral is None:
                blend_step(self.acc, self.rc, self.rd, self.rs, t.blending_exp, k == 0, stream)
                continue
            low, high = self.split
            self.sws = split_frequencies(self.rc, self.rd, self.rs, low, high, self.rgb, self.sws, stream)
            blend_step(self.bands[0], low, self.rd, self.rs, self.multispectral[0], k == 0, stream)
            blend_step(self.bands[1], high, self.rd, self.rs, self.multispectral[1], k == 0, stream)
        if self.multispectral is not None:
            # low + high colour with the low band's maps (BlendedView.cpp:86)
            add_color(self.bands[0]["color"], self.bands[1]["color"], self.acc["color"], stream)
            import torch
            with torch.cuda.stream(torch.cuda.ExternalStream(stream) if stream is not None else
                                   torch.cuda.current_stream()):
                for key in ("quality", "depth", "validity"):
                    self.acc[key].copy_(self.bands[0][key])
        self.iws = inpaint(self.acc["color"], self.acc["quality"], workspace=self.iws, stream=stream)
        enc = self.acc["color"]
        if self.resized is not None:
            resize_color(enc, self.resized, stream)
            enc = self.resized
        encode_yuv420(enc, t.bitdepth, self.out["y"], self.out["u"], self.out["v"], stream)
        return dict(self.out, color=self.acc["color"], quality=self.acc["quality"], depth=self.acc["depth"],
                    validity=self.acc["validity"])
