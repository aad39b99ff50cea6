"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use this module, always as the checker, never as the measured path.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_JSON_PATH = os.path.join(ROOT, "oracle", "_ref", "libref_json.so")

fptr = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
iptr = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcCam(C.Structure):
    _fields_ = [("proj", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("f", C.c_float * 2), ("p", C.c_float * 2),
                ("hor", C.c_float * 2), ("ver", C.c_float * 2),
                ("full_hor", C.c_int)]


class OrcBlender(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("empty", C.c_int), ("exp", C.c_float),
                ("color", C.c_void_p), ("depth", C.c_void_p), ("quality", C.c_void_p),
                ("validity", C.c_void_p), ("depth_mask", C.c_void_p)]


class OrcSource(C.Structure):
    _fields_ = [("cam", OrcCam), ("pos", C.c_float * 3), ("rot", C.c_float * 3),
                ("bitdepth_color", C.c_int), ("bitdepth_depth", C.c_int),
                ("znear", C.c_float), ("zfar", C.c_float), ("has_invalid", C.c_int),
                ("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("depth", C.c_void_p)]


class OrcTiming(C.Structure):
    _fields_ = [("t_decode", C.c_double), ("t_reproject", C.c_double), ("t_raster", C.c_double),
                ("t_blend", C.c_double), ("t_inpaint", C.c_double), ("t_encode", C.c_double),
                ("inpaint_sweeps", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_euler_to_R.argtypes = [fptr, fptr]
        L.orc_relative_pose.argtypes = [fptr, fptr, fptr, fptr, fptr, fptr]
        L.orc_generate_image_pos.argtypes = [C.POINTER(OrcCam), fptr]
        L.orc_unproject.argtypes = [C.POINTER(OrcCam), fptr, fptr, C.c_size_t, fptr]
        L.orc_affine.argtypes = [fptr, fptr, fptr, C.c_size_t, fptr]
        L.orc_project.argtypes = [C.POINTER(OrcCam), fptr, C.c_size_t, fptr, fptr]
        L.orc_project.restype = C.c_int
        L.orc_scale_uv.argtypes = [fptr, C.c_size_t, C.c_float, C.c_float, fptr]
        L.orc_reproject.argtypes = [C.POINTER(OrcCam), C.POINTER(OrcCam), fptr, fptr, fptr,
                                    C.c_float, fptr, fptr]
        L.orc_reproject.restype = C.c_int
        L.orc_transform_triangles.argtypes = [fptr, fptr, fptr, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_int, C.c_float, fptr, fptr, fptr]
        L.orc_blend_step.argtypes = [C.POINTER(OrcBlender), fptr, fptr, fptr]
        L.orc_inpaint.argtypes = [fptr, fptr, C.c_int, C.c_int, vp]
        L.orc_inpaint.restype = C.c_int
        L.orc_decode_color.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, fptr]
        L.orc_decode_depth.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                       C.c_float, C.c_int, fptr]
        L.orc_encode_yuv420.argtypes = [fptr, C.c_int, C.c_int, C.c_int, vp, vp, vp]
        L.orc_resize_color.argtypes = [fptr, C.c_int, C.c_int, fptr, C.c_int, C.c_int]
        u8ptr = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        L.orc_blur_masked.argtypes = [fptr, C.c_int, u8ptr, C.c_int, C.c_int, C.c_int, fptr]
        L.orc_blur_masked_rgb.argtypes = [fptr, u8ptr, C.c_int, C.c_int, C.c_int, fptr]
        L.orc_split_frequencies.argtypes = [fptr, fptr, fptr, C.c_int, C.c_int, C.c_int, fptr, fptr]
        L.orc_synthesize_multispectral.argtypes = [C.POINTER(OrcSource), C.c_int, C.POINTER(OrcCam), fptr, fptr,
                                                   C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, vp, vp, vp, vp,
                                                   vp, vp, vp]
        L.orc_synthesize_multispectral.restype = C.c_int
        L.orc_max_level.argtypes = [C.c_int]
        L.orc_max_level.restype = C.c_uint
        L.orc_synthesize.argtypes = [C.POINTER(OrcSource), C.c_int, C.POINTER(OrcCam), fptr, fptr,
                                     C.c_int, C.c_float, C.c_float, vp, vp, vp, vp, vp, vp, vp,
                                     C.POINTER(OrcTiming)]
        L.orc_synthesize.restype = C.c_int
        _lib = L
    return _lib


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def make_cam(proj, w, h, f=(0, 0), p=(0, 0), hor=(-180, 180), ver=(-90, 90), full_hor=None):
    c = OrcCam()
    c.proj = 1 if proj in (1, "erp", "Equirectangular") else 0
    c.w, c.h = int(w), int(h)
    c.f[0], c.f[1] = f
    c.p[0], c.p[1] = p
    c.hor[0], c.hor[1] = hor
    c.ver[0], c.ver[1] = ver
    if full_hor is None:
        full_hor = int(c.proj == 1 and int(round(hor[0])) == -180 and int(round(hor[1])) == 180
                       and float(hor[0]) == -180.0 and float(hor[1]) == 180.0)
    c.full_hor = int(full_hor)
    return c


def euler_to_R(rot):
    R = np.zeros(9, np.float32)
    lib().orc_euler_to_R(f32(rot), R)
    return R.reshape(3, 3)


def relative_pose(pos_from, rot_from, pos_to, rot_to):
    R = np.zeros(9, np.float32)
    t = np.zeros(3, np.float32)
    lib().orc_relative_pose(f32(pos_from), f32(rot_from), f32(pos_to), f32(rot_to), R, t)
    return R.reshape(3, 3), t


def generate_image_pos(cam):
    uv = np.zeros((cam.h, cam.w, 2), np.float32)
    lib().orc_generate_image_pos(C.byref(cam), uv)
    return uv


def unproject(cam, uv, depth):
    uv = f32(uv)
    depth = f32(depth)
    n = depth.size
    xyz = np.zeros(depth.shape + (3,), np.float32)
    lib().orc_unproject(C.byref(cam), uv, depth, n, xyz)
    return xyz


def affine(R, t, xyz):
    xyz = f32(xyz)
    out = np.zeros_like(xyz)
    lib().orc_affine(f32(R).ravel(), f32(t), xyz, xyz.size // 3, out)
    return out


def project(cam, xyz):
    xyz = f32(xyz)
    n = xyz.size // 3
    uv = np.zeros(xyz.shape[:-1] + (2,), np.float32)
    depth = np.zeros(xyz.shape[:-1], np.float32)
    wrap = lib().orc_project(C.byref(cam), xyz, n, uv, depth)
    return uv, depth, wrap


def reproject(src, dst, R, t, src_depth, rescale=1.0):
    src_depth = f32(src_depth)
    uv = np.zeros(src_depth.shape + (2,), np.float32)
    depth = np.zeros(src_depth.shape, np.float32)
    wrap = lib().orc_reproject(C.byref(src), C.byref(dst), f32(R).ravel(), f32(t), src_depth,
                               rescale, uv, depth)
    return uv, depth, wrap


def transform_triangles(color, depth, uv, out_w, out_h, wrap, rescale=1.0):
    color = f32(color)
    depth = f32(depth)
    uv = f32(uv)
    in_h, in_w = depth.shape
    oc = np.zeros((out_h, out_w, 3), np.float32)
    od = np.zeros((out_h, out_w), np.float32)
    osh = np.zeros((out_h, out_w), np.float32)
    lib().orc_transform_triangles(color, depth, uv, in_w, in_h, out_w, out_h, int(wrap),
                                  rescale, oc, od, osh)
    return oc, od, osh


class Blender:
    """BlendedViewSimple accumulator (restated in the oracle)."""

    def __init__(self, w, h, exp):
        self.color = np.zeros((h, w, 3), np.float32)
        self.depth = np.zeros((h, w), np.float32)
        self.quality = np.zeros((h, w), np.float32)
        self.validity = np.zeros((h, w), np.float32)
        self.depth_mask = np.zeros((h, w), np.uint8)
        self.s = OrcBlender(w, h, 1, exp, self.color.ctypes.data, self.depth.ctypes.data,
                            self.quality.ctypes.data, self.validity.ctypes.data,
                            self.depth_mask.ctypes.data)

    def blend(self, color, depth, shape):
        lib().orc_blend_step(C.byref(self.s), f32(color), f32(depth), f32(shape))


def inpaint(color, quality, want_map=False):
    color = f32(color).copy()
    quality = f32(quality)
    h, w = quality.shape
    m = np.zeros((h, w, 3), np.int32) if want_map else None
    sweeps = lib().orc_inpaint(color, quality, w, h, m.ctypes.data if want_map else None)
    return (color, m, sweeps) if want_map else color


def _plane_dtype(bitdepth):
    return np.uint8 if bitdepth <= 8 else np.uint16


def max_level(bitdepth):
    return int(lib().orc_max_level(bitdepth))


def decode_color(y, u, v, bitdepth, color_scale=None):
    h, w = y.shape
    dt = _plane_dtype(bitdepth)
    y, u, v = (np.ascontiguousarray(a, dtype=dt) for a in (y, u, v))
    if color_scale is None:
        color_scale = np.float32(1.0) / np.float32(max_level(bitdepth))
    out = np.zeros((h, w, 3), np.float32)
    lib().orc_decode_color(y.ctypes.data, u.ctypes.data, v.ctypes.data, bitdepth, w, h,
                           float(color_scale), out)
    return out


def decode_depth(raw, bitdepth, znear, zfar, has_invalid=True, depth_scale=None):
    h, w = raw.shape
    raw = np.ascontiguousarray(raw, dtype=_plane_dtype(bitdepth))
    if depth_scale is None:
        depth_scale = np.float32(1.0) / np.float32(max_level(bitdepth))
    out = np.zeros((h, w), np.float32)
    lib().orc_decode_depth(raw.ctypes.data, bitdepth, w, h, float(depth_scale), znear, zfar,
                           int(has_invalid), out)
    return out


def encode_yuv420(color, bitdepth):
    color = f32(color)
    h, w = color.shape[:2]
    dt = _plane_dtype(bitdepth)
    y = np.zeros((h, w), dt)
    u = np.zeros((h // 2, w // 2), dt)
    v = np.zeros((h // 2, w // 2), dt)
    lib().orc_encode_yuv420(color, w, h, bitdepth, y.ctypes.data, u.ctypes.data, v.ctypes.data)
    return y, u, v


def resize_color(src, dw, dh):
    src = f32(src)
    sh, sw = src.shape[:2]
    dst = np.zeros((dh, dw, 3), np.float32)
    lib().orc_resize_color(src, sw, sh, dst, dw, dh)
    return dst


def blur_masked(y, mask, rad):
    """One-channel masked box blur of the multispectral split (h x w float, u8 mask)."""
    y = f32(y)
    h, w = y.shape
    out = np.zeros((h, w), np.float32)
    lib().orc_blur_masked(y, 1, np.ascontiguousarray(mask, np.uint8), w, h, int(rad), out)
    return out


def blur_masked_rgb(color, mask, rad):
    """Three-channel masked box blur (ColorSpace RGB), h x w x 3 float."""
    color = f32(color)
    h, w = color.shape[:2]
    out = np.zeros((h, w, 3), np.float32)
    lib().orc_blur_masked_rgb(color, np.ascontiguousarray(mask, np.uint8), w, h, int(rad), out)
    return out


def split_frequencies(color, depth, shape, rgb=False):
    color, depth, shape = f32(color), f32(depth), f32(shape)
    h, w = depth.shape
    low = np.zeros((h, w, 3), np.float32)
    high = np.zeros((h, w, 3), np.float32)
    lib().orc_split_frequencies(color, depth, shape, w, h, int(bool(rgb)), low, high)
    return low, high


def synthesize(sources, tgt_cam, tgt_pos, tgt_rot, tgt_bitdepth, blending_exp, rescale=1.0, multispectral=None,
               rgb=False):
    """multispectral: None (BlendingMethod Simple, exponent blending_exp) or
    (low exponent, high exponent) for BlendingMethod Multispectral; rgb: its
    ColorSpace RGB split (all channels blurred).
    sources: list of dicts with keys cam, pos, rot, bd_color, bd_depth, znear, zfar,
    has_invalid, y, u, v, depth (numpy planes).  Returns dict of outputs."""
    n = len(sources)
    arr = (OrcSource * n)()
    keep = []
    for i, s in enumerate(sources):
        o = arr[i]
        o.cam = s["cam"]
        for k in range(3):
            o.pos[k] = s["pos"][k]
            o.rot[k] = s["rot"][k]
        o.bitdepth_color = s["bd_color"]
        o.bitdepth_depth = s["bd_depth"]
        o.znear, o.zfar = s["znear"], s["zfar"]
        o.has_invalid = int(s.get("has_invalid", True))
        planes = [np.ascontiguousarray(s[k]) for k in ("y", "u", "v", "depth")]
        keep += planes
        o.y, o.u, o.v, o.depth = (p.ctypes.data for p in planes)
    w, h = tgt_cam.w, tgt_cam.h
    # float maps at the raster size of Precision `rescale` (SynthesizedView.cpp:326-329)
    rw, rh = int(np.float32(0.5) + np.float32(w) * np.float32(rescale)), \
        int(np.float32(0.5) + np.float32(h) * np.float32(rescale))
    dt = _plane_dtype(tgt_bitdepth)
    out = {
        "color": np.zeros((rh, rw, 3), np.float32),
        "quality": np.zeros((rh, rw), np.float32),
        "depth": np.zeros((rh, rw), np.float32),
        "validity": np.zeros((rh, rw), np.float32),
        "y": np.zeros((h, w), dt),
        "u": np.zeros((h // 2, w // 2), dt),
        "v": np.zeros((h // 2, w // 2), dt),
    }
    tm = OrcTiming()
    ptrs = [out[k].ctypes.data for k in ("color", "quality", "depth", "validity", "y", "u", "v")]
    if multispectral is None:
        rc = lib().orc_synthesize(arr, n, C.byref(tgt_cam), f32(tgt_pos), f32(tgt_rot), tgt_bitdepth,
                                  blending_exp, rescale, *ptrs, C.byref(tm))
    else:
        rc = lib().orc_synthesize_multispectral(arr, n, C.byref(tgt_cam), f32(tgt_pos), f32(tgt_rot), tgt_bitdepth,
                                                float(multispectral[0]), float(multispectral[1]), int(bool(rgb)),
                                                rescale, *ptrs)
    if rc != 0:
        raise RuntimeError("orc_synthesize failed: %d" % rc)
    out["timing"] = {k: getattr(tm, k) for k, _ in OrcTiming._fields_}
    return out


_ref_json = None


def ref_json():
    """The reference's own JsonParser.cpp, built by oracle/Makefile (None if absent)."""
    global _ref_json
    if _ref_json is None:
        if not os.path.exists(REF_JSON_PATH):
            return None
        L = C.CDLL(REF_JSON_PATH)
        L.refjson_query.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        L.refjson_query.restype = C.c_int
        _ref_json = L
    return _ref_json


def ref_json_query(text, path, conv, overrides=None):
    L = ref_json()
    buf = C.create_string_buffer(4096)
    rc = L.refjson_query(text.encode(), (overrides or "").encode(), path.encode(), conv.encode(),
                         buf, 4096)
    return rc, buf.value.decode()
