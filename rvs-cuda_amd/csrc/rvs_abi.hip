// rvs_abi.hip -- extern "C" entry points declared in include/rvs_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rvs_hip.h"
#include "rvs_inpaint.h"
#include "rvs_kernels.h"
#include "rvs_stages.h"

using rvs_dev::ArrayGeom;
using rvs_dev::PairGeom;

namespace {

thread_local char g_err[512] = "";

// stage events of rvs_synthesize (rvs_profile_enable)
thread_local bool g_prof = false;
thread_local bool g_prof_valid = false;
thread_local hipEvent_t g_ev[RVS_PROF_STAGES + 1] = {};
thread_local rvs_k::RasterWS g_prof_ws;
thread_local rvs_k::RasterDims g_prof_dims;
thread_local int g_prof_nsrc = 0;

void prof_mark(int k, hipStream_t s) {
    if (g_prof) hipEventRecord(g_ev[k], s);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* where) {
    return fail(RVS_E_HIP, "%s: %s", where, hipGetErrorString(e));
}

#define HIP_TRY(expr, where)                         \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) return hip_fail(_e, where); \
    } while (0)

// ---- host float semantics of the reference (Parameters.cpp, SpaceTransformer.cpp)

void matmul33(const float* A, const float* B, float* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.f;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
}

void euler_to_R(const float rot[3], float R[9]) {
    const float radperdeg = 0.01745329252f;
    const float rz = radperdeg * rot[0], ry = radperdeg * rot[1], rx = radperdeg * rot[2];
    const float Rz[9] = {std::cos(rz), -std::sin(rz), 0.f, std::sin(rz), std::cos(rz), 0.f, 0.f, 0.f, 1.f};
    const float Ry[9] = {std::cos(ry), 0.f, std::sin(ry), 0.f, 1.f, 0.f, -std::sin(ry), 0.f, std::cos(ry)};
    const float Rx[9] = {1.f, 0.f, 0.f, 0.f, std::cos(rx), -std::sin(rx), 0.f, std::sin(rx), std::cos(rx)};
    float T[9];
    matmul33(Rz, Ry, T);
    matmul33(T, Rx, R);
}

bool cam_ok(const rvs_cam* c) {
    return c && (c->proj == 0 || c->proj == 1) && c->w > 1 && c->h > 1;
}

// Target-side constants of EquirectangularProjector::project (EquirectangularProjector.cpp:69-73)
// and the raster rescale of SynthesizedView.cpp:326-331.
void fill_target(PairGeom& g, const rvs_cam* dst, float rescale) {
    g.dst_erp = dst->proj == 1;
    g.df0n = -dst->f[0]; g.df1n = -dst->f[1];
    g.dp0 = dst->p[0]; g.dp1 = dst->p[1];
    const float degperrad = 57.295779513f;
    g.u0 = dst->w * dst->hor[1] / (dst->hor[1] - dst->hor[0]);
    g.v0 = dst->h * dst->ver[1] / (dst->ver[1] - dst->ver[0]);
    g.du_dphi = -degperrad * dst->w / (dst->hor[1] - dst->hor[0]);
    g.dv_dtheta = -degperrad * dst->h / (dst->ver[1] - dst->ver[0]);
    const int ow = (int)(0.5f + dst->w * rescale), oh = (int)(0.5f + dst->h * rescale);
    g.sx = float(ow) / dst->w;
    g.sy = float(oh) / dst->h;
    g.wrap = (dst->proj == 1 && dst->full_hor) ? 1 : 0;
    g.rescale = rescale;
}

void fill_source(PairGeom& g, const rvs_cam* src, const float* tab) {
    g.sw = src->w; g.sh = src->h;
    g.src_erp = src->proj == 1;
    g.sf0 = src->f[0]; g.sf1 = src->f[1]; g.sp0 = src->p[0]; g.sp1 = src->p[1];
    g.tab = tab;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct RasterLayout {
    size_t counts, lists, deferred, dtiles, status, total;
};

RasterLayout raster_layout(int nsrc, int out_w, int out_h, uint32_t cap) {
    const int tiles_x = (out_w + rvs_k::kTile - 1) / rvs_k::kTile;
    const int tiles_y = (out_h + rvs_k::kTile - 1) / rvs_k::kTile;
    const size_t nt = (size_t)tiles_x * tiles_y;
    RasterLayout L;
    size_t o = 0;
    L.status = o; o += align256(16 * sizeof(uint32_t));
    L.counts = o; o += align256(nsrc * nt * sizeof(uint32_t));
    L.lists = o; o += align256((size_t)nsrc * nt * cap * sizeof(uint32_t));
    L.deferred = o; o += align256((size_t)out_w * out_h);
    L.dtiles = o; o += align256(nt * sizeof(uint32_t));
    L.total = o;
    return L;
}

rvs_k::RasterDims dims_of(int out_w, int out_h, uint32_t cap) {
    rvs_k::RasterDims d;
    d.out_w = out_w; d.out_h = out_h;
    d.tiles_x = (out_w + rvs_k::kTile - 1) / rvs_k::kTile;
    d.tiles_y = (out_h + rvs_k::kTile - 1) / rvs_k::kTile;
    d.ntiles = d.tiles_x * d.tiles_y;
    d.cap = cap;
    return d;
}

rvs_k::RasterWS carve_raster(char* base, const RasterLayout& L) {
    rvs_k::RasterWS ws;
    ws.status = (uint32_t*)(base + L.status);
    ws.counts = (uint32_t*)(base + L.counts);
    ws.lists = (uint32_t*)(base + L.lists);
    ws.deferred = (uint8_t*)(base + L.deferred);
    ws.deferred_tiles = (uint32_t*)(base + L.dtiles);
    return ws;
}

// After a raster launch set: report bin overflow (synchronises the stream).
int check_raster_status(const rvs_k::RasterWS& ws, const rvs_k::RasterDims& d, int nsrc, uint32_t* needed_cap,
                        hipStream_t s) {
    uint32_t st[4];
    HIP_TRY(hipMemcpyAsync(st, ws.status, sizeof st, hipMemcpyDeviceToHost, s), "status copy");
    HIP_TRY(hipStreamSynchronize(s), "stream sync");
    if (st[1]) {
        std::vector<uint32_t> counts((size_t)nsrc * d.ntiles);
        HIP_TRY(hipMemcpy(counts.data(), ws.counts, counts.size() * 4, hipMemcpyDeviceToHost), "counts copy");
        uint32_t mx = 0;
        for (uint32_t c : counts) mx = c > mx ? c : mx;
        if (needed_cap) *needed_cap = mx;
        return fail(RVS_E_CAPACITY, "tile bin overflow: a (source, tile) bin holds %u quads, capacity %u", mx, d.cap);
    }
    if (st[3]) return fail(RVS_E_CAPACITY, "in-order replay of a tile bin of %u quads exceeds %d", st[3],
                           rvs_k::kFallbackSortCap);
    return RVS_OK;
}

}  // namespace

extern "C" {

const char* rvs_last_error(void) { return g_err; }
const char* rvs_version(void) { return "rvs-mi355x 0.1 (gfx950)"; }

int rvs_relative_pose(const float pos_from[3], const float rot_from[3], const float pos_to[3],
                      const float rot_to[3], rvs_xform* out) {
    if (!pos_from || !rot_from || !pos_to || !rot_to || !out) return fail(RVS_E_INVALID, "null argument");
    float Rf[9], Rt[9], RtT[9], nRtT[9];
    euler_to_R(rot_from, Rf);
    euler_to_R(rot_to, Rt);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) RtT[i * 3 + j] = Rt[j * 3 + i];
    matmul33(RtT, Rf, out->R);
    for (int i = 0; i < 9; ++i) nRtT[i] = RtT[i] * -1.f;
    const float d[3] = {pos_to[0] - pos_from[0], pos_to[1] - pos_from[1], pos_to[2] - pos_from[2]};
    for (int i = 0; i < 3; ++i) {
        float s = 0.f;
        for (int k = 0; k < 3; ++k) s += nRtT[i * 3 + k] * d[k];
        out->t[i] = s;
    }
    return RVS_OK;
}

size_t rvs_erp_table_len(const rvs_cam* cam) { return cam ? 2 * (size_t)cam->w + 2 * (size_t)cam->h : 0; }

int rvs_erp_table_fill(const rvs_cam* cam, float* out) {
    if (!cam_ok(cam) || cam->proj != 1 || !out) return fail(RVS_E_INVALID, "ERP camera and output required");
    const float radperdeg = 0.01745329252f;
    const float phi0 = radperdeg * cam->hor[1];
    const float theta0 = radperdeg * cam->ver[1];
    const float dphi_du = -radperdeg * (cam->hor[1] - cam->hor[0]) / cam->w;
    const float dtheta_dv = -radperdeg * (cam->ver[1] - cam->ver[0]) / cam->h;
    for (int j = 0; j < cam->w; ++j) {
        const float phi = phi0 + dphi_du * (j + 0.5f);
        out[j] = std::cos(phi);
        out[cam->w + j] = std::sin(phi);
    }
    const float eps = 1e-3f;
    for (int i = 0; i < cam->h; ++i) {
        float v = i + 0.5f;
        if (i == 0) v = eps;
        if (i == cam->h - 1) v = cam->h - eps;
        const float theta = theta0 + dtheta_dv * v;
        out[2 * cam->w + i] = std::cos(theta);
        out[2 * cam->w + cam->h + i] = std::sin(theta);
    }
    return RVS_OK;
}

int rvs_decode_yuv420(const void* y, const void* u, const void* v, int bitdepth, int w, int h,
                      float color_scale, float* out, rvs_stream_t stream) {
    if (!y || !u || !v || !out || w < 2 || h < 2 || bitdepth < 1 || bitdepth > 16)
        return fail(RVS_E_INVALID, "rvs_decode_yuv420: bad argument");
    HIP_TRY(rvs_k::launch_decode_color(y, u, v, bitdepth <= 8, w, h, color_scale, out, (hipStream_t)stream),
            "decode_yuv420");
    return RVS_OK;
}

int rvs_decode_depth(const void* raw, int bitdepth, int w, int h, float depth_scale, float znear, float zfar,
                     int has_invalid, float* out, rvs_stream_t stream) {
    if (!raw || !out || w < 1 || h < 1 || bitdepth < 1 || bitdepth > 16)
        return fail(RVS_E_INVALID, "rvs_decode_depth: bad argument");
    HIP_TRY(rvs_k::launch_decode_depth(raw, bitdepth <= 8, (size_t)w * h, depth_scale, znear, zfar, has_invalid,
                                       out, (hipStream_t)stream),
            "decode_depth");
    return RVS_OK;
}

int rvs_reproject(const rvs_cam* src, const rvs_cam* dst, const rvs_xform* xf, const float* depth,
                  const float* erp_table, float rescale, float* uv, float* vdepth, rvs_stream_t stream) {
    if (!cam_ok(src) || !cam_ok(dst) || !xf || !depth || !uv || !vdepth || !(rescale > 0.f))
        return fail(RVS_E_INVALID, "rvs_reproject: bad argument");
    if (src->proj == 1 && !erp_table) return fail(RVS_E_INVALID, "rvs_reproject: ERP source needs erp_table");
    PairGeom g;
    std::memset(&g, 0, sizeof g);
    fill_source(g, src, erp_table);
    g.depth = depth;
    g.depth_f32 = 1;
    std::memcpy(g.R, xf->R, sizeof g.R);
    std::memcpy(g.t, xf->t, sizeof g.t);
    fill_target(g, dst, rescale);
    HIP_TRY(rvs_k::launch_reproject(g, uv, vdepth, (hipStream_t)stream), "reproject");
    return RVS_OK;
}

size_t rvs_raster_workspace_size(int out_w, int out_h, uint32_t tile_cap) {
    if (tile_cap == 0) tile_cap = rvs_k::kDefaultTileCap;
    return raster_layout(1, out_w, out_h, tile_cap).total + align256(sizeof(ArrayGeom));
}

int rvs_raster(const float* color, const float* depth, const float* uv, int in_w, int in_h, int out_w, int out_h,
               int wrap, float rescale, void* workspace, size_t ws_bytes, uint32_t tile_cap, float* out_color,
               float* out_depth, float* out_shape, uint32_t* needed_cap, rvs_stream_t stream) {
    if (!color || !depth || !uv || in_w < 2 || in_h < 2 || out_w < 1 || out_h < 1 || !workspace || !out_color ||
        !out_depth || !out_shape || !(rescale > 0.f))
        return fail(RVS_E_INVALID, "rvs_raster: bad argument");
    if (tile_cap == 0) tile_cap = rvs_k::kDefaultTileCap;
    const size_t need = rvs_raster_workspace_size(out_w, out_h, tile_cap);
    if (ws_bytes < need) return fail(RVS_E_INVALID, "rvs_raster: workspace %zu < %zu bytes", ws_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    const RasterLayout L = raster_layout(1, out_w, out_h, tile_cap);
    char* base = (char*)workspace;
    rvs_k::RasterWS ws = carve_raster(base, L);
    const rvs_k::RasterDims d = dims_of(out_w, out_h, tile_cap);
    ArrayGeom* dg = (ArrayGeom*)(base + L.total);
    ArrayGeom hg;
    hg.sw = in_w; hg.sh = in_h; hg.uv = uv; hg.vd = depth; hg.color = color; hg.wrap = wrap ? 1 : 0;
    hg.rescale = rescale;
    HIP_TRY(hipMemcpyAsync(dg, &hg, sizeof hg, hipMemcpyHostToDevice, s), "geom upload");
    HIP_TRY(rvs_k::launch_reset_u32(ws.status, 16, s), "reset");
    HIP_TRY(rvs_k::launch_reset_u32(ws.counts, (size_t)d.ntiles, s), "reset");
    HIP_TRY(rvs_k::launch_bin<ArrayGeom>(dg, 1, in_w, in_h, d, ws, s), "bin");
    rvs_k::RasterOut o;
    o.color = out_color; o.depth = out_depth; o.quality = nullptr; o.validity = out_shape;
    HIP_TRY((rvs_k::launch_raster<ArrayGeom, rvs_k::MODE_RAW>(dg, 1, d, ws, 0.f, o, s)), "raster");
    return check_raster_status(ws, d, 1, needed_cap, s);
}

int rvs_blend_step(float* acc_color, float* acc_quality, float* acc_depth, float* acc_validity, uint8_t* acc_mask,
                   const float* color, const float* depth, const float* shape, int w, int h, float blending_exp,
                   int first, rvs_stream_t stream) {
    if (!acc_color || !acc_quality || !acc_depth || !acc_validity || !acc_mask || !color || !depth || !shape ||
        w < 1 || h < 1)
        return fail(RVS_E_INVALID, "rvs_blend_step: bad argument");
    HIP_TRY(rvs_k::launch_blend_step(acc_color, acc_quality, acc_depth, acc_validity, acc_mask, color, depth, shape,
                                     (size_t)w * h, blending_exp, first, (hipStream_t)stream),
            "blend_step");
    return RVS_OK;
}

size_t rvs_inpaint_workspace_size(int w, int h) { return rvs_k::inpaint_ws_bytes(w, h); }

int rvs_inpaint(float* color, const float* quality, int w, int h, void* workspace, size_t ws_bytes, int32_t* map_out,
                rvs_stream_t stream) {
    if (!color || !quality || w < 1 || h < 1 || !workspace) return fail(RVS_E_INVALID, "rvs_inpaint: bad argument");
    if (ws_bytes < rvs_k::inpaint_ws_bytes(w, h)) return fail(RVS_E_INVALID, "rvs_inpaint: workspace too small");
    const rvs_k::InpaintWS ws = rvs_k::inpaint_ws_carve(workspace, w, h);
    HIP_TRY(rvs_k::launch_inpaint(color, quality, w, h, ws, map_out, (hipStream_t)stream), "inpaint");
    return RVS_OK;
}

int rvs_encode_yuv420(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v,
                      rvs_stream_t stream) {
    if (!color || !y || !u || !v || w < 2 || h < 2 || (w & 1) || (h & 1) || bitdepth < 1 || bitdepth > 16)
        return fail(RVS_E_INVALID, "rvs_encode_yuv420: bad argument (even sizes required)");
    HIP_TRY(rvs_k::launch_encode(color, w, h, bitdepth, y, u, v, (hipStream_t)stream), "encode");
    return RVS_OK;
}

// ------------------------------------------------------------ fused path

static size_t synth_layout(int n_src, int w, int h, uint32_t cap, size_t* off_geom, size_t* off_raster,
                           size_t* off_inp, size_t* off_q) {
    size_t o = 0;
    *off_geom = o; o += align256((size_t)n_src * sizeof(PairGeom));
    *off_q = o; o += align256((size_t)w * h * sizeof(float));
    *off_raster = o; o += raster_layout(n_src, w, h, cap).total;
    *off_inp = o; o += align256(rvs_k::inpaint_ws_bytes(w, h));
    return o;
}

size_t rvs_synth_workspace_size(int n_src, const rvs_target* tgt, uint32_t tile_cap) {
    if (!tgt || n_src < 1) return 0;
    if (tile_cap == 0) tile_cap = rvs_k::kDefaultTileCap;
    size_t a, b, c, d;
    return synth_layout(n_src, tgt->cam.w, tgt->cam.h, tile_cap, &a, &b, &c, &d);
}

int rvs_synthesize(const rvs_source* srcs, int n_src, const rvs_target* tgt, void* workspace, size_t ws_bytes,
                   uint32_t tile_cap, const rvs_view_out* out, uint32_t* needed_cap, rvs_stream_t stream) {
    if (!srcs || n_src < 1 || n_src > RVS_MAX_SOURCES || !tgt || !cam_ok(&tgt->cam) || !workspace || !out ||
        !out->color)
        return fail(RVS_E_INVALID, "rvs_synthesize: bad argument");
    if (tgt->rescale != 1.0f) return fail(RVS_E_UNSUPPORTED, "rvs_synthesize: Precision != 1 is not supported");
    if (tile_cap == 0) tile_cap = rvs_k::kDefaultTileCap;
    const int w = tgt->cam.w, h = tgt->cam.h;
    size_t og, orr, oi, oq;
    const size_t need = synth_layout(n_src, w, h, tile_cap, &og, &orr, &oi, &oq);
    if (ws_bytes < need) return fail(RVS_E_INVALID, "rvs_synthesize: workspace %zu < %zu bytes", ws_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    char* base = (char*)workspace;
    std::vector<PairGeom> hg(n_src);
    int max_sw = 0, max_sh = 0;
    for (int k = 0; k < n_src; ++k) {
        const rvs_source& S = srcs[k];
        if (!cam_ok(&S.cam) || !S.y || !S.u || !S.v || !S.depth || (S.cam.proj == 1 && !S.erp_table))
            return fail(RVS_E_INVALID, "rvs_synthesize: source %d incomplete", k);
        PairGeom& g = hg[k];
        std::memset(&g, 0, sizeof g);
        fill_source(g, &S.cam, S.erp_table);
        g.depth = S.depth;
        g.depth_u8 = S.bitdepth_depth <= 8;
        g.depth_f32 = 0;
        g.dscale = tgt->depth_scale;
        g.znear = S.znear; g.zfar = S.zfar; g.has_invalid = S.has_invalid;
        g.py = S.y; g.pu = S.u; g.pv = S.v;
        g.color_u8 = S.bitdepth_color <= 8;
        g.cscale = tgt->color_scale;
        rvs_xform xf;
        rvs_relative_pose(S.pos, S.rot, tgt->pos, tgt->rot, &xf);
        std::memcpy(g.R, xf.R, sizeof g.R);
        std::memcpy(g.t, xf.t, sizeof g.t);
        fill_target(g, &tgt->cam, tgt->rescale);
        max_sw = S.cam.w > max_sw ? S.cam.w : max_sw;
        max_sh = S.cam.h > max_sh ? S.cam.h : max_sh;
    }
    g_prof_valid = false;
    PairGeom* dg = (PairGeom*)(base + og);
    HIP_TRY(rvs_k::launch_store_geoms(hg.data(), n_src, dg, s), "store geoms");
    const RasterLayout L = raster_layout(n_src, w, h, tile_cap);
    rvs_k::RasterWS ws = carve_raster(base + orr, L);
    const rvs_k::RasterDims d = dims_of(w, h, tile_cap);
    HIP_TRY(rvs_k::launch_reset_u32(ws.status, 16, s), "reset");
    HIP_TRY(rvs_k::launch_reset_u32(ws.counts, (size_t)n_src * d.ntiles, s), "reset");
    prof_mark(0, s);
    HIP_TRY(rvs_k::launch_bin<PairGeom>(dg, n_src, max_sw, max_sh, d, ws, s), "bin");
    prof_mark(1, s);
    float* quality = out->quality ? out->quality : (float*)(base + oq);
    rvs_k::RasterOut ro;
    ro.color = out->color; ro.depth = out->depth; ro.quality = quality; ro.validity = out->validity;
    HIP_TRY((rvs_k::launch_raster<PairGeom, rvs_k::MODE_BLEND>(dg, n_src, d, ws, tgt->blending_exp, ro, s)),
            "raster");
    prof_mark(2, s);
    const rvs_k::InpaintWS iws = rvs_k::inpaint_ws_carve(base + oi, w, h);
    HIP_TRY(rvs_k::launch_inpaint(out->color, quality, w, h, iws, nullptr, s), "inpaint");
    prof_mark(3, s);
    if (out->y && out->u && out->v) {
        if ((w & 1) || (h & 1)) return fail(RVS_E_UNSUPPORTED, "rvs_synthesize: odd target size cannot be encoded");
        HIP_TRY(rvs_k::launch_encode(out->color, w, h, tgt->bitdepth, out->y, out->u, out->v, s), "encode");
    }
    prof_mark(4, s);
    const int rc = check_raster_status(ws, d, n_src, needed_cap, s);
    g_prof_valid = g_prof && rc == RVS_OK;
    if (g_prof_valid) { g_prof_ws = ws; g_prof_dims = d; g_prof_nsrc = n_src; }
    return rc;
}

int rvs_profile_counters(uint32_t out[4]) {
    if (!g_prof_valid) return fail(RVS_E_INVALID, "no profiled rvs_synthesize on this thread");
    uint32_t st[4];
    HIP_TRY(hipMemcpy(st, g_prof_ws.status, sizeof st, hipMemcpyDeviceToHost), "status copy");
    std::vector<uint32_t> counts((size_t)g_prof_nsrc * g_prof_dims.ntiles);
    HIP_TRY(hipMemcpy(counts.data(), g_prof_ws.counts, counts.size() * 4, hipMemcpyDeviceToHost), "counts copy");
    uint32_t mx = 0;
    uint64_t tot = 0;
    for (uint32_t c : counts) { mx = c > mx ? c : mx; tot += c; }
    out[0] = st[0]; out[1] = st[1]; out[2] = mx; out[3] = (uint32_t)(tot > 0xffffffffull ? 0xffffffffu : tot);
    return RVS_OK;
}

int rvs_profile_enable(int on) {
    if (on && !g_ev[0])
        for (int k = 0; k <= RVS_PROF_STAGES; ++k) HIP_TRY(hipEventCreate(&g_ev[k]), "event create");
    g_prof = on != 0;
    g_prof_valid = false;
    return RVS_OK;
}

int rvs_profile_read(float ms[RVS_PROF_STAGES]) {
    if (!g_prof_valid) return fail(RVS_E_INVALID, "no profiled rvs_synthesize on this thread");
    for (int k = 0; k < RVS_PROF_STAGES; ++k) HIP_TRY(hipEventElapsedTime(&ms[k], g_ev[k], g_ev[k + 1]), "event time");
    return RVS_OK;
}

}  // extern "C"
