// rvs_abi.hip -- extern "C" entry points declared in include/rvs_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/rvs_hip.h"
#include "rvs_inpaint.h"
#include "rvs_kernels.h"
#include "rvs_stages.h"

using rvs_dev::ArrayGeom;
using rvs_dev::PairGeom;

namespace {

thread_local char g_err[512] = "";

// stage events of rvs_synthesize* (rvs_profile_enable)
thread_local bool g_prof = false;
thread_local bool g_prof_valid = false;
thread_local hipEvent_t g_ev[RVS_PROF_STAGES + 1] = {};
thread_local rvs_k::RasterWS g_prof_ws;
thread_local rvs_k::BatchDims g_prof_dims;

void prof_mark(int k, hipStream_t s) {
    if (g_prof) (void)hipEventRecord(g_ev[k], s);
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

// Source preparation buffer: binning-block bounds, then (256-byte aligned)
// the decoded colour plane (float4 per source pixel).
size_t prep_color_offset(const rvs_cam* cam) {
    return align256(rvs_k::src_bounds_floats(cam->w, cam->h) * sizeof(float));
}

// The decoded colour plane of a prepared source is packed (8 B per pixel)
// when its upsampled chroma is a multiple of 1/4 below 2^14: even sizes, bit
// depth <= 14.
bool prep_packed(const rvs_source& S) {
    return S.bitdepth_color <= 14 && ((S.cam.w | S.cam.h) & 1) == 0;
}

// Sequential sub-allocation of a workspace, 256-byte aligned.
struct Layout {
    size_t o = 0;
    size_t take(size_t bytes) {
        const size_t r = o;
        o += align256(bytes);
        return r;
    }
};

// Raster state of one launch set: status words, compact bins (counts,
// offsets, fill cursors, entry list, scan sums, kept tile ranges and vertex
// blocks with their item map),
// deferred-pixel flags, deferred-tile list and binning work list.
struct RasterPlan {
    rvs_k::BatchDims d;
    size_t status, counts, offsets, fill, lists, ranges, ktab, vbuf, tvq, imap, bsum, deferred, dtiles, work, tile_enc,
        rbuf;
};

// List capacity: the caller's, or one entry per output pixel and source
// (at least 2^20); the kernels report the exact need when it is exceeded.
uint32_t entry_capacity(uint32_t requested, uint64_t npx, int nsrc) {
    if (requested) return requested;
    const uint64_t c = npx * (uint64_t)nsrc;
    return (uint32_t)(c < (1ull << 20) ? (1ull << 20) : (c > 0xffffffffull ? 0xffffffffull : c));
}

void plan_raster(Layout& L, RasterPlan& R, int nsrc, int nviews, uint32_t ntiles, uint32_t npx, uint32_t cap,
                 int max_sw, int max_sh, uint64_t keep_max = 0, int shift = 4, bool uneven_items = true,
                 int n_unculled = 0, double culled_px_frac = 1.0) {
    rvs_k::BatchDims& d = R.d;
    d.nviews = nviews; d.nsrc = nsrc; d.ntiles = ntiles; d.npx = npx;
    d.tile_shift = shift;
    // blocks of the passes over the kept items: items of equirectangular
    // targets differ widely in cost (those near a pole deal thousands of
    // tiles) and want many short-lived blocks to even out the tail; culled
    // perspective items are alike, and fewer blocks each taking several pay
    // the block set-up (and its spilled registers) fewer times
    d.flat_blocks = uneven_items ? rvs_k::kFlatBlocks : rvs_k::kFlatBlocks / 4;
    d.cap = entry_capacity(cap, npx, nsrc);
    d.max_sw = max_sw; d.max_sh = max_sh;
    d.bins_x = (uint32_t)((max_sw - 1 + rvs_k::kBinQX - 1) / rvs_k::kBinQX);
    d.bins_y = (uint32_t)((max_sh - 1 + rvs_k::kBinQY - 1) / rvs_k::kBinQY);
    const uint64_t items = (uint64_t)d.bins_x * d.bins_y * nsrc * nviews;
    // Work items that keep their tile ranges, vertex window and quad validities (6.9 KB each):
    // at least an eighth of the items before culling; every item of the
    // n_unculled views whose items all survive (equirectangular targets are
    // not culled); and, for the others, enough for their share of the entry
    // capacity (a surviving item bins 165-335 entries in the measured scenes,
    // so capacity / 160 items cover a launch set that fills it).  The rest
    // re-evaluate vertices.
    uint64_t keep = items / 8 > 4096 ? items / 8 : 4096;
    const uint64_t need = items / (uint64_t)nviews * (uint64_t)n_unculled +
                          (uint64_t)((double)d.cap * culled_px_frac / 160.0);
    if (need > keep) keep = need;
    // RVS_KEEP_ITEMS=<n> lowers the kept range (less workspace; the items
    // beyond it take the re-evaluating passes, with the same results)
    if (const char* env = std::getenv("RVS_KEEP_ITEMS")) {
        const long long n = std::strtoll(env, nullptr, 10);
        if (n > 0 && (uint64_t)n < keep) keep = (uint64_t)n;
    }
    // a caller-imposed limit (a fixed workspace) lowers it the same way
    if (keep_max > 0 && keep_max < keep) keep = keep_max;
    d.range_cap = (uint32_t)(keep < items ? keep : items);
    const size_t nbins = (size_t)nsrc * ntiles;
    R.status = L.take(16 * sizeof(uint32_t));
    R.counts = L.take(nbins * sizeof(uint32_t));
    R.offsets = L.take(nbins * sizeof(uint32_t));
    R.fill = L.take(nbins * sizeof(uint32_t));
    R.bsum = L.take(rvs_k::scan_blocks(nbins) * sizeof(unsigned long long));
    R.lists = L.take((size_t)d.cap * sizeof(uint32_t));
    R.ranges = L.take((size_t)d.range_cap * 256 * sizeof(uint32_t));
    R.ktab = L.take((size_t)d.range_cap * (1 + rvs_k::kBinTable) * sizeof(uint2));
    d.kept_geom = 1;
    if (const char* env = std::getenv("RVS_KEPT_GEOM")) d.kept_geom = std::strtol(env, nullptr, 10) != 0;
    const size_t kg = d.kept_geom ? d.range_cap : 0;
    R.vbuf = L.take(kg * 3 * rvs_k::kBinVerts * sizeof(float));
    R.tvq = L.take(kg * 256 * sizeof(float2));
    R.imap = L.take((size_t)items * sizeof(uint32_t));
    R.deferred = L.take(npx);
    R.dtiles = L.take((size_t)ntiles * sizeof(uint32_t));
    R.work = L.take((size_t)items * sizeof(uint32_t));
    R.tile_enc = L.take((size_t)ntiles);
    // the in-order replay's per-source results of its first rbuf_tiles tiles
    // (tiles beyond them take the one-block-per-tile replay): a sixty-fourth
    // of the pixels, five floats per source
    const uint64_t tile_px = (uint64_t)1 << (2 * shift);
    uint64_t rt = (uint64_t)npx / (64 * tile_px);
    rt = rt < 16 ? 16 : rt;
    // RVS_REPLAY_TILES=<n> lowers it (0: every deferred tile in order; the
    // results are the same)
    if (const char* env = std::getenv("RVS_REPLAY_TILES")) {
        const long long n = std::strtoll(env, nullptr, 10);
        if (n >= 0 && (uint64_t)n < rt) rt = (uint64_t)n;
    }
    d.rbuf_tiles = (uint32_t)(rt < ntiles ? rt : ntiles);
    R.rbuf = L.take((size_t)d.rbuf_tiles * nsrc * tile_px * 5 * sizeof(float));
}

rvs_k::RasterWS carve_raster(char* base, const RasterPlan& R) {
    rvs_k::RasterWS ws;
    ws.status = (uint32_t*)(base + R.status);
    ws.counts = (uint32_t*)(base + R.counts);
    ws.offsets = (uint32_t*)(base + R.offsets);
    ws.fill = (uint32_t*)(base + R.fill);
    ws.bsum = (unsigned long long*)(base + R.bsum);
    ws.lists = (uint32_t*)(base + R.lists);
    ws.ranges = (uint32_t*)(base + R.ranges);
    ws.ktab = (uint2*)(base + R.ktab);
    ws.vbuf = (float*)(base + R.vbuf);
    ws.tvq = (float2*)(base + R.tvq);
    ws.imap = (uint32_t*)(base + R.imap);
    ws.deferred = (uint8_t*)(base + R.deferred);
    ws.deferred_tiles = (uint32_t*)(base + R.dtiles);
    ws.work = (uint32_t*)(base + R.work);
    ws.tile_enc = (uint8_t*)(base + R.tile_enc);
    ws.rbuf = (float*)(base + R.rbuf);
    return ws;
}

hipError_t reset_raster(const rvs_k::RasterWS& ws, const rvs_k::BatchDims& d, hipStream_t s) {
    hipError_t e = rvs_k::launch_reset_u32(ws.status, 16, s);
    if (e == hipSuccess) e = rvs_k::launch_reset_u32(ws.counts, (size_t)d.nsrc * d.ntiles, s);
    if (e == hipSuccess) e = rvs_k::launch_reset_u32(ws.fill, (size_t)d.nsrc * d.ntiles, s);
    return e;
}

// A view's tile grid and its place in the batch's global tile / pixel numbering.
// cv::resize's INTER_LINEAR taps along one axis (image_writing.cpp:189-190):
// for each destination index d, source index s = floor(f) and weight f - s of
// s + 1, f = (float)((d + 0.5) * scale - 0.5) with scale = src / dst in
// double, clamped to [0, n_src - 1] with weight 0 at the ends; the weight's
// bits in .y.  Evaluated here, on the host, in double as the reference does.
void resize_taps(int n_src, int n_dst, int2* out) {
    const double scale = (double)n_src / (double)n_dst;
    for (int d = 0; d < n_dst; ++d) {
        float f = (float)((d + 0.5) * scale - 0.5);
        int s0 = (int)std::floor(f);
        f -= (float)s0;
        if (s0 < 0) { f = 0.f; s0 = 0; }
        if (s0 >= n_src - 1) { f = 0.f; s0 = n_src - 1; }
        int fb;
        std::memcpy(&fb, &f, sizeof fb);
        out[d] = make_int2(s0, fb);
    }
}

void view_grid(rvs_k::ViewDesc& v, int w, int h, uint32_t& tiles, uint32_t& px, int shift = 4) {
    const int edge = 1 << shift;
    v.w = w; v.h = h;
    v.tile_shift = shift;
    v.tiles_x = (w + edge - 1) / edge;
    v.tiles_y = (h + edge - 1) / edge;
    v.tile_base = tiles;
    v.px_base = px;
    tiles += (uint32_t)(v.tiles_x * v.tiles_y);
    px += (uint32_t)(w * h);
}

// After a raster launch set: report list overflow (synchronises the stream).
int check_raster_status(const rvs_k::RasterWS& ws, const rvs_k::BatchDims& d, uint32_t* needed_cap, hipStream_t s) {
    uint32_t st[4];
    HIP_TRY(hipMemcpyAsync(st, ws.status, sizeof st, hipMemcpyDeviceToHost, s), "status copy");
    HIP_TRY(hipStreamSynchronize(s), "stream sync");
    if (st[3])
        return fail(RVS_E_TIMEOUT, "inpainting: a bounded wait between strip solvers gave up; the launch set's "
                                   "results are not valid");
    if (st[1]) {
        if (needed_cap) *needed_cap = st[2];
        if (st[2] == 0xffffffffu)
            return fail(RVS_E_CAPACITY, "binned entries exceed 2^32 - 1: split the batch");
        return fail(RVS_E_CAPACITY, "binned (quad, tile) entries %u exceed the list capacity %u", st[2], d.cap);
    }
    return RVS_OK;
}

// ------------------------------------------------------------- batches

// Everything a fused launch set over n_views targets and n_src sources needs,
// in workspace order.
struct SynthPlan {
    RasterPlan R;
    size_t views, geoms, bounds, quality, color, inpaint, resized, rtab, total;
    int tile_shift;       // log2 of the launch set's tile edge
    int max_w, max_h;     // largest raster (oversampled by Precision)
    int max_ew, max_eh;   // largest view
    bool any_persp, any_erp, any_rescaled;
};

// Raster size of a view with Precision `rescale` (SynthesizedView.cpp:326-329).
void raster_size(const rvs_cam& c, float rescale, int& w, int& h) {
    w = (int)(0.5f + c.w * rescale);
    h = (int)(0.5f + c.h * rescale);
}

int plan_synth(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views, uint32_t cap, SynthPlan& P,
               std::vector<rvs_k::ViewDesc>* views, uint64_t keep_max = 0) {
    if (!srcs || n_src < 1 || n_src > RVS_MAX_SOURCES || !tgts || n_views < 1 || n_views > rvs_k::kMaxViews)
        return fail(RVS_E_INVALID, "synthesize: bad source or view count");
    int max_sw = 0, max_sh = 0;
    for (int k = 0; k < n_src; ++k) {
        if (!cam_ok(&srcs[k].cam)) return fail(RVS_E_INVALID, "synthesize: source %d has a bad camera", k);
        max_sw = srcs[k].cam.w > max_sw ? srcs[k].cam.w : max_sw;
        max_sh = srcs[k].cam.h > max_sh ? srcs[k].cam.h : max_sh;
    }
    uint64_t tiles = 0, px = 0, rpx = 0, persp_px = 0, rtaps = 0;
    int n_erp = 0;   // equirectangular targets: their work items are not culled
    P.max_w = 0; P.max_h = 0; P.max_ew = 0; P.max_eh = 0; P.any_persp = false; P.any_erp = false;
    P.any_rescaled = false;
    // the launch set's tile edge: the largest any view asks for
    int shift = 4;
    for (int v = 0; v < n_views; ++v) {
        const int t = tgts[v].tile;
        if (t != 0 && t != rvs_k::kTile && t != rvs_k::kTileBig)
            return fail(RVS_E_INVALID, "synthesize: view %d: tile edge %d (0, %d or %d)", v, t, rvs_k::kTile,
                        rvs_k::kTileBig);
        if (t == rvs_k::kTileBig) shift = 5;
    }
    P.tile_shift = shift;
    const int edge = 1 << shift;
    for (int v = 0; v < n_views; ++v) {
        const rvs_target& t = tgts[v];
        if (!cam_ok(&t.cam)) return fail(RVS_E_INVALID, "synthesize: view %d has a bad camera", v);
        if (!(t.rescale > 0.f) || !(t.rescale <= 64.f))
            return fail(RVS_E_INVALID, "synthesize: view %d: Precision %g", v, (double)t.rescale);
        int rw, rh;
        raster_size(t.cam, t.rescale, rw, rh);
        if (rw < 2 || rh < 2) return fail(RVS_E_INVALID, "synthesize: view %d: raster %dx%d", v, rw, rh);
        if (rw > rvs_k::kMaxSourceSide || rh > rvs_k::kMaxSourceSide)
            return fail(RVS_E_UNSUPPORTED, "synthesize: view %d: side above %d", v, rvs_k::kMaxSourceSide);
        P.max_w = rw > P.max_w ? rw : P.max_w;
        P.max_h = rh > P.max_h ? rh : P.max_h;
        P.max_ew = t.cam.w > P.max_ew ? t.cam.w : P.max_ew;
        P.max_eh = t.cam.h > P.max_eh ? t.cam.h : P.max_eh;
        P.any_persp = P.any_persp || t.cam.proj == 0;
        P.any_erp = P.any_erp || t.cam.proj != 0;
        if (t.cam.proj != 0) ++n_erp;
        else persp_px += (uint64_t)rw * rh;
        if (rw != t.cam.w || rh != t.cam.h) {
            P.any_rescaled = true;
            rpx += (uint64_t)t.cam.w * t.cam.h;
            rtaps += (uint64_t)t.cam.w + t.cam.h;
        }
        tiles += (uint64_t)((rw + edge - 1) / edge) * ((rh + edge - 1) / edge);
        px += (uint64_t)rw * rh;
    }
    if (max_sw > rvs_k::kMaxSourceSide || max_sh > rvs_k::kMaxSourceSide)
        return fail(RVS_E_UNSUPPORTED, "synthesize: source side above %d", rvs_k::kMaxSourceSide);
    const uint64_t items = (uint64_t)((max_sw - 1 + rvs_k::kBinQX - 1) / rvs_k::kBinQX) *
                           ((max_sh - 1 + rvs_k::kBinQY - 1) / rvs_k::kBinQY) * n_src * n_views;
    if (px >= (1ull << 31) || tiles * n_src >= (1ull << 31) || items >= (1ull << 32) ||
        (uint64_t)(max_sw - 1) * (max_sh - 1) >= (1ull << 30))
        return fail(RVS_E_CAPACITY, "synthesize: batch too large (%llu pixels); split it",
                    (unsigned long long)px);
    Layout L;
    P.views = L.take((size_t)n_views * sizeof(rvs_k::ViewDesc));
    P.geoms = L.take((size_t)n_views * n_src * sizeof(PairGeom));
    size_t nb = 0;
    for (int k = 0; k < n_src; ++k) nb += align256(rvs_k::src_bounds_floats(srcs[k].cam.w, srcs[k].cam.h) * 4);
    P.bounds = L.take(P.any_persp ? nb : 0);
    P.quality = L.take((size_t)px * sizeof(float));
    // the colour map of views whose caller wants only the encoded planes
    P.color = L.take((size_t)px * 3 * sizeof(float));
    plan_raster(L, P.R, n_src, n_views, (uint32_t)tiles, (uint32_t)px, cap, max_sw, max_sh, keep_max, shift,
                P.any_erp, n_erp, px ? (double)persp_px / (double)px : 1.0);
    P.inpaint = L.take(rvs_k::inpaint_ws_bytes((size_t)px, n_views, P.max_w, P.max_h));
    // oversampled views: their colour resized to the view's size, for the encoder
    P.resized = L.take((size_t)rpx * 3 * sizeof(float));
    P.rtab = L.take((size_t)rtaps * sizeof(int2));   // their resize taps
    P.total = L.o;
    if (views) {
        views->resize(n_views);
        uint32_t tb = 0, pb = 0;
        for (int v = 0; v < n_views; ++v) {
            rvs_k::ViewDesc& d = (*views)[v];
            std::memset(&d, 0, sizeof d);
            int rw, rh;
            raster_size(tgts[v].cam, tgts[v].rescale, rw, rh);
            view_grid(d, rw, rh, tb, pb, shift);
            d.ew = tgts[v].cam.w;
            d.eh = tgts[v].cam.h;
            d.bitdepth = tgts[v].bitdepth;
            d.blend_exp = tgts[v].blending_exp;
            d.blend_inv = 1.0f / tgts[v].blending_exp;
        }
    }
    return RVS_OK;
}

void fill_pair(PairGeom& g, const rvs_source& S, const rvs_target& T) {
    std::memset(&g, 0, sizeof g);
    fill_source(g, &S.cam, S.erp_table);
    g.depth = S.depth;
    g.depth_u8 = S.bitdepth_depth <= 8;
    g.depth_f32 = 0;
    g.dscale = T.depth_scale;
    g.znear = S.znear; g.zfar = S.zfar; g.has_invalid = S.has_invalid;
    g.py = S.y; g.pu = S.u; g.pv = S.v;
    g.color_u8 = S.bitdepth_color <= 8;
    g.cscale = T.color_scale;
    const void* plane = S.prep ? (const char*)S.prep + prep_color_offset(&S.cam) : nullptr;
    g.col = plane && !prep_packed(S) ? (const float4*)plane : nullptr;
    g.colq = plane && prep_packed(S) ? (const ushort4*)plane : nullptr;
    rvs_xform xf;
    rvs_relative_pose(S.pos, S.rot, T.pos, T.rot, &xf);
    std::memcpy(g.R, xf.R, sizeof g.R);
    std::memcpy(g.t, xf.t, sizeof g.t);
    fill_target(g, &T.cam, T.rescale);
}

// Enqueues the fused launch set; returns after the last launch (no sync).
// With a finish stream sf other than s, inpainting and encoding go to sf
// after an event recorded on s behind the in-order replay.
int enqueue_synth(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views, void* workspace,
                  size_t ws_bytes, uint32_t cap, const rvs_view_out* outs, hipStream_t s, SynthPlan& P,
                  rvs_k::RasterWS& ws, hipStream_t sf = nullptr, uint64_t keep_max = 0) {
    std::vector<rvs_k::ViewDesc> views;
    const int prc = plan_synth(srcs, n_src, tgts, n_views, cap, P, &views, keep_max);
    if (prc != RVS_OK) return prc;
    if (!workspace || !outs) return fail(RVS_E_INVALID, "synthesize: null workspace or outputs");
    if (ws_bytes < P.total) return fail(RVS_E_INVALID, "synthesize: workspace %zu < %zu bytes", ws_bytes, P.total);
    for (int k = 0; k < n_src; ++k) {
        const rvs_source& S = srcs[k];
        if (!S.y || !S.u || !S.v || !S.depth || (S.cam.proj == 1 && !S.erp_table))
            return fail(RVS_E_INVALID, "synthesize: source %d incomplete", k);
    }
    char* base = (char*)workspace;
    float* quality = (float*)(base + P.quality);
    float* icolor = (float*)(base + P.color);
    float* resized = (float*)(base + P.resized);
    int2* rtab = (int2*)(base + P.rtab);
    std::vector<int2> taps;   // host copy of the resize taps, uploaded below
    for (int v = 0; v < n_views; ++v) {
        const rvs_view_out& o = outs[v];
        rvs_k::ViewDesc& d = views[v];
        const bool enc = o.y && o.u && o.v;
        if (!o.color && !enc)
            return fail(RVS_E_INVALID, "synthesize: view %d has neither a colour map nor planes", v);
        if (enc && ((d.ew & 1) || (d.eh & 1)))
            return fail(RVS_E_UNSUPPORTED, "synthesize: odd target size cannot be encoded (view %d)", v);
        float* color = o.color ? o.color : icolor + (size_t)d.px_base * 3;
        const bool rescaled = d.w != d.ew || d.h != d.eh;
        if (rescaled) {
            d.ecol = resized;
            resized += (size_t)d.ew * d.eh * 3;
            d.rtab = rtab + taps.size();
            taps.resize(taps.size() + (size_t)d.ew + d.eh);
            resize_taps(d.w, d.ew, taps.data() + taps.size() - d.ew - d.eh);
            resize_taps(d.h, d.eh, taps.data() + taps.size() - d.eh);
        } else {
            d.ecol = color;
        }
        if (enc && (d.bitdepth < 1 || d.bitdepth > 16))
            return fail(RVS_E_INVALID, "synthesize: view %d output bit depth %d", v, d.bitdepth);
        d.color = color;
        d.enc_raster = enc && !rescaled;
        d.partial_color = !o.color && !rescaled;
        d.quality = o.quality ? o.quality : quality + d.px_base;
        d.partial_quality = !o.quality;
        d.depth = o.depth;
        d.validity = o.validity;
        d.y = enc ? o.y : nullptr; d.u = enc ? o.u : nullptr; d.v = enc ? o.v : nullptr;
    }
    std::vector<PairGeom> hg((size_t)n_views * n_src);
    for (int v = 0; v < n_views; ++v)
        for (int k = 0; k < n_src; ++k) fill_pair(hg[(size_t)v * n_src + k], srcs[k], tgts[v]);
    rvs_k::ViewDesc* dviews = (rvs_k::ViewDesc*)(base + P.views);
    PairGeom* dg = (PairGeom*)(base + P.geoms);
    ws = carve_raster(base, P.R);
    const rvs_k::BatchDims& d = P.R.d;
    HIP_TRY(reset_raster(ws, d, s), "reset");
    prof_mark(0, s);
    if (P.any_persp) {
        // culling tables: the source's resident one when it matches the views'
        // depth scale, else one built here; views with another depth scale do not cull
        char* nb = base + P.bounds;
        for (int k = 0; k < n_src; ++k) {
            const rvs_source& S = srcs[k];
            const float ds = tgts[0].depth_scale;
            const float* tab = nullptr;
            if (S.prep && S.prep_depth_scale == ds) {
                tab = S.prep;
            } else {
                float* own = (float*)nb;
                HIP_TRY(rvs_k::launch_src_bounds(hg[k], own, s), "source bounds");
                tab = own;
            }
            nb += align256(rvs_k::src_bounds_floats(S.cam.w, S.cam.h) * 4);
            for (int v = 0; v < n_views; ++v) {
                PairGeom& g = hg[(size_t)v * n_src + k];
                const bool same = tgts[v].depth_scale == ds;
                g.bounds = same ? tab : nullptr;
                g.bounds_bx = rvs_k::bin_blocks_x(S.cam.w);
            }
        }
    }
    HIP_TRY(rvs_k::launch_upload(views.data(), views.size() * sizeof(rvs_k::ViewDesc), dviews, s), "view upload");
    HIP_TRY(rvs_k::launch_upload(hg.data(), hg.size() * sizeof(PairGeom), dg, s), "geometry upload");
    if (!taps.empty())
        HIP_TRY(rvs_k::launch_upload(taps.data(), taps.size() * sizeof(int2), rtab, s), "resize taps upload");
    HIP_TRY(rvs_k::launch_bin<PairGeom>(dg, dviews, d, ws, s), "bin");
    prof_mark(1, s);
    HIP_TRY((rvs_k::launch_raster<PairGeom, rvs_k::MODE_BLEND>(dg, dviews, d, ws, s)), "raster");
    prof_mark(2, s);
    HIP_TRY((rvs_k::launch_replay<PairGeom, rvs_k::MODE_BLEND>(dg, dviews, d, ws, s)), "replay");
    prof_mark(3, s);
    if (sf && sf != s) {
        // a wait captures the event as recorded at that moment, so one event
        // per host thread serves every call
        static thread_local hipEvent_t split_ev = nullptr;
        if (!split_ev) HIP_TRY(hipEventCreateWithFlags(&split_ev, hipEventDisableTiming), "split event");
        HIP_TRY(hipEventRecord(split_ev, s), "split record");
        HIP_TRY(hipStreamWaitEvent(sf, split_ev, 0), "split wait");
        s = sf;
    }
    rvs_k::InpaintWS iws = rvs_k::inpaint_ws_carve(base + P.inpaint, d.npx, n_views, P.max_w, P.max_h);
    iws.fail = ws.status + 3;   // a bounded wait of the inpainting that gave up
    iws.rflags = ws.tile_enc;   // tiles the raster left without holes or deferred pixels are skipped
    HIP_TRY(rvs_k::launch_inpaint(dviews, n_views, P.max_w, P.max_h, iws, nullptr, s), "inpaint");
    prof_mark(4, s);
    if (P.any_rescaled) HIP_TRY(rvs_k::launch_resize_views(dviews, n_views, P.max_ew, P.max_eh, s), "resize");
    HIP_TRY(rvs_k::launch_encode_views(dviews, n_views, P.max_ew, P.max_eh, ws.tile_enc, s), "encode");
    prof_mark(5, s);
    return RVS_OK;
}

// ------------------------------------------------- static device scratch

// Workspace of rvs_synthesize_noalloc: one static array per GPU, loaded with
// the library's code object, so that entry point allocates nothing.  In host
// code `scratch` names a host-side stand-in, never the device memory: every
// host use goes through the device address from hipGetSymbolAddress, taken
// once per device with that device current.
constexpr size_t kNoallocScratchBytes = RVS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(RVS_E_INVALID, "rvs_synthesize_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
        if (!p) return fail(RVS_E_HIP, "scratch address: null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RVS_OK;
}

// Kept work items fitted to the static scratch: the plan with a single kept
// item gives the fixed part; what remains holds this many kept items (their
// tile ranges, bin tables, vertex windows and validities).  0 when even the
// fixed part does not fit.
uint64_t noalloc_keep(const rvs_source* srcs, int n_src, const rvs_target* tgt, uint32_t cap, size_t& need) {
    SynthPlan P;
    if (plan_synth(srcs, n_src, tgt, 1, cap, P, nullptr) != RVS_OK) return 0;
    if (P.total <= kNoallocScratchBytes) { need = P.total; return P.R.d.range_cap ? P.R.d.range_cap : 1; }
    if (plan_synth(srcs, n_src, tgt, 1, cap, P, nullptr, 1) != RVS_OK) return 0;
    need = P.total;
    if (P.total > kNoallocScratchBytes) return 0;
    const size_t per_item = 256 * sizeof(uint32_t) + (1 + rvs_k::kBinTable) * sizeof(uint2) +
                            3 * (size_t)rvs_k::kBinVerts * sizeof(float) + 256 * sizeof(float2);
    const uint64_t keep = 1 + (kNoallocScratchBytes - P.total) / per_item;
    // alignment of the four kept arrays can cost up to 4 x 255 bytes: step back until it fits
    for (uint64_t k = keep; k >= 1; k = k > 64 ? k - 64 : k - 1) {
        if (plan_synth(srcs, n_src, tgt, 1, cap, P, nullptr, k) == RVS_OK && P.total <= kNoallocScratchBytes) {
            need = P.total;
            return k;
        }
        if (k == 1) break;
    }
    return 0;
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

size_t rvs_source_prep_size(const rvs_cam* cam) {
    if (!cam_ok(cam)) return 0;
    return prep_color_offset(cam) + (size_t)cam->w * cam->h * sizeof(float4);
}

int rvs_source_prep(rvs_source* src, float depth_scale, void* out, size_t bytes, rvs_stream_t stream) {
    if (!src || !cam_ok(&src->cam) || !out || !src->depth || !src->y || !src->u || !src->v ||
        (src->cam.proj == 1 && !src->erp_table))
        return fail(RVS_E_INVALID, "rvs_source_prep: bad argument");
    if ((uint64_t)(src->cam.w - 1) * (src->cam.h - 1) >= (1ull << 30))
        return fail(RVS_E_CAPACITY, "rvs_source_prep: source too large");
    if (bytes < rvs_source_prep_size(&src->cam)) return fail(RVS_E_INVALID, "rvs_source_prep: buffer too small");
    PairGeom g;
    std::memset(&g, 0, sizeof g);
    fill_source(g, &src->cam, src->erp_table);
    g.depth = src->depth;
    g.depth_u8 = src->bitdepth_depth <= 8;
    g.dscale = depth_scale;
    g.znear = src->znear; g.zfar = src->zfar; g.has_invalid = src->has_invalid;
    HIP_TRY(rvs_k::launch_src_bounds(g, (float*)out, (hipStream_t)stream), "source bounds");
    HIP_TRY(rvs_k::launch_prep_color(src->y, src->u, src->v, src->bitdepth_color <= 8, src->cam.w, src->cam.h,
                                     prep_packed(*src) ? 1 : 0, (char*)out + prep_color_offset(&src->cam),
                                     (hipStream_t)stream),
            "source colour");
    src->prep = (const float*)out;
    src->prep_depth_scale = depth_scale;
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

// Workspace of the raster stage entry point: one view, one precomputed source.
static size_t raster_plan(int in_w, int in_h, int out_w, int out_h, uint32_t cap, RasterPlan& R, size_t& views,
                          size_t& geom) {
    Layout L;
    views = L.take(sizeof(rvs_k::ViewDesc));
    geom = L.take(sizeof(ArrayGeom));
    const uint32_t ntiles = (uint32_t)(((out_w + rvs_k::kTile - 1) / rvs_k::kTile) *
                                       ((out_h + rvs_k::kTile - 1) / rvs_k::kTile));
    plan_raster(L, R, 1, 1, ntiles, (uint32_t)out_w * (uint32_t)out_h, cap, in_w, in_h);
    return L.o;
}

size_t rvs_raster_workspace_size(int in_w, int in_h, int out_w, int out_h, uint32_t entry_cap) {
    if (in_w < 2 || in_h < 2 || out_w < 1 || out_h < 1) return 0;
    RasterPlan R;
    size_t a, b;
    return raster_plan(in_w, in_h, out_w, out_h, entry_cap, R, a, b);
}

int rvs_raster(const float* color, const float* depth, const float* uv, int in_w, int in_h, int out_w, int out_h,
               int wrap, float rescale, void* workspace, size_t ws_bytes, uint32_t entry_cap, float* out_color,
               float* out_depth, float* out_shape, uint32_t* needed_cap, rvs_stream_t stream) {
    if (!color || !depth || !uv || in_w < 2 || in_h < 2 || out_w < 1 || out_h < 1 || !workspace || !out_color ||
        !out_depth || !out_shape || !(rescale > 0.f))
        return fail(RVS_E_INVALID, "rvs_raster: bad argument");
    if ((uint64_t)out_w * out_h >= (1ull << 31) || (uint64_t)(in_w - 1) * (in_h - 1) >= (1ull << 30) ||
        in_w > rvs_k::kMaxSourceSide || in_h > rvs_k::kMaxSourceSide)
        return fail(RVS_E_CAPACITY, "rvs_raster: image too large");
    RasterPlan R;
    size_t ov, og;
    const size_t need = raster_plan(in_w, in_h, out_w, out_h, entry_cap, R, ov, og);
    if (ws_bytes < need) return fail(RVS_E_INVALID, "rvs_raster: workspace %zu < %zu bytes", ws_bytes, need);
    hipStream_t s = (hipStream_t)stream;
    char* base = (char*)workspace;
    rvs_k::RasterWS ws = carve_raster(base, R);
    rvs_k::ViewDesc vd;
    std::memset(&vd, 0, sizeof vd);
    uint32_t tb = 0, pb = 0;
    view_grid(vd, out_w, out_h, tb, pb);
    vd.color = out_color; vd.depth = out_depth; vd.validity = out_shape;
    ArrayGeom hg;
    hg.sw = in_w; hg.sh = in_h; hg.uv = uv; hg.vd = depth; hg.color = color; hg.wrap = wrap ? 1 : 0;
    hg.rescale = rescale;
    rvs_k::ViewDesc* dv = (rvs_k::ViewDesc*)(base + ov);
    ArrayGeom* dg = (ArrayGeom*)(base + og);
    HIP_TRY(rvs_k::launch_upload(&vd, sizeof vd, dv, s), "view upload");
    HIP_TRY(rvs_k::launch_upload(&hg, sizeof hg, dg, s), "geometry upload");
    HIP_TRY(reset_raster(ws, R.d, s), "reset");
    HIP_TRY(rvs_k::launch_bin<ArrayGeom>(dg, dv, R.d, ws, s), "bin");
    HIP_TRY((rvs_k::launch_raster<ArrayGeom, rvs_k::MODE_RAW>(dg, dv, R.d, ws, s)), "raster");
    HIP_TRY((rvs_k::launch_replay<ArrayGeom, rvs_k::MODE_RAW>(dg, dv, R.d, ws, s)), "replay");
    return check_raster_status(ws, R.d, needed_cap, s);
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

size_t rvs_inpaint_workspace_size(int w, int h) {
    if (w < 1 || h < 1) return 0;
    return align256(sizeof(rvs_k::ViewDesc)) + align256(rvs_k::inpaint_ws_bytes((size_t)w * h, 1, w, h));
}

int rvs_inpaint(float* color, const float* quality, int w, int h, void* workspace, size_t ws_bytes, int32_t* map_out,
                rvs_stream_t stream) {
    if (!color || !quality || w < 1 || h < 1 || !workspace) return fail(RVS_E_INVALID, "rvs_inpaint: bad argument");
    if ((uint64_t)w * h >= (1ull << 31)) return fail(RVS_E_CAPACITY, "rvs_inpaint: image too large");
    if (ws_bytes < rvs_inpaint_workspace_size(w, h)) return fail(RVS_E_INVALID, "rvs_inpaint: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    rvs_k::ViewDesc vd;
    std::memset(&vd, 0, sizeof vd);
    uint32_t tb = 0, pb = 0;
    view_grid(vd, w, h, tb, pb);
    vd.color = color;
    vd.quality = (float*)quality;
    rvs_k::ViewDesc* dv = (rvs_k::ViewDesc*)workspace;
    const rvs_k::InpaintWS ws =
        rvs_k::inpaint_ws_carve((char*)workspace + align256(sizeof(rvs_k::ViewDesc)), (size_t)w * h, 1, w, h);
    HIP_TRY(rvs_k::launch_upload(&vd, sizeof vd, dv, s), "view upload");
    HIP_TRY(rvs_k::launch_inpaint(dv, 1, w, h, ws, map_out, s), "inpaint");
    return RVS_OK;
}

size_t rvs_resize_color_workspace_size(int sw, int sh, int dw, int dh) {
    if (sw < 1 || sh < 1 || dw < 1 || dh < 1 || (uint64_t)dw * dh >= (1ull << 31)) return 0;
    return align256(((size_t)dw + dh) * sizeof(int2));
}

int rvs_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh, void* workspace,
                     size_t ws_bytes, rvs_stream_t stream) {
    if (!src || !dst || !workspace || sw < 1 || sh < 1 || dw < 1 || dh < 1 || (uint64_t)dw * dh >= (1ull << 31))
        return fail(RVS_E_INVALID, "rvs_resize_color: bad argument");
    if (ws_bytes < rvs_resize_color_workspace_size(sw, sh, dw, dh))
        return fail(RVS_E_INVALID, "rvs_resize_color: workspace too small");
    std::vector<int2> taps((size_t)dw + dh);
    resize_taps(sw, dw, taps.data());
    resize_taps(sh, dh, taps.data() + dw);
    hipStream_t s = (hipStream_t)stream;
    int2* dtaps = (int2*)workspace;
    HIP_TRY(rvs_k::launch_upload(taps.data(), taps.size() * sizeof(int2), dtaps, s), "resize taps upload");
    HIP_TRY(rvs_k::launch_resize_color(src, sw, sh, dst, dw, dh, dtaps, s), "resize_color");
    return RVS_OK;
}

size_t rvs_split_frequencies_workspace_size(int w, int h, int rgb) {
    if (w < 1 || h < 1 || (uint64_t)w * h >= (1ull << 31)) return 0;
    return rvs_k::split_ws_bytes((size_t)w * h, rgb);
}

int rvs_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h, int rgb,
                          void* workspace, size_t ws_bytes, float* low, float* high, rvs_stream_t stream) {
    if (!color || !depth || !shape || !low || !high || !workspace || w < 1 || h < 1)
        return fail(RVS_E_INVALID, "rvs_split_frequencies: bad argument");
    if ((uint64_t)w * h >= (1ull << 31)) return fail(RVS_E_CAPACITY, "rvs_split_frequencies: image too large");
    if (ws_bytes < rvs_split_frequencies_workspace_size(w, h, rgb))
        return fail(RVS_E_INVALID, "rvs_split_frequencies: workspace too small");
    HIP_TRY(rvs_k::launch_split_frequencies(color, depth, shape, w, h, rgb, workspace, low, high,
                                            (hipStream_t)stream),
            "split_frequencies");
    return RVS_OK;
}

int rvs_add_color(const float* a, const float* b, float* out, int w, int h, rvs_stream_t stream) {
    if (!a || !b || !out || w < 1 || h < 1) return fail(RVS_E_INVALID, "rvs_add_color: bad argument");
    HIP_TRY(rvs_k::launch_add_color(a, b, out, 3 * (size_t)w * h, (hipStream_t)stream), "add_color");
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

size_t rvs_synth_batch_workspace_size(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                                      uint32_t entry_cap) {
    SynthPlan P;
    return plan_synth(srcs, n_src, tgts, n_views, entry_cap, P, nullptr) == RVS_OK ? P.total : 0;
}

int rvs_synth_plan(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views, uint32_t entry_cap,
                   size_t* ws_bytes) {
    SynthPlan P;
    const int rc = plan_synth(srcs, n_src, tgts, n_views, entry_cap, P, nullptr);
    if (ws_bytes) *ws_bytes = rc == RVS_OK ? P.total : 0;
    return rc;
}

size_t rvs_synth_workspace_size(const rvs_source* srcs, int n_src, const rvs_target* tgt, uint32_t entry_cap) {
    return rvs_synth_batch_workspace_size(srcs, n_src, tgt, 1, entry_cap);
}

int rvs_synthesize_batch_async(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                               void* workspace, size_t ws_bytes, uint32_t entry_cap, const rvs_view_out* outs,
                               rvs_stream_t stream) {
    SynthPlan P;
    rvs_k::RasterWS ws;
    g_prof_valid = false;
    const int rc = enqueue_synth(srcs, n_src, tgts, n_views, workspace, ws_bytes, entry_cap, outs,
                                 (hipStream_t)stream, P, ws);
    if (rc == RVS_OK && g_prof) { g_prof_valid = true; g_prof_ws = ws; g_prof_dims = P.R.d; }
    return rc;
}

int rvs_synthesize_batch_async_split(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                                     void* workspace, size_t ws_bytes, uint32_t entry_cap,
                                     const rvs_view_out* outs, rvs_stream_t stream, rvs_stream_t finish_stream) {
    SynthPlan P;
    rvs_k::RasterWS ws;
    g_prof_valid = false;
    return enqueue_synth(srcs, n_src, tgts, n_views, workspace, ws_bytes, entry_cap, outs, (hipStream_t)stream, P,
                         ws, (hipStream_t)finish_stream);
}

int rvs_synth_check(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views, const void* workspace,
                    uint32_t entry_cap, uint32_t* needed_cap, rvs_stream_t stream) {
    if (!workspace) return fail(RVS_E_INVALID, "rvs_synth_check: null workspace");
    SynthPlan P;
    const int prc = plan_synth(srcs, n_src, tgts, n_views, entry_cap, P, nullptr);
    if (prc != RVS_OK) return prc;
    const rvs_k::RasterWS ws = carve_raster((char*)workspace, P.R);
    return check_raster_status(ws, P.R.d, needed_cap, (hipStream_t)stream);
}

int rvs_synthesize_batch(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views, void* workspace,
                         size_t ws_bytes, uint32_t entry_cap, const rvs_view_out* outs, uint32_t* needed_cap,
                         rvs_stream_t stream) {
    SynthPlan P;
    rvs_k::RasterWS ws;
    g_prof_valid = false;
    const int rc = enqueue_synth(srcs, n_src, tgts, n_views, workspace, ws_bytes, entry_cap, outs,
                                 (hipStream_t)stream, P, ws);
    if (rc != RVS_OK) return rc;
    const int crc = check_raster_status(ws, P.R.d, needed_cap, (hipStream_t)stream);
    if (crc == RVS_OK && g_prof) { g_prof_valid = true; g_prof_ws = ws; g_prof_dims = P.R.d; }
    return crc;
}

int rvs_synthesize(const rvs_source* srcs, int n_src, const rvs_target* tgt, void* workspace, size_t ws_bytes,
                   uint32_t entry_cap, const rvs_view_out* out, uint32_t* needed_cap, rvs_stream_t stream) {
    return rvs_synthesize_batch(srcs, n_src, tgt, 1, workspace, ws_bytes, entry_cap, out, needed_cap, stream);
}

int rvs_synthesize_noalloc(const rvs_source* srcs, int n_src, const rvs_target* tgt, uint32_t entry_cap,
                           const rvs_view_out* out, uint32_t* needed_cap, rvs_stream_t stream) {
    SynthPlan P;
    const int prc = plan_synth(srcs, n_src, tgt, 1, entry_cap, P, nullptr);
    if (prc != RVS_OK) return prc;
    size_t need = 0;
    const uint64_t keep = noalloc_keep(srcs, n_src, tgt, entry_cap, need);
    if (keep == 0)
        return fail(RVS_E_CAPACITY,
                    "rvs_synthesize_noalloc: this view needs at least %zu workspace bytes, more than the %zu-byte "
                    "static scratch (use rvs_synthesize with a caller workspace)",
                    need, kNoallocScratchBytes);
    void* ws_ptr = nullptr;
    const int src = scratch_address(&ws_ptr);
    if (src != RVS_OK) return src;
    rvs_k::RasterWS ws;
    g_prof_valid = false;
    const int rc = enqueue_synth(srcs, n_src, tgt, 1, ws_ptr, kNoallocScratchBytes, entry_cap, out,
                                 (hipStream_t)stream, P, ws, nullptr, keep);
    if (rc != RVS_OK) return rc;
    const int crc = check_raster_status(ws, P.R.d, needed_cap, (hipStream_t)stream);
    if (crc == RVS_OK && g_prof) { g_prof_valid = true; g_prof_ws = ws; g_prof_dims = P.R.d; }
    return crc;
}

int rvs_profile_counters(uint32_t out[RVS_PROF_COUNTERS]) {
    if (!g_prof_valid) return fail(RVS_E_INVALID, "no profiled rvs_synthesize on this thread");
    uint32_t st[8];
    HIP_TRY(hipMemcpy(st, g_prof_ws.status, sizeof st, hipMemcpyDeviceToHost), "status copy");
    std::vector<uint32_t> counts((size_t)g_prof_dims.nsrc * g_prof_dims.ntiles);
    HIP_TRY(hipMemcpy(counts.data(), g_prof_ws.counts, counts.size() * 4, hipMemcpyDeviceToHost), "counts copy");
    uint32_t mx = 0;
    uint64_t tot = 0;
    for (uint32_t c : counts) { mx = c > mx ? c : mx; tot += c; }
    out[0] = st[0]; out[1] = st[1]; out[2] = mx; out[3] = (uint32_t)(tot > 0xffffffffull ? 0xffffffffu : tot);
    out[4] = st[4];
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
