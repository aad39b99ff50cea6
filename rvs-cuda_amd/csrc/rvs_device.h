// rvs_device.h -- device-side building blocks of the per-frame synthesis path.
//
// Vertex stage (decode depth -> unproject -> affine -> project -> uv scaling),
// triangle setup, per-pixel barycentric test and the blend step.  Each
// function restates one reference CPU function with identical float
// operations (compiled with -ffp-contract=off):
//   decode_depth      process_input_frame.cu:358-381 (nvcc contracts to one fma)
//   vertex            EquirectangularUnprojector.cpp:55-85, PerspectiveUnprojector.cpp:59-86,
//                     SynthesizedView.cpp:68-79, EquirectangularProjector.cpp:60-99,
//                     PerspectiveProjector.cpp:59-87, SynthesizedView.cpp:326-331
//   tri_setup         transform.cpp:62-96 (valid_tri, den cull, bbox)
//   tri_pixel         transform.cpp:105-117
//   fold_step         transform.cpp:119-142
//   blend_step        BlendedView.cpp:180-212 + blending.cpp:293-369
#ifndef RVS_DEVICE_H
#define RVS_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_math.h"

namespace rvs_dev {

using namespace rvs_math;

// ---------------------------------------------------------------- parameters

// One (source view, target view) pair, everything the vertex stage needs.
// Constants are computed on the host with the reference's own float
// expressions so that they carry the reference's bits.
struct PairGeom {
    // source grid
    int sw, sh;
    int src_erp;                 // 1: ERP source, 0: perspective
    float sf0, sf1, sp0, sp1;    // perspective source intrinsics
    const float* tab;            // ERP source: cos(phi)[sw], sin(phi)[sw], cos(theta)[sh], sin(theta)[sh]
    // raw depth plane (Y plane of the depth file) and its decode parameters
    const void* depth;
    int depth_u8;
    int depth_f32;               // depth already decoded (float plane): the reproject stage API
    float dscale, znear, zfar;
    int has_invalid;
    // raw colour planes
    const void *py, *pu, *pv;
    int color_u8;
    float cscale;
    // relative pose (SpaceTransformer::get_rotation / get_translation)
    float R[9];
    float t[3];
    // target projection
    int dst_erp;
    float df0n, df1n, dp0, dp1;  // perspective target: -f0, -f1, p0, p1
    float u0, v0, du_dphi, dv_dtheta;  // ERP target constants
    float sx, sy;                // cv::transform scale (output size / virtual size)
    int wrap;                    // horizontal wrapping (ERP target with full range)
    float rescale;               // g_rescale (Precision)
    // bounding boxes of the unprojected source points per binning block
    // (6 floats each, bounds_bx blocks per row) for frustum culling; null = no culling
    const float* bounds;
    int bounds_bx;
    // resident decoded colour per source pixel (color_raw: Y, U, V), or null:
    // colq holds Y, 4 U, 4 V as integers (8 B per pixel) for even source sizes
    // and bit depths <= 14, where the upsampled chroma is a multiple of 1/4;
    // col holds the floats (16 B per pixel) otherwise.  At most one is set.
    const float4* col;
    const ushort4* colq;
};

// Precomputed-array vertex source (the raster entry point on identical inputs).
struct ArrayGeom {
    int sw, sh;
    const float* uv;     // sw*sh*2
    const float* vd;     // sw*sh
    const float* color;  // sw*sh*3
    int wrap;
    float rescale;
};

// ------------------------------------------------------------ vertex stage

__device__ inline float decode_depth_raw(uint32_t raw, float scale, float znear, float zfar, int has_invalid) {
    float depth = (float)raw * scale;
    if (zfar >= 1000.f)
        depth = znear / depth;
    else
        depth = zfar * znear / __builtin_fmaf(depth, (zfar - znear), znear);
    if (has_invalid && depth == 0) depth = __builtin_nanf("");
    return depth;
}

__device__ inline uint32_t load_plane(const void* p, int u8, size_t idx) {
    return u8 ? (uint32_t)((const uint8_t*)p)[idx] : (uint32_t)((const uint16_t*)p)[idx];
}

// Source pixel (i = row, j = column) unprojected into the source camera frame
// (decode depth, then EquirectangularUnprojector / PerspectiveUnprojector).
__device__ inline float3 unproject_src(const PairGeom& g, int i, int j) {
    const size_t idx = (size_t)i * g.sw + j;
    const float d = g.depth_f32 ? ((const float*)g.depth)[idx]
                                : decode_depth_raw(load_plane(g.depth, g.depth_u8, idx), g.dscale, g.znear,
                                                   g.zfar, g.has_invalid);
    float x, y, z;
    if (g.src_erp) {
        const float cp = g.tab[j], sp = g.tab[g.sw + j];
        const float ct = g.tab[2 * g.sw + i], st = g.tab[2 * g.sw + g.sh + i];
        const float a = ct * cp;
        const float b = ct * sp;
        x = a * d;
        y = b * d;
        z = st * d;
    } else {
        if (d > 0.f) {
            const float uu = (float)j + 0.5f, vv = (float)i + 0.5f;
            x = d;
            y = -(d / g.sf0) * (uu - g.sp0);
            z = -(d / g.sf1) * (vv - g.sp1);
        } else {
            x = y = z = __builtin_nanf("");
        }
    }
    return make_float3(x, y, z);
}

// Source-frame point -> target u', v' (after the uv rescale) and target depth.
// With `und`, the radius of an equirectangular target takes only the fast
// float-float path and *und is set when it leaves the radius undecided (the
// caller then defers what depends on this vertex); without, it is exact.
__device__ inline float3 project_tgt(const PairGeom& g, float3 p, bool* und = nullptr) {
    float X = 0.f, Y = 0.f, Z = 0.f;
    X += g.R[0] * p.x; X += g.R[1] * p.y; X += g.R[2] * p.z;
    Y += g.R[3] * p.x; Y += g.R[4] * p.y; Y += g.R[5] * p.z;
    Z += g.R[6] * p.x; Z += g.R[7] * p.y; Z += g.R[8] * p.z;
    X = X + g.t[0];
    Y = Y + g.t[1];
    Z = Z + g.t[2];
    float u, v, vd;
    if (g.dst_erp) {
        float radius;
        if (und) {
            if (!norm_sq_sum_fast(X, Y, Z, radius)) *und = true;
        } else {
            radius = norm_sq_sum_d(X, Y, Z);
        }
        vd = radius;
        const float phi = atan2f(Y, X);
        const float theta = asinf(Z / radius);
        u = g.u0 + g.du_dphi * phi;
        v = g.v0 + g.dv_dtheta * theta;
    } else {
        if (X > 0.f) {
            u = g.df0n * Y / X + g.dp0;
            v = g.df1n * Z / X + g.dp1;
            vd = X;
        } else {
            u = v = vd = __builtin_nanf("");
        }
    }
    return make_float3(g.sx * u + 0.f, g.sy * v + 0.f, vd);
}

// u', v', target depth of source pixel (i = row, j = column); `und` as in
// project_tgt.
__device__ inline float3 vertex(const PairGeom& g, int i, int j, bool* und = nullptr) {
    return project_tgt(g, unproject_src(g, i, j), und);
}

__device__ inline float3 vertex(const ArrayGeom& g, int i, int j, bool* = nullptr) {
    const size_t idx = (size_t)i * g.sw + j;
    return make_float3(g.uv[2 * idx], g.uv[2 * idx + 1], g.vd[idx]);
}

// Decoded colour of a source pixel before the colour scale: Y, and U,V
// upsampled bilinearly with src = x * (in/out) and no half-pixel offset
// (process_input_frame.cu:22-81, 211-238).  Weights are 0, 1/2 or 1 for the
// even sizes used, so the products are exact; the nvcc contraction is
// restated as fmaf for the general case.
__device__ inline float3 color_raw(const PairGeom& g, int i, int j) {
    const int inW = g.sw / 2, inH = g.sh / 2;
    float up[2];
    const void* planes[2] = {g.pu, g.pv};
    if (((g.sw | g.sh) & 1) == 0) {
        // Even sizes: scale = 1/2 exactly, so srcX = j/2, x1 = j >> 1, dx in {0, 1/2}.
        // fmaf(a, 1 - dx, b * dx) is then a (dx = 0) or (a + b) / 2, exact for
        // integer samples < 2^16 (and likewise for the vertical pass): the
        // same bits as the general formula below, without its divisions.
        const int x1 = j >> 1, y1 = i >> 1;
        const int x2 = min(x1 + 1, inW - 1), y2 = min(y1 + 1, inH - 1);
        const bool hx = (j & 1) != 0, hy = (i & 1) != 0;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const float a = (float)load_plane(planes[p], g.color_u8, (size_t)y1 * inW + x1);
            float val1 = a, val2 = 0.f;
            if (hx) val1 = (a + (float)load_plane(planes[p], g.color_u8, (size_t)y1 * inW + x2)) * 0.5f;
            if (hy) {
                const float c = (float)load_plane(planes[p], g.color_u8, (size_t)y2 * inW + x1);
                val2 = hx ? (c + (float)load_plane(planes[p], g.color_u8, (size_t)y2 * inW + x2)) * 0.5f : c;
                up[p] = (val1 + val2) * 0.5f;
            } else {
                up[p] = val1;
            }
        }
    } else {
        const float scaleX = (float)inW / g.sw, scaleY = (float)inH / g.sh;
        const float srcX = j * scaleX, srcY = i * scaleY;
        const int x1 = (int)srcX, y1 = (int)srcY;
        const int x2 = (int)fminf((float)(x1 + 1), (float)(inW - 1));
        const int y2 = (int)fminf((float)(y1 + 1), (float)(inH - 1));
        const float dx = srcX - x1, dy = srcY - y1;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const float a = (float)load_plane(planes[p], g.color_u8, (size_t)y1 * inW + x1);
            const float b = (float)load_plane(planes[p], g.color_u8, (size_t)y1 * inW + x2);
            const float c = (float)load_plane(planes[p], g.color_u8, (size_t)y2 * inW + x1);
            const float e = (float)load_plane(planes[p], g.color_u8, (size_t)y2 * inW + x2);
            const float val1 = __builtin_fmaf(a, (1 - dx), b * dx);
            const float val2 = __builtin_fmaf(c, (1 - dx), e * dx);
            up[p] = __builtin_fmaf(val1, (1 - dy), val2 * dy);
        }
    }
    const float yy = (float)load_plane(g.py, g.color_u8, (size_t)i * g.sw + j);
    return make_float3(yy, up[0], up[1]);
}

// Decoded colour times the colour scale (the same product either way, from
// the resident decoded plane when the source has one).
__device__ inline float3 color(const PairGeom& g, int i, int j) {
    if (g.colq) {
        const ushort4 q = g.colq[(size_t)i * g.sw + j];
        return make_float3((float)q.x * g.cscale, ((float)q.y * 0.25f) * g.cscale, ((float)q.z * 0.25f) * g.cscale);
    }
    if (g.col) {
        const float4 r = g.col[(size_t)i * g.sw + j];
        return make_float3(r.x * g.cscale, r.y * g.cscale, r.z * g.cscale);
    }
    const float3 r = color_raw(g, i, j);
    return make_float3(r.x * g.cscale, r.y * g.cscale, r.z * g.cscale);
}

// Colours of three source pixels (a triangle's corners), the same values as
// three color() calls: the resident plane's three loads are issued together
// and then converted, so that a resolve waits on memory once, not per corner.
__device__ inline void color3(const PairGeom& g, const int r[3], const int c[3], float3 out[3]) {
    if (g.colq) {
        ushort4 q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = g.colq[(size_t)r[k] * g.sw + c[k]];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            out[k] = make_float3((float)q[k].x * g.cscale, ((float)q[k].y * 0.25f) * g.cscale,
                                 ((float)q[k].z * 0.25f) * g.cscale);
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = color(g, r[k], c[k]);
}

__device__ inline float3 color(const ArrayGeom& g, int i, int j) {
    const size_t idx = (size_t)i * g.sw + j;
    return make_float3(g.color[3 * idx], g.color[3 * idx + 1], g.color[3 * idx + 2]);
}

__device__ inline void color3(const ArrayGeom& g, const int r[3], const int c[3], float3 out[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = color(g, r[k], c[k]);
}

// ------------------------------------------------------------ triangles

// Slots of one source quad (i, j) in the reference's traversal order
// (transform.cpp:159-176): 0 = T1, 1 = T2, 2/3 = seam triangles (j == 0, wrap).
// Vertex (row, col) for slot s, corner k.
__device__ inline void slot_vertex(int slot, int k, int i, int j, int w, int& r, int& c) {
    // T1: (j,i) (j+1,i) (j,i+1)      T2: (j+1,i+1) (j,i+1) (j+1,i)
    // S1: (w-1,i) (0,i) (w-1,i+1)    S2: (0,i+1) (w-1,i+1) (0,i)
    switch (slot) {
    case 0: r = (k == 2) ? i + 1 : i;           c = (k == 1) ? j + 1 : j; break;
    case 1: r = (k == 2) ? i : i + 1;           c = (k == 1) ? j : j + 1; break;
    case 2: r = (k == 2) ? i + 1 : i;           c = (k == 1) ? 0 : w - 1; break;
    default: r = (k == 0) ? i + 1 : (k == 1 ? i + 1 : i); c = (k == 1) ? w - 1 : 0; break;
    }
}

struct Tri {
    float A0, A1, B0, B1, C0, C1;
    float dA, dB, dC;
    float den, tv;
    int xmin, xmax, ymin, ymax;
};

// Corners, vertex depths and the barycentric denominator (transform.cpp:80-85).
__device__ inline void tri_load(Tri& T, float3 a, float3 b, float3 c) {
    T.A0 = a.x; T.A1 = a.y; T.B0 = b.x; T.B1 = b.y; T.C0 = c.x; T.C1 = c.y;
    T.dA = a.z; T.dB = b.z; T.dC = c.z;
    T.den = (T.B1 - T.C1) * (T.A0 - T.C0) + (T.C0 - T.B0) * (T.A1 - T.C1);
}

// valid_tri (transform.cpp:62-75): stretch = (float) max of three double
// norms, / g_rescale; validity = clamp(10000 - 1000 * stretch, 1, 10000).  The
// clamp's max(1, q) maps NaN to 1, so the result is never 0 and
// colorize_triangle's `validity == 0` return is never taken.  The double norms
// are computed exactly in float arithmetic (tri_stretch, rvs_math.h): no FP64
// instruction on the device.
__device__ inline float tri_validity(const Tri& T, float rescale) {
    return tri_quality(tri_stretch(T.A0, T.A1, T.B0, T.B1, T.C0, T.C1), rescale);
}

// The same when the stretch is decided by the fast path; false otherwise
// (non-finite or out-of-range edges, or a square root within 2^-44 of a float
// rounding midpoint): the caller then defers what depends on it.
__device__ inline bool tri_validity_fast(const Tri& T, float rescale, float& tv) {
    float s;
    if (!tri_stretch_fast(T.A0, T.A1, T.B0, T.B1, T.C0, T.C1, s)) return false;
    tv = tri_quality(s, rescale);
    return true;
}

// colorize_triangle's early returns (transform.cpp:80-96) without the
// validity: den <= 0, empty clipped bounding box.  T.tv is left 0.
__device__ inline bool tri_setup_geom(Tri& T, float3 a, float3 b, float3 c, int out_w, int out_h) {
    tri_load(T, a, b, c);
    T.tv = 0.f;
    if (T.den <= 0.f) return false;
    T.xmin = imax(0, cvt_i32(__builtin_floorf(smin(smin(T.A0, T.B0), T.C0))));
    T.ymin = imax(0, cvt_i32(__builtin_floorf(smin(smin(T.A1, T.B1), T.C1))));
    T.xmax = imin(out_w - 1, cvt_i32(__builtin_ceilf(smax(smax(T.A0, T.B0), T.C0))));
    T.ymax = imin(out_h - 1, cvt_i32(__builtin_ceilf(smax(smax(T.A1, T.B1), T.C1))));
    return !(T.ymin > T.ymax || T.xmin > T.xmax);
}

// Returns false when colorize_triangle returns before its pixel loop.
__device__ inline bool tri_setup(Tri& T, float3 a, float3 b, float3 c, int out_w, int out_h, float rescale) {
    if (!tri_setup_geom(T, a, b, c, out_w, out_h)) return false;
    T.tv = tri_validity(T, rescale);
    return true;
}

// Narrows [xa, xb] x [ya, yb] to the pixels whose centre lies within
// dl = 0.07 + 4e-6 s of the triangle's extent, s = max(W, H) + 2; no other
// pixel can pass tri_cover.  Proof sketch (the box starts at column / row >= 0,
// so the clamps below never cut it): a centre more than dl outside the
// x-extent W has an exact barycentric coordinate below -dl/(2W) (its x offset
// from the extent is the barycentric mean of the corners' offsets, at most two
// of them negative and none below -W); the kernel's l_k differ from the exact
// ones by at most 2^-19 (W+2)(H+2)/den (rounding of the differences, products,
// sum and division), which the guard den >= 2^-14 s^3 keeps below 2^-5/s, so
// the computed one stays below -dl/(2W) + 2^-5/s < -1e-6 (likewise in y):
// dl > 2W (1e-6 + 2^-5/s) holds with 0.0075 to spare, more than the rounding
// of the bounds below for coordinates under 2^16.  Thinner triangles keep the
// box.  The binning passes narrow each triangle's
// box the same way before picking its tiles.
__device__ inline void tri_tight(const Tri& T, int& xa, int& xb, int& ya, int& yb) {
    const float minx = smin(smin(T.A0, T.B0), T.C0), maxx = smax(smax(T.A0, T.B0), T.C0);
    const float miny = smin(smin(T.A1, T.B1), T.C1), maxy = smax(smax(T.A1, T.B1), T.C1);
    const float sz = fmaxf(maxx - minx, maxy - miny) + 2.f;
    if (!(T.den >= 6.103515625e-05f * sz * sz * sz)) return;   // 2^-14; NaN or huge extents keep the box
    const float dl = 0.07f + 4e-6f * sz;
    xa = imax(xa, cvt_i32(__builtin_ceilf(fmaxf(minx - 0.5f - dl, -2.f))));
    xb = imin(xb, cvt_i32(__builtin_floorf(fminf(maxx - 0.5f + dl, 1e9f))));
    ya = imax(ya, cvt_i32(__builtin_ceilf(fmaxf(miny - 0.5f - dl, -2.f))));
    yb = imin(yb, cvt_i32(__builtin_floorf(fminf(maxy - 0.5f + dl, 1e9f))));
}

// Numerators of the barycentric coordinates l1, l2 at pixel (x, y).
__device__ inline void tri_num(const Tri& T, int x, int y, float& n1, float& n2) {
    const float offset = 0.5f;
    const float px = (float)x + offset, py = (float)y + offset;
    n1 = (T.B1 - T.C1) * (px - T.C0) + (T.C0 - T.B0) * (py - T.C1);
    n2 = (T.C1 - T.A1) * (px - T.C0) + (T.A0 - T.C0) * (py - T.C1);
}

// Coverage from the numerators (transform.cpp:105-117).
__device__ inline bool tri_cover(const Tri& T, float n1, float n2, float& l1, float& l2, float& l3) {
    l1 = n1 / T.den;
    l2 = n2 / T.den;
    l3 = 1.f - l1 - l2;
    const float eps = 1e-6f;
    return l1 >= -eps && l2 >= -eps && l3 >= -eps;
}

// Barycentric coordinates at pixel (x, y); returns coverage.
__device__ inline bool tri_pixel(const Tri& T, int x, int y, float& l1, float& l2, float& l3) {
    float n1, n2;
    tri_num(T, x, y, n1, n2);
    return tri_cover(T, n1, n2, l1, l2, l3);
}

// Division-free rejection of pixels that tri_cover cannot accept (den > 0):
// n < fl(-2e-6 * den) gives n / den < -1e-6 after rounding (also when the
// bound is subnormal or 0: then n <= -2^-149 and den < 2^-130), and with
// n1, n2 >= 0 and a normal den, n1 + n2 > fl(1.00001 * den) gives
// l1 + l2 > 1 + 9e-6 after rounding, so l3 < -1e-6.  A NaN or infinite den
// never rejects.
__device__ inline bool tri_surely_outside(const Tri& T, float n1, float n2) {
    const float t = -2e-6f * T.den;
    if (n1 < t || n2 < t) return true;
    return n1 >= 0.f && n2 >= 0.f && T.den >= 1e-30f && n1 + n2 > T.den * 1.00001f;
}

__device__ inline float tri_depth(const Tri& T, float l1, float l2, float l3) {
    return T.dA * l1 + T.dB * l2 + T.dC * l3;
}

__device__ inline float3 tri_color(float3 cA, float3 cB, float3 cC, float l1, float l2, float l3) {
    return make_float3(cA.x * l1 + cB.x * l2 + cC.x * l3,
                       cA.y * l1 + cB.y * l2 + cC.y * l3,
                       cA.z * l1 + cB.z * l2 + cC.z * l3);
}

// A candidate whose score v/d^3 is computed without overflow, underflow or
// NaN: the fold (ratio^3 * shape < validity) then orders it like v/d^3, up to
// a relative error far below kAmbiguity.
__device__ inline bool depth_regular(float d) {
    return d >= 9.5367431640625e-07f && d <= 1048576.0f;   // [2^-20, 2^20]
}

// Relative score band inside which two candidates are replayed in order.
#define RVS_AMBIGUITY 3.814697265625e-06f   /* 2^-18 */

// Per-pixel state of the sequential fold (transform.cpp:119-142).
struct FoldState {
    float new_d;
    float shape;
    int new_prol;
    uint32_t ij;     // last accepted candidate: packed source quad (row << 16 | column)
    int slot;        // and its slot (-1 = none accepted)
};

__device__ inline void fold_init(FoldState& s) {
    s.new_d = __builtin_inff();
    s.shape = 0.f;
    s.new_prol = 1;
    s.ij = 0u;
    s.slot = -1;
}

__device__ inline void fold_step(FoldState& s, float d, float tv, bool prol, uint32_t ij, int slot) {
    const float ratio = d / s.new_d;
    const bool is_valid = ratio * ratio * ratio * s.shape < tv;
    if ((is_valid || s.new_prol) && !prol) {
        s.new_d = d; s.new_prol = 0; s.shape = tv; s.ij = ij; s.slot = slot;
    } else if (is_valid && s.new_prol && prol) {
        s.new_d = d; s.shape = tv; s.ij = ij; s.slot = slot;
    }
}

// ---------------------------------------------------------------- blending

struct BlendState {
    float3 c;
    float q, d, v;
    int mask;   // m_depth_mask (0 or nonzero)
};

__device__ inline float rvs_pow(float x, float a) {
    return a == 1.0f ? x : powf(x, a);
}

// First view: assign(color, depth, quality = shape/depth, validity = shape) and
// m_depth_mask = depth > 0 (BlendedView.cpp:181-186, SynthesizedView.cpp:440).
__device__ inline void blend_first(BlendState& s, float3 c, float d, float shape) {
    s.c = c;
    s.d = d;
    s.q = shape / d;
    s.v = shape;
    s.mask = d > 0.f;
}

// a_inv = 1.0f / a, computed once per view on the host (the same float division).
__device__ inline void blend_step(BlendState& s, float3 c, float d, float shape, float a, float a_inv) {
    const float q_new = shape / d;
    const float q_acc = s.q;
    const bool m_new = d > 0.f;
    float3 out;
    float outQ;
    int outMask;
    if (a < 0) {
        out = make_float3(0.f, 1.f, 0.f);
        outQ = 0.f;
        outMask = 0;
        if (q_acc > outQ) { out = s.c; outQ = q_acc; outMask = s.mask; }
        if (q_new > outQ) { out = c; outQ = q_new; outMask = m_new; }
    } else {
        float sumW = 0.0f, inpSum = 0.0f;
        float3 fin = make_float3(0.f, 0.f, 0.f), inp = make_float3(0.f, 0.f, 0.f);
        // view 0: the accumulator, view 1: the new view
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float qi = i == 0 ? q_acc : q_new;
            const bool mi = i == 0 ? (s.mask != 0) : m_new;
            const float3 ci = i == 0 ? s.c : c;
            float quality = qi;
            const bool isValidDepth = qi > 0 && !mi;
            if (quality > 0 && isValidDepth) {
                quality = rvs_pow(quality, a);
                sumW += quality;
                fin.x += ci.x * quality; fin.y += ci.y * quality; fin.z += ci.z * quality;
            } else if (quality > 0 && qi > 0) {
                quality = rvs_pow(quality, a);
                inpSum += quality;
                inp.x += ci.x * quality; inp.y += ci.y * quality; inp.z += ci.z * quality;
            }
        }
        if (sumW == 0) {
            if (inpSum == 0.0f) {
                out = make_float3(0.f, 1.f, 0.f);
                outMask = 1;
                outQ = 0.0f;
            } else {
                const float r = 1.f / inpSum;
                out = make_float3(inp.x * r, inp.y * r, inp.z * r);
                outMask = 1;
                outQ = rvs_pow(inpSum, a_inv);
            }
        } else {
            const float r = 1.f / sumW;
            out = make_float3(fin.x * r, fin.y * r, fin.z * r);
            outMask = 0;
            outQ = rvs_pow(sumW, a_inv);
        }
    }
    if (q_new > q_acc) s.d = d;
    s.v = (s.v < shape) ? shape : s.v;
    s.c = out;
    s.q = outQ;
    s.mask = outMask;
}

// ------------------------------------------------------------ resize

// write_color's cv::resize(color, color, virtualSize) (image_writing.cpp:
// 189-190), OpenCV's INTER_LINEAR on a float image, at destination pixel
// (x, y): an exact 2x downscale takes OpenCV's area path, the 2x2 block sum
// left to right, top to bottom, times 0.25; any other size the half-pixel
// centred bilinear rule -- f = (float)((d + 0.5) * (src / dst) - 0.5) in
// double, s = floor(f), f -= s, clamped to [0, n - 1] with f = 0 at the ends,
// weights (1 - f, f), rows first.  OpenCV's own operation order is not pinned
// here.  The taps (s, f bits) of every destination column and row come from
// tables the host computes (resize_taps): the double expression has no
// float-only equivalent that is exact in every case (a value on a float
// rounding midpoint rounds by the double's own error).
__device__ inline float3 resize_linear_px(const float* src, int sw, int sh, int dw, int dh, int x, int y,
                                          const int2* tx, const int2* ty) {
    auto at = [&](int r, int c) {
        const float* p = src + 3 * ((size_t)r * sw + c);
        return make_float3(p[0], p[1], p[2]);
    };
    if (sw == 2 * dw && sh == 2 * dh) {
        const float3 a = at(2 * y, 2 * x), b = at(2 * y, 2 * x + 1), c = at(2 * y + 1, 2 * x), d = at(2 * y + 1, 2 * x + 1);
        return make_float3((((a.x + b.x) + c.x) + d.x) * 0.25f, (((a.y + b.y) + c.y) + d.y) * 0.25f,
                           (((a.z + b.z) + c.z) + d.z) * 0.25f);
    }
    const int2 ax = tx[x], ay = ty[y];
    const int x0 = ax.x, y0 = ay.x;
    const int x1 = x0 + 1 < sw ? x0 + 1 : x0, y1 = y0 + 1 < sh ? y0 + 1 : y0;
    const float fx = __int_as_float(ax.y), fy = __int_as_float(ay.y);
    const float gx = 1.f - fx, gy = 1.f - fy;
    const float3 a = at(y0, x0), b = at(y0, x1), c = at(y1, x0), d = at(y1, x1);
    const float3 r0 = make_float3(a.x * gx + b.x * fx, a.y * gx + b.y * fx, a.z * gx + b.z * fx);
    const float3 r1 = make_float3(c.x * gx + d.x * fx, c.y * gx + d.y * fx, c.z * gx + d.z * fx);
    return make_float3(r0.x * gy + r1.x * fy, r0.y * gy + r1.y * fy, r0.z * gy + r1.z * fy);
}

}  // namespace rvs_dev

#endif
