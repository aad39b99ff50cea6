// rvs_stages.hip -- element-wise stages: decode, reproject, blend step, encode.
#include <hip/hip_runtime.h>

#include "rvs_stages.h"

namespace rvs_k {

using namespace rvs_dev;

// read_color_YUV -> resizeImage + mergeFrame (image_loading.cpp:74-145).
__global__ void k_decode_color(const void* y, const void* u, const void* v, int u8, int w, int h, float cs,
                               float* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (x >= w || r >= h) return;
    PairGeom g;
    g.sw = w; g.sh = h; g.py = y; g.pu = u; g.pv = v; g.color_u8 = u8; g.cscale = cs; g.col = nullptr;
    g.colq = nullptr;
    const float3 c = color(g, r, x);
    const size_t o = (size_t)r * w + x;
    out[3 * o] = c.x; out[3 * o + 1] = c.y; out[3 * o + 2] = c.z;
}

// Resident decoded colour plane of a source (color_raw per pixel): floats, or
// (packed) Y, 4 U, 4 V as 16-bit integers -- exact for even sizes and codes
// below 2^14, whose upsampled chroma is a multiple of 1/4.
__global__ void k_prep_color(const void* y, const void* u, const void* v, int u8, int w, int h, int packed,
                             void* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (x >= w || r >= h) return;
    PairGeom g;
    g.sw = w; g.sh = h; g.py = y; g.pu = u; g.pv = v; g.color_u8 = u8; g.col = nullptr; g.colq = nullptr;
    const float3 c = color_raw(g, r, x);
    const size_t o = (size_t)r * w + x;
    if (packed)
        ((ushort4*)out)[o] = make_ushort4((unsigned short)c.x, (unsigned short)(c.y * 4.f),
                                          (unsigned short)(c.z * 4.f), 0);
    else
        ((float4*)out)[o] = make_float4(c.x, c.y, c.z, 0.f);
}

// read_depth_YUV -> formatDepth (image_loading.cpp:147-213).
__global__ void k_decode_depth(const void* raw, int u8, size_t n, float scale, float znear, float zfar,
                               int has_invalid, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = decode_depth_raw(load_plane(raw, u8, i), scale, znear, zfar, has_invalid);
}

// SynthesizedView::computeBasicMode stage (a) (SynthesizedView.cpp:315-331).
__global__ void k_reproject(PairGeom g, float* __restrict__ uv, float* __restrict__ vd) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= g.sw || i >= g.sh) return;
    const float3 r = vertex(g, i, j);
    const size_t o = (size_t)i * g.sw + j;
    uv[2 * o] = r.x; uv[2 * o + 1] = r.y;
    vd[o] = r.z;
}

// BlendedViewSimple::blendBasicMode on whole maps.
__global__ void k_blend_step(float* acc_c, float* acc_q, float* acc_d, float* acc_v, uint8_t* acc_m,
                             const float* __restrict__ c, const float* __restrict__ d,
                             const float* __restrict__ shape, size_t n, float a, float a_inv, int first) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 cv = make_float3(c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    BlendState s;
    if (first) {
        blend_first(s, cv, d[i], shape[i]);
        s.mask = s.mask ? 255 : 0;
    } else {
        s.c = make_float3(acc_c[3 * i], acc_c[3 * i + 1], acc_c[3 * i + 2]);
        s.q = acc_q[i]; s.d = acc_d[i]; s.v = acc_v[i]; s.mask = acc_m[i];
        blend_step(s, cv, d[i], shape[i], a, a_inv);
    }
    acc_c[3 * i] = s.c.x; acc_c[3 * i + 1] = s.c.y; acc_c[3 * i + 2] = s.c.z;
    acc_q[i] = s.q; acc_d[i] = s.d; acc_v[i] = s.v;
    acc_m[i] = (uint8_t)(s.mask ? (s.mask == 1 ? 1 : 255) : 0);
}

// write_color (image_writing.cpp:142-203) for the 2x2 block (bx, by).
template <typename T>
__device__ inline void encode_block(const float* __restrict__ color, int w, int bx, int by, unsigned maxv,
                                    T* __restrict__ Y, T* __restrict__ U, T* __restrict__ V) {
    const int hw = w / 2;
    uint32_t su = 0, sv = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int x = 2 * bx + (k & 1), y = 2 * by + (k >> 1);
        const size_t o = (size_t)y * w + x;
        const float* c = color + 3 * o;
        Y[o] = (T)quantize_u32(c[0], maxv);
        su += (uint32_t)(T)quantize_u32(c[1], maxv);
        sv += (uint32_t)(T)quantize_u32(c[2], maxv);
    }
    U[(size_t)by * hw + bx] = (T)((su + 2) >> 2);
    V[(size_t)by * hw + bx] = (T)((sv + 2) >> 2);
}

// One thread per 2x2 block.
template <typename T>
__global__ void k_encode(const float* __restrict__ color, int w, int h, unsigned maxv, T* __restrict__ Y,
                         T* __restrict__ U, T* __restrict__ V) {
    const int bx = blockIdx.x * blockDim.x + threadIdx.x;
    const int by = blockIdx.y;
    if (bx >= w / 2 || by >= h / 2) return;
    encode_block(color, w, bx, by, maxv, Y, U, V);
}

// Batch of views (grid.z = view); views without output planes are skipped.
// The planes encode ecol at the view's size (ew x eh).
__global__ void k_encode_views(const ViewDesc* __restrict__ views, const uint8_t* __restrict__ tile_enc) {
    const ViewDesc& vd = views[blockIdx.z];
    const int bx = blockIdx.x * blockDim.x + threadIdx.x;
    const int by = blockIdx.y;
    if (!vd.y || bx >= vd.ew / 2 || by >= vd.eh / 2) return;
    // tiles without holes or deferred pixels were encoded by the raster pass
    if (tile_enc && vd.enc_raster &&
        (tile_enc[vd.tile_base + (uint32_t)(((2 * by) >> vd.tile_shift) * vd.tiles_x + ((2 * bx) >> vd.tile_shift))] & 1))
        return;
    const unsigned maxv = (1u << vd.bitdepth) - 1u;
    if (vd.bitdepth <= 8)
        encode_block(vd.ecol, vd.ew, bx, by, maxv, (uint8_t*)vd.y, (uint8_t*)vd.u, (uint8_t*)vd.v);
    else
        encode_block(vd.ecol, vd.ew, bx, by, maxv, (uint16_t*)vd.y, (uint16_t*)vd.u, (uint16_t*)vd.v);
}

// Oversampled views: colour (w x h) resized to the view's size (ew x eh).
__global__ void k_resize_views(const ViewDesc* __restrict__ views) {
    const ViewDesc& vd = views[blockIdx.z];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (vd.ecol == vd.color || x >= vd.ew || y >= vd.eh) return;
    const float3 c = resize_linear_px(vd.color, vd.w, vd.h, vd.ew, vd.eh, x, y, vd.rtab, vd.rtab + vd.ew);
    float* o = vd.ecol + 3 * ((size_t)y * vd.ew + x);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}

// The same for one colour image (the stage entry point; taps: dw column
// entries, then dh row entries).
__global__ void k_resize_color(const float* __restrict__ src, int sw, int sh, float* __restrict__ dst, int dw,
                               int dh, const int2* __restrict__ taps) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    const float3 c = resize_linear_px(src, sw, sh, dw, dh, x, y, taps, taps + dw);
    float* o = dst + 3 * ((size_t)y * dw + x);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}

hipError_t launch_resize_views(const ViewDesc* views, int nviews, int max_ew, int max_eh, hipStream_t s) {
    hipLaunchKernelGGL(k_resize_views, dim3((max_ew + 255) / 256, max_eh, nviews), dim3(256), 0, s, views);
    return hipGetLastError();
}

hipError_t launch_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh, const int2* taps,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_resize_color, dim3((dw + 255) / 256, dh), dim3(256), 0, s, src, sw, sh, dst, dw, dh, taps);
    return hipGetLastError();
}

hipError_t launch_decode_color(const void* y, const void* u, const void* v, int u8, int w, int h, float cs,
                               float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_decode_color, dim3((w + 255) / 256, h), dim3(256), 0, s, y, u, v, u8, w, h, cs, out);
    return hipGetLastError();
}

hipError_t launch_prep_color(const void* y, const void* u, const void* v, int u8, int w, int h, int packed,
                             void* out, hipStream_t s) {
    hipLaunchKernelGGL(k_prep_color, dim3((w + 255) / 256, h), dim3(256), 0, s, y, u, v, u8, w, h, packed, out);
    return hipGetLastError();
}

hipError_t launch_decode_depth(const void* raw, int u8, size_t n, float scale, float znear, float zfar,
                               int has_invalid, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_decode_depth, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, raw, u8, n, scale, znear,
                       zfar, has_invalid, out);
    return hipGetLastError();
}

hipError_t launch_reproject(const PairGeom& g, float* uv, float* vd, hipStream_t s) {
    hipLaunchKernelGGL(k_reproject, dim3((g.sw + 255) / 256, g.sh), dim3(256), 0, s, g, uv, vd);
    return hipGetLastError();
}

hipError_t launch_blend_step(float* acc_c, float* acc_q, float* acc_d, float* acc_v, uint8_t* acc_m, const float* c,
                             const float* d, const float* shape, size_t n, float a, int first, hipStream_t s) {
    const float a_inv = 1.0f / a;
    hipLaunchKernelGGL(k_blend_step, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, acc_c, acc_q, acc_d, acc_v,
                       acc_m, c, d, shape, n, a, a_inv, first);
    return hipGetLastError();
}

hipError_t launch_encode(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v, hipStream_t s) {
    const unsigned maxv = (1u << bitdepth) - 1u;
    const dim3 grid((w / 2 + 255) / 256, h / 2);
    if (bitdepth <= 8)
        hipLaunchKernelGGL(k_encode<uint8_t>, grid, dim3(256), 0, s, color, w, h, maxv, (uint8_t*)y, (uint8_t*)u,
                           (uint8_t*)v);
    else
        hipLaunchKernelGGL(k_encode<uint16_t>, grid, dim3(256), 0, s, color, w, h, maxv, (uint16_t*)y, (uint16_t*)u,
                           (uint16_t*)v);
    return hipGetLastError();
}

hipError_t launch_encode_views(const ViewDesc* views, int nviews, int max_w, int max_h, const uint8_t* tile_enc,
                               hipStream_t s) {
    const dim3 grid((max_w / 2 + 255) / 256, max_h / 2, nviews);
    hipLaunchKernelGGL(k_encode_views, grid, dim3(256), 0, s, views, tile_enc);
    return hipGetLastError();
}

}  // namespace rvs_k
