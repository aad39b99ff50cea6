// rvs_stages.h -- launchers of the element-wise stage kernels.
#ifndef RVS_STAGES_H
#define RVS_STAGES_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_kernels.h"

namespace rvs_k {

// packed: 16-bit Y, 4 U, 4 V per pixel (even sizes, bit depth <= 14), else float4.
hipError_t launch_prep_color(const void* y, const void* u, const void* v, int u8, int w, int h, int packed,
                             void* out, hipStream_t s);
hipError_t launch_decode_color(const void* y, const void* u, const void* v, int u8, int w, int h, float cs,
                               float* out, hipStream_t s);
hipError_t launch_decode_depth(const void* raw, int u8, size_t n, float scale, float znear, float zfar,
                               int has_invalid, float* out, hipStream_t s);
hipError_t launch_reproject(const rvs_dev::PairGeom& g, float* uv, float* vd, hipStream_t s);
hipError_t launch_blend_step(float* acc_c, float* acc_q, float* acc_d, float* acc_v, uint8_t* acc_m,
                             const float* c, const float* d, const float* shape, size_t n, float a, int first,
                             hipStream_t s);
hipError_t launch_encode(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v,
                         hipStream_t s);
// Encodes every view of a batch that has output planes (ViewDesc y, u, v) from
// ecol at the views' own sizes (max_w x max_h: the largest ew x eh); blocks of
// tiles that tile_enc (may be null) marks as encoded by k_raster (bit 0) are skipped.
hipError_t launch_encode_views(const ViewDesc* views, int nviews, int max_w, int max_h, const uint8_t* tile_enc,
                               hipStream_t s);
// write_color's resize (image_writing.cpp:189-190) of one float3 image.
// taps: the resize_taps tables of the columns (dw entries) then the rows (dh)
hipError_t launch_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh, const int2* taps,
                               hipStream_t s);
// Multispectral blending's frequency split of one float3 view (rgb: blur all
// three channels, else Y only); ws: split_ws_bytes(w * h, rgb) bytes.
size_t split_ws_bytes(size_t npx, int rgb);
hipError_t launch_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h,
                                    int rgb, void* ws, float* low, float* high, hipStream_t s);
// out[i] = a[i] + b[i] over n floats.
hipError_t launch_add_color(const float* a, const float* b, float* out, size_t n, hipStream_t s);

}  // namespace rvs_k

#endif
