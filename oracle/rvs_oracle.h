/*
 * rvs_oracle.h -- CPU restatement of the reference view synthesizer's CPU
 * synthesis path (EnzoN7/RVS-CUDA, rvs/src).  TEST INFRASTRUCTURE ONLY: the
 * oracle is the checker for the HIP path.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Every function restates one reference function, cited file:line, with the
 * same float operations in the same order (built with -ffp-contract=off, no
 * fast-math).  OpenCV is not available, so the OpenCV conventions the reference
 * relies on are restated explicitly (see rvs_oracle.cpp, "OpenCV conventions").
 *
 * Parity status per function (how each is pinned is listed in tests/):
 *   orc_euler_to_R, orc_relative_pose      pinned: closed forms + orthonormality
 *   orc_generate_image_pos                 pinned: reference unit_test.cpp:90-101,222-231
 *   orc_unproject, orc_project             pinned: reference unit_test.cpp:103-188,233-308
 *                                           and an FP64 restatement
 *   orc_decode_color / orc_decode_depth     pinned: closed forms (exact bilinear
 *                                           weights, near/far end points)
 *   orc_transform_triangles                 pinned: brute-force tiny meshes, identity
 *                                           warp, analytic homography of a plane;
 *                                           seam triangles (wrap) and the prolongation
 *                                           branch against a brute force
 *   orc_blend_step                          pinned: identity / mean / empty colour cases,
 *                                           exact quality ties, power-sum quality;
 *                                           x * fl(1/S) (OpenCV Vec/float) unpinned
 *   orc_synthesize                          pinned: exact identity warp, holes filled from
 *                                           the L1-nearest covered pixel with the
 *                                           quality mask, first view's scales, input
 *                                           order symmetry except the depth rule
 *   orc_inpaint                             pinned: brute-force L1 distance + literal
 *                                           sweep replay on tiny grids
 *   orc_encode_yuv420                       Y pinned (closed form); U,V chroma rule
 *                                           restates OpenCV 4.9 resize(0.5) area-fast
 *                                           path: "parity unpinned" (OpenCV absent)
 *   orc_blur_masked, orc_blur_masked_rgb,   pinned: brute-force window sums on tiny
 *   orc_split_frequencies                   images, masked-out pixels 0, clipped
 *                                           windows' counts, (int) truncation of the
 *                                           Y sum, float rounding of the RGB sums, a
 *                                           non-finite channel dropping the pixel,
 *                                           low + high = input
 *   orc_synthesize_multispectral            pinned: equal exponents with a radius-0
 *                                           blur reduce to simple blending
 *   orc_resize_color                        write_color's cv::resize to the output size
 *                                           (Precision != 1): pinned by closed forms
 *                                           (constants, 2x block means of integer
 *                                           images, centre taps of a 3x downscale,
 *                                           linear ramps); OpenCV's operation order
 *                                           "parity unpinned" (OpenCV absent)
 */
#ifndef RVS_ORACLE_H
#define RVS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Camera intrinsics as read by Parameters::readFrom (Parameters.cpp:100-129). */
typedef struct {
    int proj;         /* 0 = Perspective, 1 = Equirectangular */
    int w, h;         /* Parameters::getSize() */
    float f[2];       /* Focal (perspective) */
    float p[2];       /* Principle_point minus crop origin (perspective) */
    float hor[2];     /* Hor_range, degrees (ERP) */
    float ver[2];     /* Ver_range, degrees (ERP) */
    int full_hor;     /* Hor_range == [-180, 180] as integers (Parameters.cpp:420-426) */
} orc_cam;

/* ---- geometry (stage a) ---- */
/* Parameters::getRotationMatrix -> EulerAnglesToRotationMatrix (Parameters.cpp:67-97,151-155) */
void orc_euler_to_R(const float rot_deg[3], float R[9]);
/* SpaceTransformer::get_rotation / get_translation (SpaceTransformer.cpp:77-92) */
void orc_relative_pose(const float pos_from[3], const float rot_from[3],
                       const float pos_to[3], const float rot_to[3],
                       float R[9], float t[3]);
/* Unprojector::generateImagePos (Unprojector.cpp:63-74) + ERP override
 * (EquirectangularUnprojector.cpp:87-98).  uv: h*w*2 floats. */
void orc_generate_image_pos(const orc_cam* cam, float* uv);
/* {Equirectangular,Perspective}Unprojector::unproject.  xyz: n*3 floats. */
void orc_unproject(const orc_cam* cam, const float* uv, const float* depth, size_t n, float* xyz);
/* affine_transform (SynthesizedView.cpp:68-79) */
void orc_affine(const float R[9], const float t[3], const float* xyz_in, size_t n, float* xyz_out);
/* {Equirectangular,Perspective}Projector::project; returns the wrapping flag
 * (1 = horizontal).  uv: n*2, depth: n. */
int orc_project(const orc_cam* cam, const float* xyz, size_t n, float* uv, float* depth);
/* cv::transform(virtual_uv, scaled_uv, Matx22f(sx,0,0,sy)) (SynthesizedView.cpp:326-331) */
void orc_scale_uv(const float* uv, size_t n, float sx, float sy, float* out);
/* Whole stage (a) for one (source, target) pair as SynthesizedView::computeBasicMode
 * does it (SynthesizedView.cpp:315-331): uv (n*2), depth (n) on the source grid. */
int orc_reproject(const orc_cam* src, const orc_cam* dst, const float R[9], const float t[3],
                  const float* src_depth, float rescale, float* uv, float* depth);

/* ---- raster (stage b) ---- */
/* detail::transform_trianglesMethod (transform.cpp:148-181) with colorize_triangle
 * (:77-145) and valid_tri (:62-75).  color: in_w*in_h*3, depth/uv on the source grid.
 * Outputs at out_w*out_h: color*3, depth, shape (= triangle_shape / validity). */
void orc_transform_triangles(const float* color, const float* depth, const float* uv,
                             int in_w, int in_h, int out_w, int out_h, int wrap, float rescale,
                             float* out_color, float* out_depth, float* out_shape);

/* ---- blend (stage c) ---- */
/* BlendedViewSimple accumulator (BlendedView.cpp:180-212). */
typedef struct {
    int w, h;
    int empty;          /* m_is_empty */
    float exp;          /* m_blending_exp */
    float* color;       /* w*h*3 */
    float* depth;       /* w*h */
    float* quality;     /* w*h */
    float* validity;    /* w*h */
    uint8_t* depth_mask;/* w*h, m_depth_mask */
} orc_blender;
/* Blend one synthesized view (color, depth, shape as produced by the raster):
 * SynthetisedViewTriangle::transform assigns quality = shape/depth, validity = shape
 * (SynthesizedView.cpp:440), then BlendedViewSimple::blendBasicMode. */
void orc_blend_step(orc_blender* b, const float* color, const float* depth, const float* shape);

/* ---- multispectral blending (BlendingMethod "Multispectral") ---- */
/* detail::calcBlurring of one channel (blending.cpp:62-104) over the masked
 * integral image of IntegralImage2D<float, 1> (IntegralImage2D.h:738-770):
 * window (2 rad + 1)^2 clipped to the image, sums in double in the
 * reference's recurrence order, (int) of the window sum, float division by
 * the count of finite values under mask == 255; 0 where mask == 0.
 * y: every `stride`-th float of the image (3: the Y channel of float3). */
void orc_blur_masked(const float* y, int stride, const uint8_t* mask, int w, int h, int rad, float* out);
/* The three-channel calcBlurring (blending.cpp:109-159) over
 * IntegralImage2D<float, 3> (IntegralImage2D.h:385-420): a pixel counts when
 * mask == 255 and the double sum of its channels is finite; each channel's
 * double window sum is rounded to float (no integer truncation) and divided by
 * the count; 0 where mask == 0.  color, out: float3. */
void orc_blur_masked_rgb(const float* color, const uint8_t* mask, int w, int h, int rad, float* out);
/* BlendedViewMultiSpec::blend's split (BlendedView.cpp:70-80, split_frequencies
 * blending.cpp:371-394): mask = shape/depth > 0, radius max(w, h) / 20.
 * rgb == 0 (ColorSpace YUV): Y blurred, low = (Y blurred, U, V); rgb != 0
 * (ColorSpace RGB): all three channels blurred.  high = colour - low. */
void orc_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h, int rgb,
                           float* low, float* high);

/* ---- inpaint (stage d) ---- */
/* detail::inpaintBasicMode (inpainting.cpp:213-223): mask = !(quality > 0)
 * (View.cpp:198-203), compute_nearest sweeps (:104-136), colour copy (:179-198).
 * color is updated in place.  map_out (optional, w*h*3 ints: src x, src y, dist).
 * Returns the number of sweeps performed. */
int orc_inpaint(float* color, const float* quality, int w, int h, int32_t* map_out);

/* ---- decode / encode (stages a1, a2, e) ---- */
/* read_color_YUV + resizeUV + mergeYUV (image_loading.cpp:74-145,
 * process_input_frame.cu:22-81,211-238).  y: w*h, u/v: (w/2)*(h/2) of u8 (bitdepth<=8)
 * or u16.  color_scale = fl(1/max_level) as image_loading.cpp:134.  out: w*h*3. */
void orc_decode_color(const void* y, const void* u, const void* v, int bitdepth,
                      int w, int h, float color_scale, float* out);
/* read_depth_YUV + normalize (image_loading.cpp:147-213, process_input_frame.cu:346-385) */
void orc_decode_depth(const void* raw, int bitdepth, int w, int h, float depth_scale,
                      float znear, float zfar, int has_invalid, float* out);
/* write_color (image_writing.cpp:182-203): quantization (:142-180) then
 * write_color_YUV (:72-89).  y: w*h, u/v: (w/2)*(h/2), u8 or u16 by bitdepth. */
void orc_encode_yuv420(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v);

/* write_color (image_writing.cpp:189-190): cv::resize(color, color, virtualSize)
 * with the default INTER_LINEAR on a float3 image.  OpenCV 4.9 (imgproc/
 * resize.cpp) takes its area path for an exact 2x downscale (each 2x2 block,
 * summed left to right and top to bottom, times 0.25) and otherwise the
 * half-pixel-centred bilinear rule: f = (float)((d + 0.5) * scale - 0.5),
 * s = floor(f), f -= s; s < 0 -> (0, f = 0); s >= n - 1 -> (n - 1, f = 0);
 * weights (1 - f, f) along rows, then along columns. */
void orc_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh);
/* max_level (image_loading.cpp:227-231) */
unsigned orc_max_level(int bitdepth);

/* ---- whole per-frame path (Pipeline::computeView, Pipeline.cpp:169-301, CPU branch) ---- */
typedef struct {
    orc_cam cam;
    float pos[3], rot[3];
    int bitdepth_color, bitdepth_depth;
    float znear, zfar;
    int has_invalid;
    const void *y, *u, *v;   /* raw colour planes */
    const void *depth;       /* raw depth plane (w*h) */
} orc_source;

typedef struct {
    double t_decode, t_reproject, t_raster, t_blend, t_inpaint, t_encode;  /* seconds */
    int inpaint_sweeps;
} orc_timing;

/* Synthesize one target view from n sources in InputCameraNames order.
 * out_color (w*h*3, optional), out_quality/out_depth/out_validity (w*h, optional:
 * the blender's maps before inpainting), out_y/u/v: encoded planes. */
int orc_synthesize(const orc_source* srcs, int n_src, const orc_cam* tgt,
                   const float tgt_pos[3], const float tgt_rot[3], int tgt_bitdepth,
                   float blending_exp, float rescale,
                   float* out_color, float* out_quality, float* out_depth, float* out_validity,
                   void* out_y, void* out_u, void* out_v, orc_timing* timing);
/* The same with BlendedViewMultiSpec (BlendedView.cpp:62-88): every view is
 * split into frequency bands, the low band blended with exp_low, the high band
 * with exp_high; the result is low + high colour with the low band's maps. */
int orc_synthesize_multispectral(const orc_source* srcs, int n_src, const orc_cam* tgt, const float tgt_pos[3],
                                 const float tgt_rot[3], int tgt_bitdepth, float exp_low, float exp_high, int rgb,
                                 float rescale, float* out_color, float* out_quality, float* out_depth,
                                 float* out_validity, void* out_y, void* out_u, void* out_v);

#ifdef __cplusplus
}
#endif
#endif
