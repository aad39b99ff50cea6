/*
 * rvs_hip.h -- C ABI of the MI355X (gfx950) per-frame view-synthesis path.
 *
 * This is the drop-in boundary that replaces the reference's CUDA library
 * (the headers rvs_cuda_lib/src/<stage>.cuh).  Plain C types only.  Conventions, which differ from
 * the reference on purpose:
 *   - every device buffer is owned by the caller: no call allocates, frees or
 *     consumes its inputs (the reference's wrappers cudaMalloc their outputs and
 *     cudaFree their inputs, e.g. blend_img.cu:200-206);
 *   - every call is stream-ordered on the hipStream_t it is given (NULL = the
 *     default stream) and returns without synchronising, except where stated;
 *   - parameters travel as kernel arguments: no __constant__ globals, so calls
 *     on different streams may run concurrently (the reference's
 *     cudaMemcpyToSymbol parameters make it non-reentrant);
 *   - results follow the reference's CPU path (rvs/src), which is the parity
 *     target, not the reference's CUDA kernels (which compute different results);
 *   - 0 on success, a negative RVS_E_* code otherwise; rvs_last_error() gives a
 *     message for the calling thread.  No C++ exception crosses this boundary.
 */
#ifndef RVS_HIP_H
#define RVS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* rvs_stream_t;   /* hipStream_t */

#define RVS_OK 0
#define RVS_E_INVALID (-1)   /* bad argument */
#define RVS_E_HIP (-2)       /* HIP runtime error */
#define RVS_E_CAPACITY (-3)  /* the binned-entry list or a workspace limit was exceeded; see rvs_last_error */
#define RVS_E_UNSUPPORTED (-4)
#define RVS_E_TIMEOUT (-5)   /* a bounded device-side wait gave up; the call's results are not valid */

#define RVS_MAX_SOURCES 64

/* Camera intrinsics (Parameters, rvs/src/Parameters.hpp:79-160). */
typedef struct {
    int proj;        /* 0 = Perspective, 1 = Equirectangular */
    int w, h;        /* Parameters::getSize() */
    float f[2];      /* Focal (perspective) */
    float p[2];      /* Principle_point minus crop origin (perspective) */
    float hor[2];    /* Hor_range in degrees (ERP) */
    float ver[2];    /* Ver_range in degrees (ERP) */
    int full_hor;    /* Parameters::isFullHorRange() */
} rvs_cam;

/* Relative pose source -> target (SpaceTransformer::get_rotation/get_translation,
 * rvs/src/SpaceTransformer.cpp:77-92), row-major R. */
typedef struct {
    float R[9];
    float t[3];
} rvs_xform;

const char* rvs_last_error(void);
/* Library version string. */
const char* rvs_version(void);

/* ------------------------------------------------------------------ host-side
 * Pure host functions (no GPU): the float semantics of the reference. */

/* R, t as Parameters::getRotationMatrix + SpaceTransformer compute them in float
 * (Parameters.cpp:67-97,151-155; SpaceTransformer.cpp:77-92). */
int rvs_relative_pose(const float pos_from[3], const float rot_from[3], const float pos_to[3],
                      const float rot_to[3], rvs_xform* out);
/* Number of floats of the ERP unprojection table of a camera: 2*w + 2*h. */
size_t rvs_erp_table_len(const rvs_cam* cam);
/* Fill the table on the host: cos(phi_j), sin(phi_j) for every column and
 * cos(theta_i), sin(theta_i) for every row, computed with the reference's float
 * expressions (EquirectangularUnprojector.cpp:62-80, rows 0 and h-1 at v = 1e-3
 * and h - 1e-3, :87-98).  The caller copies it to the device. */
int rvs_erp_table_fill(const rvs_cam* cam, float* host_out);

/* ---------------------------------------------------------- stage entry points
 * One call per stage of the reference's rvs_cuda_lib, caller-owned buffers. */

/* Replaces resizeImage + mergeFrame (process_input_frame.cuh:9-11): raw planar
 * YUV420 (u8 when bitdepth <= 8, else u16) -> interleaved float3 Y,U,V scaled by
 * color_scale (fl(1/max_level) as image_loading.cpp:134).  out: w*h*3 floats. */
int rvs_decode_yuv420(const void* y, const void* u, const void* v, int bitdepth, int w, int h,
                      float color_scale, float* out, rvs_stream_t stream);

/* Replaces formatDepth (process_input_frame.cuh:13): raw depth plane -> metric
 * depth; NaN where has_invalid and the result is 0. */
int rvs_decode_depth(const void* raw, int bitdepth, int w, int h, float depth_scale, float znear,
                     float zfar, int has_invalid, float* out, rvs_stream_t stream);

/* Replaces unprojectERP_projectERP / unprojectERP_projectPerspective + scaleUV
 * (unproject_project.cuh:12-20, scale_uv.cuh:12), for ERP and perspective
 * sources and targets: source depth (src->w*src->h) -> target uv (x2) and
 * target depth, on the source grid.  erp_table: device copy of
 * rvs_erp_table_fill(src) when src is ERP, else NULL.  rescale = Precision. */
int rvs_reproject(const rvs_cam* src, const rvs_cam* dst, const rvs_xform* xf, const float* depth,
                  const float* erp_table, float rescale, float* uv, float* vdepth,
                  rvs_stream_t stream);

/* Replaces transform_trianglesMethod (transform_triangle.cuh:12-16) with the
 * CPU semantics of transform.cpp:148-181: tile-binned rasterisation whose
 * per-pixel winner equals the reference's sequential fold bit for bit.
 * color: src w*h*3, depth/uv on the source grid (as produced by
 * rvs_reproject).  Outputs at out_w*out_h: color*3, depth (+inf uncovered),
 * shape (triangle validity, 0 uncovered).  workspace: rvs_raster_workspace_size
 * (0 for invalid sizes).
 * entry_cap: capacity of the binned (quad, tile) entry list, 0 = one entry per
 * output pixel (at least 2^20).  Synchronises the stream; RVS_E_CAPACITY
 * reports the needed entry_cap through needed_cap (if not NULL) when the
 * entries exceeded it (the outputs are then not written). */
size_t rvs_raster_workspace_size(int in_w, int in_h, int out_w, int out_h, uint32_t entry_cap);
int rvs_raster(const float* color, const float* depth, const float* uv, int in_w, int in_h,
               int out_w, int out_h, int wrap, float rescale, void* workspace, size_t ws_bytes,
               uint32_t entry_cap, float* out_color, float* out_depth, float* out_shape,
               uint32_t* needed_cap, rvs_stream_t stream);

/* Replaces blendImages (blend_img.cuh:12-17) with the CPU semantics of
 * BlendedViewSimple::blendBasicMode (BlendedView.cpp:180-212): blends one
 * synthesized view (color, depth, shape) into the accumulator (acc_*), in place.
 * first != 0 adopts the view (the reference's m_is_empty branch).
 * acc_mask: u8 per pixel (m_depth_mask). */
int rvs_blend_step(float* acc_color, float* acc_quality, float* acc_depth, float* acc_validity,
                   uint8_t* acc_mask, const float* color, const float* depth, const float* shape,
                   int w, int h, float blending_exp, int first, rvs_stream_t stream);

/* Replaces inpaintImg (inpaint_img.cuh:12) with the CPU semantics of
 * detail::inpaintBasicMode (inpainting.cpp:104-223): holes are pixels whose
 * quality is not > 0; each takes the colour of the nearest known pixel chosen
 * exactly as the reference's alternating raster sweeps choose it.  color is
 * updated in place.  map_out (optional device int32 w*h*3): source x, source
 * y, distance.  workspace: rvs_inpaint_workspace_size. */
size_t rvs_inpaint_workspace_size(int w, int h);
int rvs_inpaint(float* color, const float* quality, int w, int h, void* workspace,
                size_t ws_bytes, int32_t* map_out, rvs_stream_t stream);

/* write_color's resize of an oversampled colour image to the output size
 * (image_writing.cpp:189-190, cv::resize with INTER_LINEAR): an exact 2x
 * downscale averages each 2x2 block (OpenCV's area path), other sizes use
 * half-pixel-centred bilinear weights.  OpenCV's operation order is not
 * pinned (no OpenCV here); the oracle restates the same rule.  float3 images.
 * The workspace (rvs_resize_color_workspace_size bytes) receives the per
 * column and per row taps, computed on the host in double as cv::resize does;
 * the call returns once they are enqueued (the host copy is not kept). */
size_t rvs_resize_color_workspace_size(int sw, int sh, int dw, int dh);
int rvs_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh, void* workspace,
                     size_t ws_bytes, rvs_stream_t stream);

/* BlendedViewMultiSpec::blend's frequency split (BlendedView.cpp:70-80;
 * split_frequencies, blending.cpp:371-394; calcBlurring, :62-159) of one
 * synthesized view (color float3, depth, shape as rvs_raster writes them):
 * mask = shape / depth > 0; a box blur of radius max(w, h) / 20 over the
 * finite values under the mask, divided by their count (0 outside the mask);
 * rgb == 0 (ColorSpace YUV) blurs Y, truncating each window sum to an integer
 * as the reference does, and keeps U, V in the low band; rgb != 0 (ColorSpace
 * RGB) blurs all three channels.  low = blurred colour, high = color - low.
 * Window sums are exact fixed-point sums (2^-40 units) where the reference
 * sums in double, so a sum lying within rounding distance of an integer (YUV)
 * or of a float midpoint (RGB) may come out one step apart; colour values
 * must satisfy |v| < 2^22.  workspace: rvs_split_frequencies_workspace_size
 * (12 bytes per pixel for YUV, 28 for RGB; 0 for invalid sizes). */
size_t rvs_split_frequencies_workspace_size(int w, int h, int rgb);
int rvs_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h, int rgb,
                          void* workspace, size_t ws_bytes, float* low, float* high, rvs_stream_t stream);

/* out = a + b over w*h float3 pixels: the multispectral result, low band plus
 * high band colour (BlendedView.cpp:86). */
int rvs_add_color(const float* a, const float* b, float* out, int w, int h, rvs_stream_t stream);

/* Replaces quantization (process_input_frame.cuh:15) with the CPU semantics of
 * write_color (image_writing.cpp:142-203): Y = (T)(v*max) for every pixel, U,V
 * quantised at full resolution then 2x2 averaged as (a+b+c+d+2)>>2. */
int rvs_encode_yuv420(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v,
                      rvs_stream_t stream);

/* ------------------------------------------------------ fused per-view path */

/* One input view resident on the device (InputView, rvs/src/View.hpp:156-174). */
typedef struct {
    rvs_cam cam;
    int bitdepth_color, bitdepth_depth;
    float znear, zfar;        /* Depth_range */
    int has_invalid;          /* HasInvalidDepth */
    float pos[3], rot[3];     /* Position, Rotation (yaw, pitch, roll in degrees) */
    const void *y, *u, *v;    /* device colour planes (u8 or u16) */
    const void* depth;        /* device depth plane (Y plane only, w*h) */
    const float* erp_table;   /* device rvs_erp_table_fill(cam) for ERP sources, else NULL */
    const float* prep;        /* device rvs_source_prep table, or NULL (then built per call) */
    float prep_depth_scale;   /* depth_scale the prep table was built with */
} rvs_source;

/* Resident per-source preparation, built once per loaded source frame and
 * reused by every view while the frame's planes are unchanged:
 *   - binning metadata: for every block of 64 x 4 source quads, the bounding
 *     box of its unprojected points (source camera frame).  A perspective
 *     target skips the blocks that lie behind it or beyond an image side (with
 *     error margins; the result is unchanged).  Valid for targets with the
 *     same depth_scale;
 *   - the decoded colour plane (Y and upsampled U, V before the colour scale,
 *     16 bytes per source pixel), read by the winner resolve instead of the
 *     raw YUV420 planes (the same values).
 * rvs_source_prep fills `out` (device, rvs_source_prep_size bytes) and sets
 * src->prep and src->prep_depth_scale.  rvs_synthesize* build the binning
 * table in their workspace when src->prep is NULL or its depth_scale differs
 * from the target's, and decode colours from the raw planes when src->prep is
 * NULL. */
size_t rvs_source_prep_size(const rvs_cam* cam);
int rvs_source_prep(rvs_source* src, float depth_scale, void* out, size_t bytes, rvs_stream_t stream);

/* The virtual view to synthesize (VirtualCameraNames[0] after the pose trace). */
typedef struct {
    rvs_cam cam;
    float pos[3], rot[3];
    int bitdepth;             /* output BitDepthColor */
    float blending_exp;       /* BlendingFactor */
    float rescale;            /* Precision: rasterise, blend and inpaint at (int)(0.5f + w * rescale) x
                                 (int)(0.5f + h * rescale), encode after resizing to w x h
                                 (SynthesizedView.cpp:326-331, image_writing.cpp:189-190) */
    float color_scale;        /* fl(1/max_level) of the first loaded view (image_loading.cpp:134) */
    float depth_scale;        /* idem for depth (image_loading.cpp:193) */
    int tile;                 /* raster tile edge: 0 or 16 (16 x 16 tiles), or 32; a launch set uses
                                 the largest its views ask for (results are the same; the autotuner
                                 picks the faster per workload) */
} rvs_target;

/* Device outputs of one synthesized view; any pointer may be NULL, but a view
 * needs the colour map or the three planes.  The float maps are at the raster
 * size, (int)(0.5f + w * rescale) x (int)(0.5f + h * rescale) (w x h at
 * Precision 1); the encoded planes at w x h.  A view with planes and no colour
 * map keeps its colour in the workspace: tiles with neither holes nor pixels
 * left to the in-order replay are then encoded by the raster pass directly and
 * their colour is stored only where inpainting can read it (less traffic, the
 * same planes). */
typedef struct {
    float* color;             /* raster size * 3, inpainted (float3 Y,U,V) */
    float* quality;           /* raster size, blended quality before inpainting */
    float* depth;             /* raster size, blended depth */
    float* validity;          /* raster size, blended validity */
    void *y, *u, *v;          /* encoded YUV420 planes at w x h (u8/u16 by target bitdepth) */
} rvs_view_out;

/* Workspace bytes for rvs_synthesize over these sources and target (0 when an
 * argument is invalid).  entry_cap = capacity of the binned (quad, tile) entry
 * list of the launch set (0 = one entry per output pixel and source, at least
 * 2^20); bins are compact, so only the total matters. */
size_t rvs_synth_workspace_size(const rvs_source* srcs, int n_src, const rvs_target* tgt,
                                uint32_t entry_cap);

/* Pipeline::computeView for one virtual view (Pipeline.cpp:169-301, CPU
 * semantics): for each source in order, reproject + rasterise + blend, then
 * inpaint and encode.  Sources must already be resident.  Returns
 * RVS_E_CAPACITY (and the needed entry_cap through needed_cap, if not NULL)
 * when the binned entries exceeded entry_cap; this check synchronises the
 * stream.
 * Equivalent to rvs_synthesize_batch with one view. */
int rvs_synthesize(const rvs_source* srcs, int n_src, const rvs_target* tgt, void* workspace,
                   size_t ws_bytes, uint32_t entry_cap, const rvs_view_out* out,
                   uint32_t* needed_cap, rvs_stream_t stream);

/* Batches: n_views virtual views (e.g. the poses of a pose trace, or requests
 * admitted together by a server) synthesized from the same resident sources by
 * ONE launch set.  Every stage runs once over all views (tiles and pixels are
 * numbered across the batch), so per-launch costs are paid once per batch and
 * small views still fill the device.  Views may differ in camera and size.
 * Results are those of n_views separate rvs_synthesize calls.  At most
 * 4096 views and 2^31 output pixels per batch. */
size_t rvs_synth_batch_workspace_size(const rvs_source* srcs, int n_src, const rvs_target* tgts,
                                      int n_views, uint32_t entry_cap);
/* The same plan with its outcome: RVS_OK and the workspace bytes in *ws_bytes,
 * or the code a launch with these arguments would fail with (RVS_E_INVALID,
 * RVS_E_UNSUPPORTED, RVS_E_CAPACITY; rvs_last_error says why).  Touches no
 * device.  No counterpart in the reference, whose wrappers allocate per call:
 * for request loops that must tell a malformed request from one that does not
 * fit their memory budget. */
int rvs_synth_plan(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                   uint32_t entry_cap, size_t* ws_bytes);
int rvs_synthesize_batch(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                         void* workspace, size_t ws_bytes, uint32_t entry_cap,
                         const rvs_view_out* outs, uint32_t* needed_cap, rvs_stream_t stream);
/* Enqueue only: no synchronisation and no host-device copies, so the call can be
 * captured into a HIP graph (hipStreamBeginCapture) and replayed.  Descriptors
 * travel as kernel arguments: a replay re-runs the captured poses and buffers.
 * Check for list overflow afterwards with rvs_synth_check (same arguments);
 * an overflowed batch must be re-run with the reported entry_cap. */
int rvs_synthesize_batch_async(const rvs_source* srcs, int n_src, const rvs_target* tgts,
                               int n_views, void* workspace, size_t ws_bytes, uint32_t entry_cap,
                               const rvs_view_out* outs, rvs_stream_t stream);
/* The same launch set in two stream-ordered parts: binning, rasterisation and
 * the in-order replay on `stream`, then (after an event recorded there)
 * inpainting and encoding on `finish_stream`.  No counterpart in the
 * reference (its stages run back to back, Pipeline.cpp:169-301): for request
 * loops whose batches use separate workspaces, so that one batch's inpainting
 * and encoding overlap the next batch's binning and rasterisation on
 * `stream`.  Check overflow with rvs_synth_check on `finish_stream`. */
int rvs_synthesize_batch_async_split(const rvs_source* srcs, int n_src, const rvs_target* tgts,
                                     int n_views, void* workspace, size_t ws_bytes, uint32_t entry_cap,
                                     const rvs_view_out* outs, rvs_stream_t stream,
                                     rvs_stream_t finish_stream);
int rvs_synth_check(const rvs_source* srcs, int n_src, const rvs_target* tgts, int n_views,
                    const void* workspace, uint32_t entry_cap, uint32_t* needed_cap,
                    rvs_stream_t stream);

/* rvs_synthesize without a caller workspace and without allocating: the
 * workspace is a static device array of the library (RVS_NOALLOC_SCRATCH_BYTES,
 * 128 MiB, held on every GPU the library runs on until the process ends),
 * whose device address is looked up once per device on the first call there.
 * No counterpart in the reference (its wrappers cudaMalloc per call): for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid it).  Same arguments and results as rvs_synthesize; the binned
 * vertex windows kept between passes are limited to what fits (the rest are
 * re-evaluated, with the same results).  A view whose fixed workspace exceeds
 * the scratch is refused with RVS_E_CAPACITY and a message giving the bytes
 * it needs.  Synchronises the stream.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other threads or streams) race. */
#define RVS_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int rvs_synthesize_noalloc(const rvs_source* srcs, int n_src, const rvs_target* tgt, uint32_t entry_cap,
                           const rvs_view_out* out, uint32_t* needed_cap, rvs_stream_t stream);

/* Per-stage device times of the last rvs_synthesize* call on the calling
 * thread, measured with HIP events recorded on its stream between the stages
 * when enabled: [0] binning (incl. culling tables built per call, culling),
 * [1] tile raster + blend (k_raster), [2] in-order replay of deferred pixels,
 * [3] inpainting, [4] encoding; milliseconds.  After an _async call,
 * synchronise the stream first. */
#define RVS_PROF_STAGES 5
int rvs_profile_enable(int on);
int rvs_profile_read(float ms[RVS_PROF_STAGES]);
/* Counters of the last profiled call: [0] tiles with pixels replayed in order,
 * [1] list overflow flag, [2] largest (source, tile) bin, [3] total binned
 * (quad, tile) entries over all sources, [4] binning work items left after
 * culling.  Synchronises the device. */
#define RVS_PROF_COUNTERS 5
int rvs_profile_counters(uint32_t out[RVS_PROF_COUNTERS]);

#ifdef __cplusplus
}
#endif
#endif
