/*
 * rvs_host.h -- C ABI of the host-side library (librvs_host.so): synthetic scene
 * rendering and the reference-format front end (JSON configuration, camera
 * parameters, pose traces, raw YUV files) used by the drop-in Pipeline.
 */
#ifndef RVS_HOST_H
#define RVS_HOST_H

#include <stdint.h>
#include "rvs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Render a seeded synthetic scene (kind 0: textured rectangles before a plane at
 * 10 m; kind 1: textured spheres inside a box) for one camera.  Colour planes
 * y (w*h), u, v ((w/2)*(h/2)) and the depth plane (w*h) are u8 when the bit depth
 * is <= 8, else u16; depth is quantised inverse depth in [znear, zfar].
 * threads <= 0 uses every hardware thread.  Returns 0 on success. */
int rvs_scene_render(int kind, uint64_t seed, const rvs_cam* cam, const float pos[3], const float rot[3],
                     int bitdepth_color, int bitdepth_depth, float znear, float zfar, void* y, void* u, void* v,
                     void* depth, int threads);

/* Message of the last failing call of this library on the calling thread. */
const char* rvs_host_last_error(void);

/* Parse `text` with the configuration JSON reader, walk `path` ("a/0/b"; numeric
 * segments index arrays, "?key" uses optional()), convert with `as` in
 * {double, float, int, string, bool, size, type, truthy}; overrides_text (may be
 * empty) is applied to every object reached.  Returns 0 and the value in `out`,
 * or 1 and the error text. */
int rvs_json_query(const char* text, const char* overrides_text, const char* path, const char* as, char* out,
                   int cap);
/* Parse a pose trace (PoseTraces.cpp:54-92): writes x,y,z,yaw,pitch,roll per row
 * (up to max_rows) and returns the row count, or -1. */
int rvs_pose_trace_parse(const char* text, float* out, int max_rows);
/* Load a view-synthesis configuration (and the files it names) and describe it
 * as JSON in `out`.  Returns 0, or 1 with the error text. */
int rvs_config_describe(const char* path, char* out, int cap);
/* Run the configuration-driven Application (Pipeline::execute): reads the input
 * YUV files, writes the configured outputs.  stagewise != 0 runs the
 * view-by-view plugin path instead of the fused one.  Returns the number of
 * frames or -1 (rvs_host_last_error). */
int rvs_app_run(const char* config_path, const char* source_path, int stagewise);

#ifdef __cplusplus
}
#endif
#endif
