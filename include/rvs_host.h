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

#ifdef __cplusplus
}
#endif
#endif
