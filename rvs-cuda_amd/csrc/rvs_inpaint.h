// rvs_inpaint.h -- workspace layout of the exact inpainting kernels.
#ifndef RVS_INPAINT_H
#define RVS_INPAINT_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace rvs_k {

struct InpaintWS {
    int* label;     // w*h: -1 known pixel, else union-find parent (root after flattening)
    int* dist;      // w*h: distance of the map (inpainting.cpp:182-188)
    int* src;       // w*h: linear index of the chosen known pixel
    int* cid;       // w*h: component id of a root
    int* comp;      // 4 per component: x0, x1, y0, y1
    int* root;      // root pixel of a component
    int* counters;  // [0] components, [1] work counter
};

size_t inpaint_ws_bytes(int w, int h);
InpaintWS inpaint_ws_carve(void* base, int w, int h);
hipError_t launch_inpaint(float* color, const float* quality, int w, int h, const InpaintWS& ws,
                          int32_t* map_out, hipStream_t s);

}  // namespace rvs_k

#endif
