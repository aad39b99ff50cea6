// rvs_inpaint.h -- workspace layout and launcher of the exact inpainting kernels.
#ifndef RVS_INPAINT_H
#define RVS_INPAINT_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "rvs_kernels.h"

namespace rvs_k {

constexpr int kInpaintBlocks = 1024;   // persistent component solvers
constexpr int kInpaintBigBlocks = 256;  // solvers of large components (one 1024-thread block each)
constexpr int kStripBlocks = 512;       // persistent strip solvers (one wave each)
constexpr int kHoleTileBlocks = 2048;   // persistent blocks of the passes over the hole tiles

// Indexed by global pixel (a view's pixels start at its px_base).
struct InpaintWS {
    // per pixel, one 16-byte record (a line step of the solvers loads all of
    // it at once): x distance of the map (inpainting.cpp:182-188; holes only),
    // y global index of the chosen known pixel (holes only), z label (-1 known
    // pixel, else union-find parent; the root after flattening), w component id
    // (of a root)
    int4* px;
    int* comp;      // 4 per component: x0, x1, y0, y1 (view coordinates)
    int* root;      // root pixel of a component
    int* counters;  // [0] components, [2] work counter of the small-component pass, [3] strip tasks,
                    // [4] / [5] tickets of the forward / backward strip sweeps, [6] components left to
                    // the one-block solver, [7] strip carry slots in use, [8] tiles with holes
    uint8_t* thole; // [view][tile row][tile column] of 32 x 32 pixel tiles: 1 when the tile has a hole;
                    // the px records are stored only in such tiles
    int* htiles;    // the tiles with holes (index into thole), in no particular order
    int ttx, tpv;   // tiles per row and per view of the thole grid (sized for the largest view)
    const uint8_t* rflags;  // raster tile flags of the views (bit 1: a hole or a deferred pixel), or
                            // null; a 32 x 32 tile whose raster tiles all lack bit 1 has no hole
    int4* tasks;    // strip tasks: component, strip | strips << 8 | cols << 16, first carry slot
    unsigned long long* slots[2];  // per (task, line) scan carries of the forward / backward sweep
    int* legacy;    // large components the strip plan could not place
    uint32_t* fail; // set when a bounded wait gives up (may be null)
    int task_cap, slot_cap;
};

size_t inpaint_ws_bytes(size_t npx, int nviews, int max_w, int max_h);
InpaintWS inpaint_ws_carve(void* base, size_t npx, int nviews, int max_w, int max_h);
// Inpaints every view of the batch in place (views[v].color from views[v].quality).
hipError_t launch_inpaint(const ViewDesc* views, int nviews, int max_w, int max_h, const InpaintWS& ws,
                          int32_t* map_out, hipStream_t s);

}  // namespace rvs_k

#endif
