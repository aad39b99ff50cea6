// rvs_kernels.h -- launch-side declarations shared by the .hip translation units.
//
// A launch set synthesizes a batch of views (a single view is a batch of one)
// from one set of resident sources.  Target tiles of all views are numbered
// globally (view v owns tiles [tile_base, tile_base + tiles_x*tiles_y)), and so
// are output pixels (view v owns [px_base, px_base + w*h)) for the inpainting
// workspace.  Per (view, source) pair the vertex stage reads one descriptor G
// (PairGeom for resident raw sources, ArrayGeom for precomputed inputs) at
// geoms[view * nsrc + source].
#ifndef RVS_KERNELS_H
#define RVS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_device.h"

namespace rvs_k {

constexpr int kTile = 16;            // default target tile edge (pixels), and the edge of the sub-tiles
                                     // the in-order replay works on
constexpr int kTileBig = 32;         // the other tile edge a launch set can use (BatchDims::tile_shift)
constexpr int kBinQX = 64;           // binning work item: 64 x 4 source quads
constexpr int kBinQY = 4;
constexpr int kBinTable = 128;       // LDS bin table entries per binning block (power of two)
constexpr int kFallbackChunk = 128;  // list entries per LDS chunk in the in-order replay
constexpr int kFallbackSortCap = 8192;  // longest (source, tile) list the replay sorts in LDS (longer
                                        // ones are sorted in place in the entry list)
constexpr int kMaxViews = 4096;
constexpr int kBinVerts = (kBinQY + 1) * (kBinQX + 1);   // vertices of one binning work item
constexpr int kBinBlocks = 2048;    // persistent binning blocks (work items past the kept range)
constexpr int kFlatBlocks = 65536;  // persistent blocks of the passes over the kept work items
constexpr int kRasterGroup = 4;     // sources rasterised together in k_raster's one-phase instantiation
constexpr int kRasterGroupMulti = 4; // sources per phase of the looping instantiation (more sources)
constexpr int kRasterCache = 1536;  // binned quads per phase whose coverage-pass results stay in LDS
constexpr int kRasterCacheOne = 960; // the same for launch sets of up to kRasterGroup sources (one phase)
constexpr int kMaxSourceSide = 65536;  // bins hold quads as (row << 16) | column; also the
                                       // largest target side (tile coordinates fit 16 bits)
constexpr int kFallbackBlocks = 1024;

// One target view of a batch.
struct ViewDesc {
    int w, h, tiles_x, tiles_y;
    uint32_t tile_base;      // first global tile
    uint32_t px_base;        // first global pixel
    int bitdepth;
    float blend_exp;
    float blend_inv;         // 1.0f / blend_exp
    float* color;            // w*h*3 (blend) / raster colour (raw)
    float* quality;          // w*h (blend); may be null for raw
    float* depth;            // w*h
    float* validity;         // w*h (blend: validity, raw: triangle shape)
    void *y, *u, *v;         // encoded planes, may be null
    int ew, eh;              // size of the encoded planes: the view's own size (w, h are the raster's,
                             // oversampled by Precision)
    float* ecol;             // colour the planes encode: color, or its resize to ew x eh
    const int2* rtab;        // resize taps of an oversampled view: ew columns, then eh rows
    int tile_shift;          // log2 of the launch set's tile edge
    int enc_raster;          // planes wanted, not oversampled: tiles without holes or deferred
                             // pixels are encoded by k_raster itself
    int partial_color;       // color is internal: tiles encoded by k_raster write it only on
                             // their border (all an inpainting next door can read)
    int partial_quality;     // quality is internal: written only in tiles with a hole or a
                             // deferred pixel (the only ones the inpainting reads)
};

// Device-side state of one launch set.
struct RasterWS {
    uint32_t* counts;        // [global tile][source] entries per bin (counting pass)
    uint32_t* offsets;       // [global tile][source] first entry of each bin in lists
    uint32_t* fill;          // [global tile][source] fill cursors
    uint32_t* lists;         // [cap] quad codes (row << 16 | column), bins back to back
    uint32_t* ranges;        // [range_cap work items][256] packed tile range of each quad
    uint2* ktab;             // [range_cap work items][1 + kBinTable]: the counting pass's LDS bin
                             // table of the item, [0].x = n used slots, then n of (slot | count << 8, tile)
    float* vbuf;             // [range_cap work items][kBinQY + 1][kBinQX + 1][3] vertices (u', v', depth)
    float2* tvq;             // [range_cap work items][256] validity of triangles T1, T2 of each quad
    uint32_t* imap;          // [view][source][item] work index of a kept item, ~0 otherwise
    unsigned long long* bsum;// scan block sums
    uint8_t* deferred;       // [global pixel] pixel handled by the in-order replay
    uint32_t* deferred_tiles;// [global tile]
    uint32_t* work;          // binning work items (view, source, block)
    uint8_t* tile_enc;       // [global tile] bit 0: encoded by k_raster; bit 1: a hole or a deferred
                             // pixel (blending launches); written by every tile
    float* rbuf;             // [rbuf_tiles][source][tile pixel][5] the replay's per-source colour,
                             // depth and validity of deferred pixels
    uint32_t* status;        // [0] n_deferred, [1] list overflow, [2] entries (saturated), [3] inpainting
                             // wait gave up, [4] work items
};

struct BatchDims {
    int nviews, nsrc;
    uint32_t ntiles;         // all views
    uint32_t npx;            // all views
    uint32_t cap;            // list capacity (binned (quad, tile) entries of the launch set)
    uint32_t range_cap;      // work items whose tile ranges and vertices are kept between the passes
    int max_sw, max_sh;      // largest source grid
    uint32_t bins_x, bins_y; // binning work items per source (for the largest source)
    int tile_shift;          // log2 of the tile edge: 4 (16 x 16 tiles, 256-thread raster blocks) or
                             // 5 (32 x 32, 1024-thread blocks)
    uint32_t flat_blocks;    // blocks of the passes over the kept work items
    uint32_t rbuf_tiles;     // deferred tiles whose sources are replayed in parallel (RasterWS::rbuf)
    uint32_t kept_geom;      // kept work items also store vertex windows and validities (vbuf, tvq)
};

enum { MODE_BLEND = 0, MODE_RAW = 1 };

hipError_t launch_reset_u32(uint32_t* p, size_t n, hipStream_t s);
// Colour of every oversampled view resized to its encoded size (views[v].ecol);
// views rasterised at their own size are skipped.
hipError_t launch_resize_views(const ViewDesc* views, int nviews, int max_ew, int max_eh, hipStream_t s);
// Scan blocks needed for nbins bins (bsum entries).
size_t scan_blocks(size_t nbins);
// Raw bytes into device memory through kernel arguments (stream-ordered, capturable).
hipError_t launch_upload(const void* host, size_t bytes, void* dst, hipStream_t s);

// Binning blocks per row / column of a source grid.
inline int bin_blocks_x(int sw) { return (sw - 1 + kBinQX - 1) / kBinQX; }
inline int bin_blocks_y(int sh) { return (sh - 1 + kBinQY - 1) / kBinQY; }
// Floats of a source's culling table: one box (min xyz, max xyz) per binning block.
inline size_t src_bounds_floats(int sw, int sh) { return (size_t)bin_blocks_x(sw) * bin_blocks_y(sh) * 6; }

// Fills the culling table of one source (only its source-side fields are read).
hipError_t launch_src_bounds(const rvs_dev::PairGeom& g, float* out, hipStream_t s);

template <class G>
hipError_t launch_bin(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws, hipStream_t s);
// Tile rasterisation + blending (k_raster), then the in-order replay of the
// deferred pixels (k_fallback); the replay must follow the raster launch.
template <class G, int MODE>
hipError_t launch_raster(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws,
                         hipStream_t s);
template <class G, int MODE>
hipError_t launch_replay(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws,
                         hipStream_t s);

}  // namespace rvs_k

#endif
