// rvs_kernels.h -- launch-side declarations shared by the .hip translation units.
#ifndef RVS_KERNELS_H
#define RVS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_device.h"

namespace rvs_k {

constexpr int kTile = 16;            // target tile edge (pixels); one 256-thread block per tile
constexpr int kBinQX = 64;           // binning block: 64 x 4 source quads
constexpr int kBinQY = 4;
constexpr uint32_t kDefaultTileCap = 1024;
constexpr int kFallbackChunk = 128;  // list entries per LDS chunk in the in-order replay
constexpr int kFallbackSortCap = 8192;  // longest (source, tile) list the replay sorts in LDS

// Device-side state of one raster launch set.
struct RasterWS {
    uint32_t* counts;        // [nsrc][ntiles] bin fill counters
    uint32_t* lists;         // [nsrc][ntiles][cap] quad indices
    uint8_t* deferred;       // [out_h][out_w] pixel handled by the in-order replay
    uint32_t* deferred_tiles;// [ntiles]
    uint32_t* status;        // [0] n_deferred, [1] overflow flag, [2] max count, [3] replay overflow
};

struct RasterDims {
    int out_w, out_h, tiles_x, tiles_y, ntiles;
    uint32_t cap;
};

// Output maps of the raster (single source) or of the blend (all sources).
struct RasterOut {
    float* color;       // w*h*3
    float* depth;       // w*h
    float* quality;     // blend: quality; raw: unused
    float* validity;    // blend: validity; raw: shape
};

enum { MODE_BLEND = 0, MODE_RAW = 1 };

hipError_t launch_reset_u32(uint32_t* p, size_t n, hipStream_t s);

template <class G>
hipError_t launch_bin(const G* geoms, int nsrc, int max_sw, int max_sh, const RasterDims& d,
                      const RasterWS& ws, hipStream_t s);
template <class G, int MODE>
hipError_t launch_raster(const G* geoms, int nsrc, const RasterDims& d, const RasterWS& ws,
                         float blend_exp, const RasterOut& out, hipStream_t s);

}  // namespace rvs_k

#endif
