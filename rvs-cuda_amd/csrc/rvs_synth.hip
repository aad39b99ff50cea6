// rvs_synth.hip -- triangle binning, tile rasterisation and blending for gfx950.
//
// Reference semantics: detail::transform_trianglesMethod (rvs/src/transform.cpp:148-181)
// visits source quads in raster order and folds every covered target pixel
// sequentially: a candidate (depth d, triangle validity v) replaces the state
// (D, V) when fl(fl(fl(r*r)*r)*V) < v with r = fl(d/D), strict, so ties keep the
// earlier triangle.  Mathematically that is the argmax of v/d^3 with the lowest
// traversal index on ties, but the float test is not transitive for
// near-equal scores.  This file computes it for a batch of target views that
// share one set of resident sources:
//   0. k_bin_cull: binning work items (view, source, 64x4 source quads) whose
//      source points all lie behind a perspective view or beyond one side of
//      its image are dropped (bounding boxes of the unprojected points come
//      from k_src_bounds and are tested with error margins); the rest form a
//      work list.
//   1. k_bin (counting pass, scan, fill pass): every surviving source quad
//      is appended to the 16x16 or 32x32 target tiles its triangles' narrowed
//      (reference-exact) bounding boxes touch; the counting pass keeps each
//      work item's vertex block and validities for the raster.
//   2. k_raster: one 256- or 1024-thread block per target tile, four sources per
//      phase (one phase, 8 waves per SIMD, for launch sets of up to four
//      sources).  Every covered pixel takes an LDS 64-bit atomicMax, key1 = max
//      of (score bits << 32 | candidate id), and a 32-bit one, key2 = max
//      score over the candidates that are not the maximum (the loser of each
//      key1 update).  A pixel whose runner-up scores within 2^-18 of the
//      winner (exact ties included), or that saw a candidate with non-finite
//      or extreme values (key2 = all ones), is deferred.  Every other pixel
//      provably ends the sequential fold on the argmax (the float test errs
//      by less than 2^-21 relative), so its winner is re-evaluated -- from
//      the kept vertices and validity of the work item the coverage pass
//      left in LDS -- and blended.
//   3. k_fallback_src / k_fallback_blend / k_fallback: deferred pixels replay
//      the reference fold literally, in traversal order (per-pixel candidate
//      lists gathered in LDS, or a sort of the tile's bins when a list
//      overflows), each source of a tile in its own block and then blended,
//      or (beyond the workspace for that) a block per tile folding its
//      sources in turn.
// Blending across sources (BlendedViewSimple::blendBasicMode) runs per pixel in
// registers inside the same block, in InputCameraNames order.
#include <hip/hip_runtime.h>

#include "rvs_kernels.h"

namespace rvs_k {

using namespace rvs_dev;

__global__ void k_reset_u32(uint32_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_reset_u32(uint32_t* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_reset_u32, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// Descriptor uploads travel as kernel arguments, so a launch set stays
// stream-ordered and capturable in a graph (no host-side memcpy).
constexpr uint32_t kBlobBytes = 3584;
struct Blob {
    uint32_t n;
    uint8_t bytes[kBlobBytes];
};

__global__ void k_upload(Blob b, uint8_t* dst) {
    for (uint32_t i = threadIdx.x; i < b.n; i += blockDim.x) dst[i] = b.bytes[i];
}

hipError_t launch_upload(const void* host, size_t bytes, void* dst, hipStream_t s) {
    const uint8_t* src = (const uint8_t*)host;
    for (size_t off = 0; off < bytes; off += kBlobBytes) {
        Blob b;
        b.n = (uint32_t)((bytes - off) < (size_t)kBlobBytes ? (bytes - off) : (size_t)kBlobBytes);
        for (uint32_t i = 0; i < b.n; ++i) b.bytes[i] = src[off + i];
        hipLaunchKernelGGL(k_upload, dim3(1), dim3(256), 0, s, b, (uint8_t*)dst + off);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// View owning global tile T (views are sorted by tile_base).
__device__ inline int find_view(const ViewDesc* views, int nviews, uint32_t T) {
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].tile_base <= T) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Tile index of block b: each XCD (blocks are dealt round-robin over 8 XCDs)
// gets a contiguous band of tiles, so neighbouring tiles share that XCD's L2.
__device__ inline uint32_t tile_of_block(uint32_t b, uint32_t ntiles) {
    const uint32_t per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

// ------------------------------------------------------------------ quads

// The (up to four) triangles of source quad (i, j), with the reference's
// vertex gating (transform.cpp:161-176).  alive bit k set = colorize_triangle
// is called for slot k and gets past its early returns.
struct Quad {
    Tri t[4];
    uint32_t alive;
    uint32_t prol;      // bit k: prol flag of slot k (transform.cpp:121-123)
};

__device__ inline bool finite3(float3 v) { return is_finite(v.x) && is_finite(v.y) && is_finite(v.z); }

__device__ inline void quad_add(Quad& Q, int slot, float3 a, float3 b, float3 c, int out_w, int out_h, float rescale,
                                bool validity) {
    Tri T;
    const bool ok = validity ? tri_setup(T, a, b, c, out_w, out_h, rescale) : tri_setup_geom(T, a, b, c, out_w, out_h);
    Q.t[slot] = T;
    if (!ok) return;
    Q.alive |= 1u << slot;
    if ((a.z > 0.f) || (b.z > 0.f) || (c.z > 0.f)) Q.prol |= 1u << slot;
}

// validity = false leaves t[k].tv at 0 (binning needs only which triangles
// survive and their bounding boxes).
template <class G, class VF>
__device__ inline void quad_build(Quad& Q, const G& g, int i, int j, VF vtx, int out_w, int out_h,
                                  bool validity = true) {
    Q.alive = 0; Q.prol = 0;
    Q.t[0].tv = 0.f; Q.t[1].tv = 0.f;
    const float3 v00 = vtx(i, j), v01 = vtx(i, j + 1), v10 = vtx(i + 1, j), v11 = vtx(i + 1, j + 1);
    if (!(v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f)) return;
    if (v00.z > 0.f && v00.x > 0.f) quad_add(Q, 0, v00, v01, v10, out_w, out_h, g.rescale, validity);
    if (v11.z > 0.f && v11.x > 0.f) quad_add(Q, 1, v11, v10, v01, out_w, out_h, g.rescale, validity);
    if (g.wrap && j == 0) {
        const float3 l0 = vtx(i, g.sw - 1), l1 = vtx(i + 1, g.sw - 1);
        quad_add(Q, 2, l0, v00, l1, out_w, out_h, g.rescale, validity);
        quad_add(Q, 3, v10, l1, v00, out_w, out_h, g.rescale, validity);
    }
}

// Tile range [tx0, tx1] x [ty0, ty1] of the bounding boxes, narrowed to the
// pixels each can cover (tri_tight), of the triangles of source quad (i, j)
// that survive the gating and colorize_triangle's early returns (the union
// quad_build + quad_tiles compute), one triangle at a time; with tv, also the
// validities of T1 and T2 (0 when not alive or left undecided by the fast
// path).  False when no triangle can cover a pixel.
template <class G, class VF>
__device__ inline bool quad_range_tv(const G& g, int i, int j, VF vtx, int out_w, int out_h, int shift, int& tx0,
                                     int& tx1, int& ty0, int& ty1, float2* tv) {
    if (tv) *tv = make_float2(0.f, 0.f);
    const float3 v01 = vtx(i, j + 1), v10 = vtx(i + 1, j);
    if (!(v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f)) return false;
    int xmin = INT_MAX, xmax = INT_MIN, ymin = INT_MAX, ymax = INT_MIN;
    auto one = [&](float3 a, float3 b, float3 c, float* tvo) {
        Tri T;
        if (!tri_setup_geom(T, a, b, c, out_w, out_h)) return;
        // tiles only where the triangle can cover a pixel (tri_tight; a triangle
        // that can cover none is never rasterised, so its validity is unused)
        int xa = T.xmin, xb = T.xmax, ya = T.ymin, yb = T.ymax;
        tri_tight(T, xa, xb, ya, yb);
        if (xa <= xb && ya <= yb) {
            xmin = min(xmin, xa); xmax = max(xmax, xb);
            ymin = min(ymin, ya); ymax = max(ymax, yb);
        }
        // fast path only: an undecided validity stays 0 (unknown), and the
        // raster then defers the pixels that depend on it
        if (tvo && !tri_validity_fast(T, g.rescale, *tvo)) *tvo = 0.f;
    };
    const float3 v00 = vtx(i, j);
    if (v00.z > 0.f && v00.x > 0.f) one(v00, v01, v10, tv ? &tv->x : nullptr);
    {
        const float3 v11 = vtx(i + 1, j + 1);
        if (v11.z > 0.f && v11.x > 0.f) one(v11, v10, v01, tv ? &tv->y : nullptr);
    }
    if (g.wrap && j == 0) {
        const float3 l0 = vtx(i, g.sw - 1), l1 = vtx(i + 1, g.sw - 1);
        one(l0, v00, l1, nullptr);
        one(v10, l1, v00, nullptr);
    }
    if (xmin > xmax) return false;
    tx0 = xmin >> shift; tx1 = xmax >> shift; ty0 = ymin >> shift; ty1 = ymax >> shift;   // boxes start at >= 0
    return true;
}

// Calls fn(slot, triangle, prol) for every triangle of source quad (i, j)
// that gets past the vertex gating and colorize_triangle's early returns, in
// slot order (the quad_build gating, one triangle at a time); vertices come
// from vtx (exact), and the triangle's validity is left 0 for fn to evaluate
// (tri_validity) when it needs it.
template <class G, class VF, class F>
__device__ inline void quad_tris(const G& g, int i, int j, VF vtx, int out_w, int out_h, F fn) {
    const float3 v00 = vtx(i, j), v01 = vtx(i, j + 1), v10 = vtx(i + 1, j);
    const float3 v11 = vtx(i + 1, j + 1);
    if (!(v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f)) return;
    auto one = [&](int slot, float3 a, float3 b, float3 c) {
        Tri T;
        if (tri_setup_geom(T, a, b, c, out_w, out_h)) fn(slot, T, (a.z > 0.f) || (b.z > 0.f) || (c.z > 0.f));
    };
    if (v00.z > 0.f && v00.x > 0.f) one(0, v00, v01, v10);
    if (v11.z > 0.f && v11.x > 0.f) one(1, v11, v10, v01);
    if (g.wrap && j == 0) {
        const float3 l0 = vtx(i, g.sw - 1), l1 = vtx(i + 1, g.sw - 1);
        one(2, l0, v00, l1);
        one(3, v10, l1, v00);
    }
}

// Bins hold source quads as (row << 16) | column: ascending order of these
// codes is the reference's traversal order (rows, then columns).
__device__ inline void unpack_ij(uint32_t ij, int& i, int& j) {
    i = (int)(ij >> 16);
    j = (int)(ij & 0xffffu);
}

// ------------------------------------------------------------ culling

// Culling table of one source: per binning block (kBinQY x kBinQX quads), the
// bounding box (source camera frame) of the finite unprojected points of its
// vertex window (kBinQY+1 rows x kBinQX+1 columns) as min x, y, z, max x, y, z;
// no finite point -> min +inf, max -inf.  One wave per block.
__global__ void __launch_bounds__(256) k_src_bounds(PairGeom g, int bins_x, int bins_y, float* out) {
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= bins_x * bins_y) return;   // whole wave
    const int i0 = (item / bins_x) * kBinQY, j0 = (item % bins_x) * kBinQX;
    float v[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(),
                  -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int k = lane; k < (kBinQY + 1) * (kBinQX + 1); k += 64) {
        const int i = i0 + k / (kBinQX + 1), j = j0 + k % (kBinQX + 1);
        if (i >= g.sh || j >= g.sw) continue;
        const float3 p = unproject_src(g, i, j);
        if (!finite3(p)) continue;
        v[0] = fminf(v[0], p.x); v[1] = fminf(v[1], p.y); v[2] = fminf(v[2], p.z);
        v[3] = fmaxf(v[3], p.x); v[4] = fmaxf(v[4], p.y); v[5] = fmaxf(v[5], p.z);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = fminf(v[c], __shfl_xor(v[c], off));
#pragma unroll
        for (int c = 3; c < 6; ++c) v[c] = fmaxf(v[c], __shfl_xor(v[c], off));
    }
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 6; ++c) out[(size_t)item * 6 + c] = v[c];
}

hipError_t launch_src_bounds(const PairGeom& g, float* out, hipStream_t s) {
    const int bx = bin_blocks_x(g.sw), by = bin_blocks_y(g.sh);
    hipLaunchKernelGGL(k_src_bounds, dim3((unsigned)((bx * by + 3) / 4)), dim3(256), 0, s, g, bx, by, out);
    return hipGetLastError();
}

// True when no triangle built from points inside box b can be alive for a
// perspective target: every point is behind the camera (X < 0), or beyond one
// image side by more than two pixels.  For X > 0 the side tests are linear in
// the point, u < -m <=> f0*Y - (p0 + m)*X > 0 (u = -f0*Y/X + p0), so their
// minimum over the box is at a corner.  A triangle is alive only if all its
// vertices have X > 0 (vertex gating), and a vertex with u' < -1 everywhere
// (or > W) gives xmin > xmax in tri_setup.  The margins cover the float error
// of the corner evaluation here and of the vertex stage (both below 1e-6 of
// the magnitudes involved); overflow or NaN never culls.
__device__ inline bool box_outside(const PairGeom& g, const float* b, int out_w, int out_h) {
    if (!(b[0] <= b[3])) return true;   // no finite point: every vertex is gated
    float X[8], Y[8], Z[8];
    float mag = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float x = (k & 1) ? b[3] : b[0], y = (k & 2) ? b[4] : b[1], z = (k & 4) ? b[5] : b[2];
        X[k] = g.R[0] * x + g.R[1] * y + g.R[2] * z + g.t[0];
        Y[k] = g.R[3] * x + g.R[4] * y + g.R[5] * z + g.t[1];
        Z[k] = g.R[6] * x + g.R[7] * y + g.R[8] * z + g.t[2];
        mag = fmaxf(mag, fabsf(x) + fabsf(y) + fabsf(z));
    }
    mag += fabsf(g.t[0]) + fabsf(g.t[1]) + fabsf(g.t[2]);
    const float e = 1e-5f * mag;
    const float f0 = -g.df0n, f1 = -g.df1n, p0 = g.dp0, p1 = g.dp1;
    const float mu = 2.f / g.sx, mv = 2.f / g.sy;                       // two output pixels
    const float ur = ((float)out_w + 2.f) / g.sx, vr = ((float)out_h + 2.f) / g.sy;
    const float tl = e * (fabsf(f0) + fabsf(p0) + mu), tr = e * (fabsf(f0) + fabsf(p0) + ur);
    const float tt = e * (fabsf(f1) + fabsf(p1) + mv), tb = e * (fabsf(f1) + fabsf(p1) + vr);
    bool behind = true, left = true, right = true, top = true, bottom = true;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        behind = behind && (X[k] < -e);
        left = left && (f0 * Y[k] - (p0 + mu) * X[k] > tl);
        right = right && (-f0 * Y[k] - (ur - p0) * X[k] > tr);
        top = top && (f1 * Z[k] - (p1 + mv) * X[k] > tt);
        bottom = bottom && (-f1 * Z[k] - (vr - p1) * X[k] > tb);
    }
    return behind || left || right || top || bottom;
}

template <class G>
__device__ inline bool item_culled(const G&, int, int, int, int) { return false; }
template <>
__device__ inline bool item_culled<PairGeom>(const PairGeom& g, int bx, int by, int out_w, int out_h) {
    if (!g.bounds || g.dst_erp || g.wrap) return false;
    return box_outside(g, g.bounds + ((size_t)by * g.bounds_bx + bx) * 6, out_w, out_h);
}

// One thread per (view, source, item): append the items that may cover the view
// to the work list (code = (view * nsrc + source) * items + item).
// One block row per (view, source) pair (blockIdx.y, striding): the pair's
// descriptor is then uniform and read with scalar loads.  Survivors are
// appended to the work list with one atomic per 1024-thread block (the waves'
// counts are combined in LDS): the list's counter is a single address, and
// same-address device atomics are serialised at the memory side.
constexpr int kCullThreads = 1024;

template <class G>
__global__ void __launch_bounds__(kCullThreads) k_bin_cull(const G* __restrict__ geoms,
                                                           const ViewDesc* __restrict__ views, BatchDims d,
                                                           RasterWS ws) {
    __shared__ uint32_t s_cnt[kCullThreads / 64];
    __shared__ uint32_t s_base;
    const uint32_t per = d.bins_x * d.bins_y;
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npairs = (uint32_t)d.nsrc * (uint32_t)d.nviews;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (uint32_t vs = blockIdx.y; vs < npairs; vs += gridDim.y) {
        const size_t gid = (size_t)vs * per + item;
        bool keep = false;
        if (item < per) {
            const int v = (int)(vs / (uint32_t)d.nsrc);
            const G& g = geoms[vs];
            const int bx = (int)(item % d.bins_x), by = (int)(item / d.bins_x);
            if (by * kBinQY < g.sh - 1 && bx * kBinQX < g.sw - 1) keep = !item_culled(g, bx, by, views[v].w, views[v].h);
        }
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < kCullThreads / 64; ++w) tot += s_cnt[w];
            s_base = tot ? atomicAdd(&ws.status[4], tot) : 0u;
        }
        __syncthreads();
        uint32_t base = s_base;
        for (int w = 0; w < wave; ++w) base += s_cnt[w];
        uint32_t wi = 0xffffffffu;
        if (keep) {
            const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
            const uint32_t pos = base + (uint32_t)__popcll(m & lt);
            ws.work[pos] = (uint32_t)gid;
            if (pos < d.range_cap) wi = pos;
        }
        if (item < per) ws.imap[gid] = wi;
        __syncthreads();   // s_cnt / s_base free for the next pair
    }
}

// ------------------------------------------------------------------ binning
//
// Bins are compact: a counting pass (k_bin<COUNT>) finds every surviving
// quad's tile range and counts the (quad, tile) entries per (tile, source)
// bin; an exclusive scan turns the counts into bin offsets; fill passes write
// the quad codes, taking slots through one fill cursor per bin.  For the first
// range_cap work items the counting pass keeps each quad's tile range (and the
// item's vertex window and triangle validities, for the raster), so their fill
// pass (k_fill_kept) needs no vertex work; later items re-evaluate their
// vertices in their fill pass (k_bin<FILL>).  The order of a bin's entries is
// free: the raster's keys commute and the in-order replay sorts them.  When
// the entries exceed the list capacity the fill passes and the raster are
// skipped and the needed capacity is reported.
//
// Blocks first gather their entries per bin in an LDS table (a few to a few
// dozen distinct tiles per 64x4 quad block), then take one global atomic per
// bin: device-scope atomics are performed at the memory side, so their number
// and the chains of dependent ones set these kernels' time.  The counting
// pass needs no atomic's return value (its atomics do not hold the block);
// the fill passes take one returning atomic per bin.

constexpr uint32_t kRangeNone = 0xffffffffu;   // no surviving triangle
constexpr uint32_t kRangeBig = 0xfffffffeu;    // range too wide to pack: re-evaluated

__device__ inline uint32_t pack_range(int tx0, int tx1, int ty0, int ty1) {
    const int ntx = tx1 - tx0, nty = ty1 - ty0;
    if (ntx > 14 || nty > 14) return kRangeBig;
    return (uint32_t)tx0 | ((uint32_t)ty0 << 12) | ((uint32_t)ntx << 24) | ((uint32_t)nty << 28);
}

__device__ inline void unpack_range(uint32_t r, int& tx0, int& tx1, int& ty0, int& ty1) {
    tx0 = (int)(r & 0xfffu);
    ty0 = (int)((r >> 12) & 0xfffu);
    tx1 = tx0 + (int)((r >> 24) & 0xfu);
    ty1 = ty0 + (int)(r >> 28);
}

// Vertex (row lr, column lc) of a kept work item's vertex block (the counting
// pass evaluated it with vertex(); the bits are the same).
__device__ inline float3 kept_vertex(const float* vb, int lr, int lc) {
    const float* p = vb + 3 * (lr * (kBinQX + 1) + lc);
    return make_float3(p[0], p[1], p[2]);
}

// Work index of the kept work item holding quad (i, j) of pair vs, or ~0.
// (always ~0 when kept items store no vertex windows)
__device__ inline uint32_t kept_index(const RasterWS& ws, const BatchDims& d, uint32_t vs, int i, int j) {
    if (!d.kept_geom) return 0xffffffffu;
    const uint32_t item = (uint32_t)(i / kBinQY) * d.bins_x + (uint32_t)(j / kBinQX);
    return ws.imap[(size_t)vs * (d.bins_x * d.bins_y) + item];
}

// Vertex block of quad (i, j) of pair vs when its work item was kept, else null.
__device__ inline const float* kept_block(const RasterWS& ws, const BatchDims& d, uint32_t vs, int i, int j) {
    const uint32_t wi = kept_index(ws, d, vs, i, j);
    return wi == 0xffffffffu ? nullptr : ws.vbuf + (size_t)wi * (3 * kBinVerts);
}

// Validities of triangles T1, T2 of quad (i, j) of kept item wi (the counting
// pass evaluated them exactly; only alive triangles' values are meaningful).
__device__ inline float2 kept_validity(const RasterWS& ws, uint32_t wi, int i, int j) {
    return ws.tvq[(size_t)wi * 256 + (uint32_t)((i % kBinQY) * kBinQX + (j % kBinQX))];
}

constexpr uint32_t kTableEmpty = 0xffffffffu;

// Slot of global tile t in the block's LDS table, inserting it; -1 when the
// table is full of other tiles (keys are never removed while an item runs,
// so a tile refused once stays absent and find_slot agrees).
__device__ inline int table_slot(uint32_t* keys, uint32_t t) {
    uint32_t h = (t * 2654435761u) >> 25;
    for (int probe = 0; probe < kBinTable; ++probe) {
        const uint32_t k = keys[h];
        if (k == t) return (int)h;
        if (k == kTableEmpty) {
            const uint32_t old = atomicCAS(&keys[h], kTableEmpty, t);
            if (old == kTableEmpty || old == t) return (int)h;
        }
        h = (h + 1) & (kBinTable - 1);
    }
    return -1;
}

__device__ inline int find_slot(const uint32_t* keys, uint32_t t) {
    uint32_t h = (t * 2654435761u) >> 25;
    for (int probe = 0; probe < kBinTable; ++probe) {
        const uint32_t k = keys[h];
        if (k == t) return (int)h;
        if (k == kTableEmpty) return -1;
        h = (h + 1) & (kBinTable - 1);
    }
    return -1;
}

// Tile range [tx0, tx1] x [ty0, ty1] touched by the narrowed bounding boxes
// of a quad's surviving triangles; false when none can cover a pixel.
__device__ inline bool quad_tiles(const Quad& Q, int shift, int& tx0, int& tx1, int& ty0, int& ty1) {
    if (!Q.alive) return false;
    int xmin = INT_MAX, xmax = INT_MIN, ymin = INT_MAX, ymax = INT_MIN;
    for (int k = 0; k < 4; ++k)
        if (Q.alive & (1u << k)) {
            int xa = Q.t[k].xmin, xb = Q.t[k].xmax, ya = Q.t[k].ymin, yb = Q.t[k].ymax;
            tri_tight(Q.t[k], xa, xb, ya, yb);   // as quad_range_tv: the pixels it can cover
            if (xa > xb || ya > yb) continue;
            xmin = min(xmin, xa); xmax = max(xmax, xb);
            ymin = min(ymin, ya); ymax = max(ymax, yb);
        }
    if (xmin > xmax) return false;
    tx0 = xmin >> shift; tx1 = xmax >> shift; ty0 = ymin >> shift; ty1 = ymax >> shift;
    return true;
}

// Block-cooperative walk over the tiles of the threads' quad ranges: entry e
// of the block's concatenated ranges goes to thread e % 256, which finds the
// owning quad by a binary search of the block-wide inclusive scan of the
// ranges' tile counts.  One quad with a wide range (a stretched triangle
// across a depth edge, a quad near an equirectangular pole) then costs every
// lane of the block a few iterations instead of holding its wave.  Ranges of
// at most kLaneTiles tiles (nearly all) are walked by their own lane, and a
// block without a wider one skips the scan.
struct RangeWalk {
    uint32_t incl[256];   // inclusive prefix of the tile counts
    uint32_t org[256];    // tx0 | ty0 << 16
    uint32_t wid[256];    // tx1 - tx0 + 1
    uint32_t wsum[4];     // wave totals
};

constexpr uint32_t kLaneTiles = 4;   // ranges up to this many tiles stay with their own lane

// Publishes this thread's range (n = 0: none); all 256 threads call it, and
// the walk may start after the __syncthreads() that follows.
__device__ inline void range_publish(RangeWalk& W, int tid, bool on, int tx0, int tx1, int ty0, int ty1) {
    const uint32_t n = on ? (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1) : 0u;
    const int lane = tid & 63, wave = tid >> 6;
    uint32_t x = n;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) W.wsum[wave] = x;
    W.org[tid] = on ? ((uint32_t)tx0 | ((uint32_t)ty0 << 16)) : 0u;
    W.wid[tid] = on ? (uint32_t)(tx1 - tx0 + 1) : 1u;
    __syncthreads();
    for (int w = 0; w < wave; ++w) x += W.wsum[w];
    W.incl[tid] = x;
}

template <class F>
__device__ inline void range_walk(const RangeWalk& W, int tid, F fn) {
    const uint32_t total = W.incl[255];
    for (uint32_t e = (uint32_t)tid; e < total; e += 256u) {
        int lo = 0, hi = 255;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (W.incl[mid] > e) hi = mid; else lo = mid + 1;
        }
        const uint32_t l = e - (lo ? W.incl[lo - 1] : 0u);
        const uint32_t o = W.org[lo], w = W.wid[lo];
        fn(lo, (int)(o & 0xffffu) + (int)(l % w), (int)(o >> 16) + (int)(l / w));
    }
}

// FLAT: blocks stride over the kept work items (counting pass only);
// otherwise persistent blocks stride over the items after the kept ones.
template <class G, bool FILL, bool FLAT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_bin(const G* __restrict__ geoms, const ViewDesc* __restrict__ views,
                                             BatchDims d, RasterWS ws) {
    __shared__ float3 V[kBinQY + 1][kBinQX + 1];
    __shared__ uint32_t t_key[kBinTable];
    __shared__ uint32_t t_n[kBinTable];      // entries per bin
    __shared__ uint32_t t_base[kBinTable];   // fill write cursor per bin (fill pass)
    __shared__ RangeWalk walk;
    __shared__ uint32_t t_used;              // used table slots (kept items)
    const int tid = threadIdx.x;
    const uint32_t per = d.bins_x * d.bins_y;
    if (FILL && ws.status[1]) return;   // list capacity exceeded: nothing to fill
    const uint32_t nwork = ws.status[4];
    // Work items after the kept ones are alike in cost: a static stride over
    // them keeps every block busy without a shared work counter.  The fill
    // pass of the kept items is k_fill_kept.
    static_assert(!(FILL && FLAT), "kept items are filled by k_fill_kept");
    const uint64_t lim = FLAT ? min(nwork, d.range_cap) : nwork;
    const uint64_t step = gridDim.x;
    for (uint64_t wi64 = (FLAT ? 0u : d.range_cap) + blockIdx.x; wi64 < lim; wi64 += step) {
        const uint32_t wi = (uint32_t)wi64;
        const uint32_t code = ws.work[wi];
        const uint32_t item = code % per, vs = code / per;
        const int s = (int)(vs % (uint32_t)d.nsrc), v = (int)(vs / (uint32_t)d.nsrc);
        const G& g = geoms[vs];
        const ViewDesc& vd = views[v];
        const int i0 = (int)(item / d.bins_x) * kBinQY, j0 = (int)(item % d.bins_x) * kBinQX;
        const int i = i0 + tid / kBinQX, j = j0 + tid % kBinQX;
        const bool stored = !FILL && wi < d.range_cap;   // same for the whole block
        const bool stored_geom = stored && d.kept_geom;
        if (tid < kBinTable) {
            t_key[tid] = kTableEmpty;
            t_n[tid] = 0;
        }
        bool more = false;
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        // a kept item also records its quads' triangle validities for the raster
        // (exact; this pass waits on memory far more than it computes)
        auto quad_range = [&](auto vtx) {
            if (stored_geom) {
                float2 tv;
                more = quad_range_tv(g, i, j, vtx, vd.w, vd.h, d.tile_shift, tx0, tx1, ty0, ty1, &tv);
                ws.tvq[(size_t)wi * 256 + tid] = tv;
            } else {
                more = quad_range_tv(g, i, j, vtx, vd.w, vd.h, d.tile_shift, tx0, tx1, ty0, ty1, nullptr);
            }
        };
        {
            float* vb = stored_geom ? ws.vbuf + (size_t)wi * (3 * kBinVerts) : nullptr;
            for (int k = tid; k < kBinVerts; k += 256) {
                const int rr = k / (kBinQX + 1), c = k % (kBinQX + 1);
                const int ii = i0 + rr, jj = j0 + c;
                if (ii < g.sh && jj < g.sw) {
                    const float3 x = vertex(g, ii, jj);
                    V[rr][c] = x;
                    if (vb) { vb[3 * k] = x.x; vb[3 * k + 1] = x.y; vb[3 * k + 2] = x.z; }
                }
            }
            __syncthreads();   // vertices loaded, table cleared
            if (i < g.sh - 1 && j < g.sw - 1)
                quad_range([&](int ii, int jj) -> float3 {
                    const int rr = ii - i0, c = jj - j0;
                    if (rr >= 0 && rr <= kBinQY && c >= 0 && c <= kBinQX) return V[rr][c];
                    return vertex(g, ii, jj);
                });
        }
        auto tile_of = [&](int tx, int ty) { return vd.tile_base + (uint32_t)(ty * vd.tiles_x + tx); };
        auto bin_of = [&](uint32_t t) { return (size_t)t * d.nsrc + s; };
        // the range of a kept quad, for its fill pass
        if (!FILL && stored)
            ws.ranges[(size_t)wi * 256 + tid] = more ? pack_range(tx0, tx1, ty0, ty1) : kRangeNone;
        // entries: narrow ranges by their own lane, wide ones dealt to the block
        auto quad_of = [&](int lane) {
            return ((uint32_t)(i0 + lane / kBinQX) << 16) | (uint32_t)(j0 + lane % kBinQX);
        };
        const bool wide = more && (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1) > kLaneTiles;
        const bool any_wide = __syncthreads_or(wide);
        if (any_wide) {
            range_publish(walk, tid, wide, tx0, tx1, ty0, ty1);
            __syncthreads();   // ranges scanned
        }
        auto each_tile = [&](auto fn) {
            if (more && !wide)
                for (int ty = ty0; ty <= ty1; ++ty)
                    for (int tx = tx0; tx <= tx1; ++tx) fn(tid, tx, ty);
            if (any_wide) range_walk(walk, tid, fn);
        };
        each_tile([&](int lane, int tx, int ty) {
            const uint32_t t = tile_of(tx, ty);
            const int k = table_slot(t_key, t);
            if (k >= 0) {
                atomicAdd(&t_n[k], 1u);
            } else if (FILL) {   // table full: a slot of its own
                const size_t bin = bin_of(t);
                ws.lists[(size_t)ws.offsets[bin] + atomicAdd(&ws.fill[bin], 1u)] = quad_of(lane);
            } else {
                atomicAdd(&ws.counts[bin_of(t)], 1u);
            }
        });
        __syncthreads();   // table complete
        if (tid < kBinTable && t_n[tid]) {
            const size_t bin = bin_of(t_key[tid]);
            if (FILL) t_base[tid] = ws.offsets[bin] + atomicAdd(&ws.fill[bin], t_n[tid]);
            else atomicAdd(&ws.counts[bin], t_n[tid]);
        }
        if (stored) {
            // the item's table for its fill pass: the used slots (slot, count,
            // tile) in any order, their number first
            if (tid == 0) t_used = 0u;
            __syncthreads();
            uint2* kt = ws.ktab + (size_t)wi * (1 + kBinTable);
            if (tid < kBinTable && t_n[tid])
                kt[1 + atomicAdd(&t_used, 1u)] = make_uint2((uint32_t)tid | (t_n[tid] << 8), t_key[tid]);
            __syncthreads();
            if (tid == 0) kt[0] = make_uint2(t_used, 0u);
        }
        if (FILL) {
            __syncthreads();   // bases known
            each_tile([&](int lane, int tx, int ty) {
                const int k = find_slot(t_key, tile_of(tx, ty));
                if (k >= 0) ws.lists[atomicAdd(&t_base[k], 1u)] = quad_of(lane);
            });
        }
        __syncthreads();   // LDS free for the next item
    }
}

// Fill pass over the kept work items, a block per item (striding when there
// are more items than blocks) and a thread per quad: entries take slots
// through the item's LDS bin table, rebuilt from what the counting pass
// stored, and one fill-cursor atomic per bin.  A quad whose range was too
// wide to pack is re-binned from the kept vertices.
template <class G>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_fill_kept(const G* __restrict__ geoms, const ViewDesc* __restrict__ views,
                                                   BatchDims d, RasterWS ws) {
    __shared__ uint32_t t_key[kBinTable];
    __shared__ uint32_t t_base[kBinTable];
    __shared__ RangeWalk walk;
    if (ws.status[1]) return;   // list capacity exceeded: nothing to fill
    const uint32_t lim = min(ws.status[4], d.range_cap);
    const int tid = threadIdx.x;
    const uint32_t per = d.bins_x * d.bins_y;
    for (uint32_t wi = blockIdx.x; wi < lim; wi += gridDim.x) {
        if (tid < kBinTable) t_key[tid] = kTableEmpty;
        const uint32_t code = ws.work[wi];
        const uint32_t item = code % per, vs = code / per;
        const int s = (int)(vs % (uint32_t)d.nsrc), v = (int)(vs / (uint32_t)d.nsrc);
        const ViewDesc& vd = views[v];
        const int i0 = (int)(item / d.bins_x) * kBinQY, j0 = (int)(item % d.bins_x) * kBinQX;
        const int i = i0 + tid / kBinQX, j = j0 + tid % kBinQX;
        const uint32_t r = ws.ranges[(size_t)wi * 256 + tid];
        bool more = false;
        int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
        if (r == kRangeBig) {
            const G& g = geoms[vs];
            const float* vb = d.kept_geom ? ws.vbuf + (size_t)wi * (3 * kBinVerts) : nullptr;
            Quad Q;
            quad_build(Q, g, i, j,
                       [&](int ii, int jj) -> float3 {
                           const int rr = ii - i0, c = jj - j0;
                           if (vb && rr >= 0 && rr <= kBinQY && c >= 0 && c <= kBinQX) return kept_vertex(vb, rr, c);
                           return vertex(g, ii, jj);
                       },
                       vd.w, vd.h, false);
            more = quad_tiles(Q, d.tile_shift, tx0, tx1, ty0, ty1);
        } else if (r != kRangeNone) {
            unpack_range(r, tx0, tx1, ty0, ty1);
            more = true;
        }
        // entries (narrow ranges by their own lane, wide ones dealt to the
        // block): LDS bin table, one fill-cursor atomic per bin
        auto tile_of = [&](int tx, int ty) { return vd.tile_base + (uint32_t)(ty * vd.tiles_x + tx); };
        auto quad_of = [&](int lane) {
            return ((uint32_t)(i0 + lane / kBinQX) << 16) | (uint32_t)(j0 + lane % kBinQX);
        };
        const bool wide = more && (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1) > kLaneTiles;
        const bool any_wide = __syncthreads_or(wide);   // also orders the table clear
        if (any_wide) {
            range_publish(walk, tid, wide, tx0, tx1, ty0, ty1);
            __syncthreads();   // ranges scanned
        }
        auto each_tile = [&](auto fn) {
            if (more && !wide)
                for (int ty = ty0; ty <= ty1; ++ty)
                    for (int tx = tx0; tx <= tx1; ++tx) fn(tid, tx, ty);
            if (any_wide) range_walk(walk, tid, fn);
        };
        // the counting pass's table of this item, at the same slots (so
        // find_slot sees the same probe sequences), with one fill-cursor
        // atomic per used slot; a tile that found the table full there is
        // absent here too and takes a slot of its own per entry
        {
            const uint2* kt = ws.ktab + (size_t)wi * (1 + kBinTable);
            const uint32_t nused = kt[0].x;
            if (tid < (int)nused) {
                const uint2 e = kt[1 + tid];
                const uint32_t slot = e.x & 0xffu, n = e.x >> 8;
                t_key[slot] = e.y;
                const size_t bin = (size_t)e.y * d.nsrc + s;
                t_base[slot] = ws.offsets[bin] + atomicAdd(&ws.fill[bin], n);
            }
        }
        __syncthreads();   // table and bases known
        each_tile([&](int lane, int tx, int ty) {
            const uint32_t t = tile_of(tx, ty);
            const int k = find_slot(t_key, t);
            if (k >= 0) {
                ws.lists[atomicAdd(&t_base[k], 1u)] = quad_of(lane);
            } else {
                const size_t bin = (size_t)t * d.nsrc + s;
                ws.lists[(size_t)ws.offsets[bin] + atomicAdd(&ws.fill[bin], 1u)] = quad_of(lane);
            }
        });
        __syncthreads();   // table free for the next item
    }
}

// Exclusive scan of the bin counts into bin offsets (64-bit block sums; the
// total, saturated to 32 bits, goes to status[2], and status[1] is set when
// it exceeds the list capacity).
constexpr int kScanPer = 2048;   // bins per scan block (256 threads x 8)

__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* __restrict__ counts, uint32_t* offsets,
                                                     size_t n, unsigned long long* bsum) {
    __shared__ unsigned long long part[256];
    const size_t base = (size_t)blockIdx.x * kScanPer + (size_t)threadIdx.x * 8;
    uint32_t c[8];
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        c[k] = base + k < n ? counts[base + k] : 0u;
        sum += c[k];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {   // inclusive Hillis-Steele scan
        const unsigned long long add = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - sum;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (base + k < n) offsets[base + k] = (uint32_t)run;
        run += c[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = part[255];
}

__global__ void __launch_bounds__(1024) k_scan_top(unsigned long long* bsum, uint32_t nb, uint32_t cap,
                                                   uint32_t* status) {
    __shared__ unsigned long long part[1024];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0ull;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const unsigned long long x = b < nb ? bsum[b] : 0ull;
        part[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const unsigned long long add = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nb) bsum[b] = carry + part[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        status[2] = carry > 0xffffffffull ? 0xffffffffu : (uint32_t)carry;
        if (carry > (unsigned long long)cap) status[1] = 1u;
    }
}

__global__ void k_scan_add(uint32_t* offsets, size_t n, const unsigned long long* __restrict__ bsum) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) offsets[i] += (uint32_t)bsum[i / kScanPer];
}

size_t scan_blocks(size_t nbins) { return (nbins + kScanPer - 1) / kScanPer; }

template <class G>
hipError_t launch_bin(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws, hipStream_t s) {
    const size_t total = (size_t)d.bins_x * d.bins_y * d.nsrc * d.nviews;
    const uint32_t npairs = (uint32_t)d.nsrc * (uint32_t)d.nviews;
    hipLaunchKernelGGL(k_bin_cull<G>,
                       dim3((d.bins_x * d.bins_y + kCullThreads - 1) / kCullThreads, npairs < 65535u ? npairs : 65535u),
                       dim3(kCullThreads), 0, s, geoms, views, d, ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t blocks = total < (size_t)kBinBlocks ? total : (size_t)kBinBlocks;
    // kept items: persistent blocks striding over them (blocks past the
    // surviving items return at once).  A block's set-up, including the few
    // registers these kernels spill at occupancy 8, is paid once per block, so
    // a few items per block beat a block per item; fewer, longer-lived blocks
    // lose to the tail when items differ in cost (equirectangular targets:
    // items near a pole deal thousands of tiles), so the plan picks the count.
    // A counter that blocks take items from is slower still: every take is a
    // same-address atomic.
    const size_t fb = d.flat_blocks ? d.flat_blocks : (uint32_t)kFlatBlocks;
    const unsigned flat = (unsigned)(d.range_cap < fb ? d.range_cap : fb);
    if (d.range_cap)
        hipLaunchKernelGGL((k_bin<G, false, true>), dim3(flat), dim3(256), 0, s, geoms, views, d, ws);
    if (total > d.range_cap)
        hipLaunchKernelGGL((k_bin<G, false, false>), dim3((unsigned)blocks), dim3(256), 0, s, geoms, views, d, ws);
    const size_t nbins = (size_t)d.ntiles * d.nsrc;
    const size_t nb = scan_blocks(nbins);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(256), 0, s, ws.counts, ws.offsets, nbins, ws.bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, ws.bsum, (uint32_t)nb, d.cap, ws.status);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, s, ws.offsets, nbins,
                       ws.bsum);
    if (d.range_cap)
        hipLaunchKernelGGL(k_fill_kept<G>, dim3(flat), dim3(256), 0, s, geoms, views, d, ws);
    if (total > d.range_cap)
        hipLaunchKernelGGL((k_bin<G, true, false>), dim3((unsigned)blocks), dim3(256), 0, s, geoms, views, d, ws);
    return hipGetLastError();
}

// ------------------------------------------------------------ tile raster

// The lane's index in its wave, computed where it is called (not merged with
// other reads of it, so it need not stay live in a VGPR in between).
__device__ inline int lane_now() {
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// Winner's colour, depth and validity at pixel (x, y) for slot `slot` of
// source quad (i, j): the state the fold holds after accepting that candidate.
// FAST: vertex radii and the validity take only their fast float-float paths,
// and the result is false when one of them is undecided (the caller defers the
// pixel; the coverage pass, which evaluated the same triangle, deferred it
// already); otherwise they are exact and the result is true.
template <bool FAST, class G>
__device__ __forceinline__ bool resolve(const G& g, int i, int j, int slot, int x, int y, float3& c, float& dep,
                               float& shape, int out_w, int out_h, const float* vb) {
    int r[3], cc[3];
    float3 v[3], col[3];
    bool und = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) slot_vertex(slot, k, i, j, g.sw, r[k], cc[k]);
    color3(g, r, cc, col);   // issued first: its loads overlap the vertices'
#pragma unroll
    for (int k = 0; k < 3; ++k)
        v[k] = (vb && slot < 2) ? kept_vertex(vb, i % kBinQY + (r[k] - i), j % kBinQX + (cc[k] - j))
                               : vertex(g, r[k], cc[k], FAST ? &und : nullptr);
    Tri T;
    if (FAST) {
        tri_setup_geom(T, v[0], v[1], v[2], out_w, out_h);
        if (!tri_validity_fast(T, g.rescale, T.tv) || und) return false;
    } else {
        tri_setup(T, v[0], v[1], v[2], out_w, out_h, g.rescale);
    }
    float l1, l2, l3;
    tri_pixel(T, x, y, l1, l2, l3);
    dep = tri_depth(T, l1, l2, l3);
    shape = T.tv;
    c = tri_color(col[0], col[1], col[2], l1, l2, l3);
    return true;
}

// The same from the corner vertices and validity the coverage pass cached
// (identical float operations, without re-evaluating vertices or norms).
template <class G>
__device__ __forceinline__ void resolve_cached(const G& g, int i, int j, int slot, float3 a, float3 b, float3 cv, float tv,
                                      int x, int y, float3& c, float& dep, float& shape) {
    int r[3], cc[3];
    float3 col[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) slot_vertex(slot, k, i, j, g.sw, r[k], cc[k]);
    color3(g, r, cc, col);
    Tri T;
    tri_load(T, a, b, cv);
    T.tv = tv;
    float l1, l2, l3;
    tri_pixel(T, x, y, l1, l2, l3);
    dep = tri_depth(T, l1, l2, l3);
    shape = tv;
    c = tri_color(col[0], col[1], col[2], l1, l2, l3);
}

// key2 value of a pixel that must take the in-order replay (above every
// score: scores are finite positive floats).
constexpr uint32_t kDeferKey = 0xffffffffu;

// Pass over one triangle of a binned quad for the tile at (tx0, ty0): per
// covered pixel, push the candidate key (score bits << 32 | candidate id)
// into key1 and the displaced score into key2.  Candidates with non-finite
// inputs or extreme depths mark the pixel for the in-order replay instead.
// Returns the triangle's validity, or 0 when no pixel of the tile can be
// covered (then it is not evaluated: the reference computes it before the
// bounding box test, but it never causes a return, and a triangle that covers
// no pixel is never a winner, so skipping it changes nothing).
template <int TILE>
__device__ __forceinline__ float raster_tri(float3 a, float3 b, float3 c, uint32_t id, bool prol_false_possible, int tx0,
                                   int ty0, int out_w, int out_h, float rescale, unsigned long long* key1,
                                   uint32_t* key2, bool undecided, float tv_known) {
    Tri T;
    if (!tri_setup_geom(T, a, b, c, out_w, out_h)) return 0.f;
    const bool prol = (a.z > 0.f) || (b.z > 0.f) || (c.z > 0.f);
    bool special = !finite3(a) || !finite3(b) || !finite3(c) || !is_finite(T.den) ||
                   (prol_false_possible && !prol) || undecided;
    int xa = max(T.xmin, tx0), xb = min(T.xmax, tx0 + TILE - 1);
    int ya = max(T.ymin, ty0), yb = min(T.ymax, ty0 + TILE - 1);
    if (!special) tri_tight(T, xa, xb, ya, yb);
    if (xa > xb || ya > yb) return 0.f;   // no pixel of this tile can be covered
    auto push = [&](int x, int y) {
        float n1, n2;
        tri_num(T, x, y, n1, n2);
        float l1, l2, l3;
        if (!tri_cover(T, n1, n2, l1, l2, l3)) return;
        const int p = (y - ty0) * TILE + (x - tx0);
        const float dd = tri_depth(T, l1, l2, l3);
        if (special || !depth_regular(dd)) {
            atomicMax(&key2[p], kDeferKey);
            return;
        }
        // the score only orders candidates outside the ambiguity band:
        // v_rcp_f32 (1 ulp) keeps its error within 5 * 2^-24 relative
        const float f = T.tv * __builtin_amdgcn_rcpf(dd * dd * dd);
        const unsigned long long k = ((unsigned long long)f2u(f) << 32) | id;
        const unsigned long long old = atomicMax(&key1[p], k);
        // the candidate that is not (or no longer) the maximum goes to the runner-up slot
        atomicMax(&key2[p], (uint32_t)((old < k ? old : k) >> 32));
    };
    if (xb - xa < 8 && yb - ya < 8) {
        // Two steps, so that lanes with few candidates do not wait through the
        // exact test of lanes with large boxes: the division-free rejection
        // marks the candidate pixels of the box (8 x 8 bit mask), then only
        // those take the exact coverage test and the key updates.
        uint64_t mask = 0ull;
        for (int y = ya; y <= yb; ++y)
            for (int x = xa; x <= xb; ++x) {
                float n1, n2;
                tri_num(T, x, y, n1, n2);
                if (!tri_surely_outside(T, n1, n2)) mask |= 1ull << ((y - ya) * 8 + (x - xa));
            }
        if (!mask) return 0.f;   // no candidate pixel: the validity is never needed
        // a validity kept by the counting pass (> 0), else the fast path; one it
        // leaves undecided defers the pixels the triangle covers
        if (tv_known > 0.f) T.tv = tv_known;
        else if (!tri_validity_fast(T, rescale, T.tv)) special = true;
        while (mask) {
            const int bit = __builtin_ctzll(mask);
            mask &= mask - 1ull;
            push(xa + (bit & 7), ya + (bit >> 3));
        }
    } else {
        if (tv_known > 0.f) T.tv = tv_known;
        else if (!tri_validity_fast(T, rescale, T.tv)) special = true;
        for (int y = ya; y <= yb; ++y)
            for (int x = xa; x <= xb; ++x) {
                float n1, n2;
                tri_num(T, x, y, n1, n2);
                if (!tri_surely_outside(T, n1, n2)) push(x, y);
            }
    }
    return T.tv;
}

// Per-phase LDS copy of what the coverage pass looked up for each binned
// quad: the quad code and its kept work item (~0: none), so that the winners'
// corner vertices and validities are read from the kept blocks without the
// work-item map.
template <int CAP>
struct RasterCache {
    static constexpr int cap = CAP;
    uint32_t ij[CAP];
    uint32_t wi[CAP];
};

// Candidate ids: (entry << 2) | slot, entry = index of the quad in the phase's
// concatenated bins.  A non-deferred pixel's winner beats every other
// candidate by more than the float test's error, so any unique id works; equal
// scores are deferred to the in-order replay.
template <int TILE, class G, class Cache>
__device__ __forceinline__ void raster_quad(const G& g, uint32_t ij, uint32_t e, int tx0, int ty0, int out_w, int out_h,
                                   unsigned long long* key1, uint32_t* key2, Cache* C,
                                   const RasterWS& ws, const BatchDims& d, uint32_t vs) {
    int i, j;
    unpack_ij(ij, i, j);
    const uint32_t wi = kept_index(ws, d, vs, i, j);
    const float* vb = wi == 0xffffffffu ? nullptr : ws.vbuf + (size_t)wi * (3 * kBinVerts);
    const float2 tvk = vb ? kept_validity(ws, wi, i, j) : make_float2(0.f, 0.f);
    const int li = i % kBinQY, lj = j % kBinQX;
    // vertices evaluated here take the fast radius path; an undecided one
    // defers the pixels of the quad's triangles (kept vertices are exact)
    bool und = false;
    auto vtx = [&](int di, int dj) -> float3 {
        return vb ? kept_vertex(vb, li + di, lj + dj) : vertex(g, i + di, j + dj, &und);
    };
    // vertices are fetched just before their triangles need them; keys
    // commute, so only the set of candidates matters here
    const float3 v01 = vtx(0, 1), v10 = vtx(1, 0);
    const bool gate = v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f;
    const uint32_t s0 = e << 2;
    {
        const float3 v00 = vtx(0, 0);
        if (gate && v00.z > 0.f && v00.x > 0.f)
            raster_tri<TILE>(v00, v01, v10, s0, false, tx0, ty0, out_w, out_h, g.rescale, key1, key2, und,
                             tvk.x);
    }
    {
        const float3 v11 = vtx(1, 1);
        if (gate && v11.z > 0.f && v11.x > 0.f)
            raster_tri<TILE>(v11, v10, v01, s0 | 1u, false, tx0, ty0, out_w, out_h, g.rescale, key1, key2, und,
                             tvk.y);
    }
    if (gate && g.wrap && j == 0) {
        const float3 v00 = vtx(0, 0);
        const float3 l0 = vertex(g, i, g.sw - 1, &und), l1 = vertex(g, i + 1, g.sw - 1, &und);
        raster_tri<TILE>(l0, v00, l1, s0 | 2u, true, tx0, ty0, out_w, out_h, g.rescale, key1, key2, und, 0.f);
        raster_tri<TILE>(v10, l1, v00, s0 | 3u, true, tx0, ty0, out_w, out_h, g.rescale, key1, key2, und, 0.f);
    }
    if (C) {
        C->ij[e] = ij;
        C->wi[e] = wi;
    }
}

template <int MODE>
__device__ inline void write_pixel(const ViewDesc& vd, size_t o, const BlendState& st, float3 rc, float rd, float rs,
                                   bool color = true, bool quality = true) {
    if (MODE == MODE_RAW) {
        vd.color[3 * o] = rc.x; vd.color[3 * o + 1] = rc.y; vd.color[3 * o + 2] = rc.z;
        vd.depth[o] = rd;
        vd.validity[o] = rs;
    } else {
        if (color) { vd.color[3 * o] = st.c.x; vd.color[3 * o + 1] = st.c.y; vd.color[3 * o + 2] = st.c.z; }
        if (vd.depth) vd.depth[o] = st.d;
        if (quality) vd.quality[o] = st.q;
        if (vd.validity) vd.validity[o] = st.v;
    }
}

// write_color (image_writing.cpp:142-203) of one pixel of a tile, from the
// blended colour in registers: Y here, U and V by the top-left pixel of each
// 2x2 block from its wave neighbours (a tile row is at most half a wave, so
// the block's four pixels are lanes l, l+1, l+TILE, l+TILE+1 of one wave).  The
// same quantisation and chroma rule as k_encode_views; tiles start at even
// coordinates and encoded views have even sizes, so a block is inside or
// outside as a whole.
template <int TILE>
__device__ inline void encode_tile_pixel(const ViewDesc& vd, float3 c, int px, int py, bool inside, int lx, int ly) {
    const unsigned maxv = (1u << vd.bitdepth) - 1u;
    const bool u8 = vd.bitdepth <= 8;
    uint32_t qy = quantize_u32(c.x, maxv), qu = quantize_u32(c.y, maxv), qv = quantize_u32(c.z, maxv);
    // the plane's element type truncates, as (T) does in k_encode_views
    if (u8) { qy &= 0xffu; qu &= 0xffu; qv &= 0xffu; } else { qy &= 0xffffu; qu &= 0xffffu; qv &= 0xffffu; }
    static_assert(2 * TILE <= 64, "a 2x2 block's two rows must lie in one wave");
    const uint32_t su = qu + (uint32_t)__shfl_down((int)qu, 1) + (uint32_t)__shfl_down((int)qu, TILE) +
                        (uint32_t)__shfl_down((int)qu, TILE + 1);
    const uint32_t sv = qv + (uint32_t)__shfl_down((int)qv, 1) + (uint32_t)__shfl_down((int)qv, TILE) +
                        (uint32_t)__shfl_down((int)qv, TILE + 1);
    if (!inside) return;
    const size_t o = (size_t)py * vd.ew + px;
    const bool corner = ((lx | ly) & 1) == 0;
    const size_t oc = (size_t)(py >> 1) * (vd.ew >> 1) + (px >> 1);
    if (u8) {
        ((uint8_t*)vd.y)[o] = (uint8_t)qy;
        if (corner) { ((uint8_t*)vd.u)[oc] = (uint8_t)((su + 2) >> 2); ((uint8_t*)vd.v)[oc] = (uint8_t)((sv + 2) >> 2); }
    } else {
        ((uint16_t*)vd.y)[o] = (uint16_t)qy;
        if (corner) {
            ((uint16_t*)vd.u)[oc] = (uint16_t)((su + 2) >> 2);
            ((uint16_t*)vd.v)[oc] = (uint16_t)((sv + 2) >> 2);
        }
    }
}

// MULTI: more sources than one phase holds (the phases loop, and the blend
// state lives across them); otherwise one phase, which keeps the coverage pass
// free of the blend state's registers.
// TILE: the tile edge; a block has a thread per pixel (256 or 1024 threads),
// and the caches scale with the tile's area.
template <class G, int MODE, bool MULTI, int TILE>
__global__ void __launch_bounds__(TILE * TILE) __attribute__((amdgpu_waves_per_eu(MULTI ? (TILE > kTile ? 4 : 6) : 8))) k_raster(const G* __restrict__ geoms, const ViewDesc* __restrict__ views,
                                                BatchDims d, RasterWS ws) {
    constexpr int NT = TILE * TILE;
    constexpr int AREA = NT / (kTile * kTile);
    constexpr int GRP = MULTI ? kRasterGroupMulti : kRasterGroup;   // sources per phase
    __shared__ unsigned long long key1[GRP][NT];
    __shared__ uint32_t key2[GRP][NT];
    __shared__ uint32_t pref[GRP + 1];
    __shared__ uint32_t boff[GRP];   // list offsets of the phase's bins
    __shared__ int s_deferred;   // any pixel of the tile deferred
    __shared__ int s_hole;       // any pixel of the tile a hole (quality not > 0)
    // one phase: a smaller cache keeps a 16 x 16 block within 20 KB of LDS (8 blocks per CU)
    __shared__ RasterCache<(MULTI ? kRasterCache : kRasterCacheOne) * AREA> cache;
    const uint32_t T = tile_of_block(blockIdx.x, d.ntiles);
    if (T >= d.ntiles || ws.status[1]) return;   // bins not filled: capacity reported to the host
    const int v = find_view(views, d.nviews, T);
    const ViewDesc& vd = views[v];
    const int tile = (int)(T - vd.tile_base);
    const int tid = threadIdx.x;
    const int tx0 = (tile % vd.tiles_x) * TILE, ty0 = (tile / vd.tiles_x) * TILE;
    // The thread's index is re-derived after the coverage pass (wave base in
    // an SGPR + lane id) instead of being held in a VGPR through it.
    const int wave_base = __builtin_amdgcn_readfirstlane(tid & ~63);
    const G* gv = geoms + (size_t)v * d.nsrc;
    bool deferred = false;
    BlendState st;
    float3 rc = make_float3(0.f, 0.f, 0.f);
    float rd = __builtin_inff(), rs = 0.f;
    for (int s0 = 0; s0 < (MULTI ? d.nsrc : 1); s0 += GRP) {
        const int gs = min(GRP, d.nsrc - s0);
        for (int k = 0; k < gs; ++k) {
            key1[k][tid] = 0ull;
            key2[k][tid] = 0u;
        }
        if (tid == 0) {
            if (s0 == 0) { s_deferred = 0; s_hole = 0; }
            // the phase's counts and offsets loaded together, then summed
            uint32_t cnt[GRP], off[GRP];
#pragma unroll
            for (int k = 0; k < GRP; ++k) {
                const size_t b = (size_t)T * d.nsrc + s0 + (k < gs ? k : 0);
                cnt[k] = ws.counts[b];
                off[k] = ws.offsets[b];
            }
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < GRP; ++k)
                if (k < gs) {
                    pref[k] = acc;
                    boff[k] = off[k];
                    acc += cnt[k];
                }
            pref[gs] = acc;
        }
        __syncthreads();
        const uint32_t N = __builtin_amdgcn_readfirstlane(pref[gs]);   // uniform: kept in an SGPR
        const bool cached = N <= (uint32_t)cache.cap;
        // a wave-uniform trip count; the entry index is re-derived each trip
        const uint32_t trips = (N + (uint32_t)NT - 1u) / (uint32_t)NT;
        for (uint32_t it = 0; it < trips; ++it) {
            const uint32_t e = it * (uint32_t)NT + (uint32_t)(wave_base + lane_now());
            if (e >= N) continue;
            int k = 0;
            while (k + 1 < gs && e >= pref[k + 1]) ++k;
            const uint32_t ij = ws.lists[(size_t)boff[k] + (e - pref[k])];
            raster_quad<TILE>(gv[s0 + k], ij, e, tx0, ty0, vd.w, vd.h, key1[k], key2[k], cached ? &cache : nullptr,
                              ws, d, (uint32_t)(v * d.nsrc + s0 + k));
        }
        __syncthreads();
        const int t2 = wave_base + lane_now();
        const int px = tx0 + (t2 % TILE), py = ty0 + (t2 / TILE);
        const bool inside = px < vd.w && py < vd.h;
        if (inside) {
            for (int k = 0; k < gs; ++k) {
                const int s = s0 + k;
                if (deferred) break;
                const unsigned long long kw = key1[k][t2];
                const uint32_t k2 = key2[k][t2];
                bool amb = k2 == kDeferKey;
                if (!amb && kw != 0ull && k2 != 0u) {
                    const float fw = u2f((uint32_t)(kw >> 32)), f2 = u2f(k2);
                    amb = f2 >= fw - fw * RVS_AMBIGUITY;
                }
                if (amb) { deferred = true; break; }
                float3 c = make_float3(0.f, 0.f, 0.f);
                float dep = __builtin_inff(), shape = 0.f;
                if (kw != 0ull) {
                    const uint32_t id = (uint32_t)kw, e = id >> 2;
                    const int slot = (int)(id & 3u);
                    const G& g = gv[s];
                    int i, j;
                    const float* vb = nullptr;
                    uint32_t wk = 0xffffffffu;
                    if (cached) {
                        unpack_ij(cache.ij[e], i, j);
                        if (slot < 2) wk = cache.wi[e];
                        if (wk != 0xffffffffu) vb = ws.vbuf + (size_t)wk * (3 * kBinVerts);
                    }
                    if (vb) {
                        // T1: v00 v01 v10; T2: v11 v10 v01 (corner offsets of slot_vertex);
                        // the validity the counting pass kept (the coverage pass used it)
                        const int li = i % kBinQY, lj = j % kBinQX;
                        const float3 a = kept_vertex(vb, li + slot, lj + slot);
                        const float3 b = kept_vertex(vb, li + slot, lj + 1 - slot);
                        const float3 cc = kept_vertex(vb, li + 1 - slot, lj + slot);
                        const float2 tvk = kept_validity(ws, wk, i, j);
                        resolve_cached(g, i, j, slot, a, b, cc, slot ? tvk.y : tvk.x, px, py, c, dep, shape);
                    } else {
                        const uint32_t ij = cached ? cache.ij[e] : ws.lists[(size_t)boff[k] + (e - pref[k])];
                        unpack_ij(ij, i, j);
                        if (!resolve<true>(g, i, j, slot, px, py, c, dep, shape, vd.w, vd.h,
                                           kept_block(ws, d, (uint32_t)(v * d.nsrc + s), i, j))) {
                            deferred = true;
                            break;
                        }
                    }
                }
                if (MODE == MODE_RAW) {
                    rc = c; rd = dep; rs = shape;
                } else if (s == 0) {
                    blend_first(st, c, dep, shape);
                } else {
                    blend_step(st, c, dep, shape, vd.blend_exp, vd.blend_inv);
                }
            }
        }
        __syncthreads();
    }
    const int t3 = wave_base + lane_now();
    const int lx = t3 % TILE, ly = t3 / TILE;
    const int px = tx0 + lx, py = ty0 + ly;
    const bool inside = px < vd.w && py < vd.h;
    const size_t o = (size_t)py * vd.w + px;
    // a tile with a deferred pixel or a hole (quality not > 0: the inpainting
    // mask) is finished by the replay / inpainting and encoded after them;
    // any other tile of a view that wants planes is encoded here
    if (deferred) s_deferred = 1;
    if (MODE == MODE_BLEND && inside && !deferred && !(st.q > 0.f)) s_hole = 1;
    __syncthreads();
    const bool enc = MODE == MODE_BLEND && vd.enc_raster && !s_deferred && !s_hole;
    if (inside && !deferred) {
        // an internal colour map of a tile encoded here is read only on the
        // tile's border (by the inpainting of a neighbouring tile)
        const bool border = lx == 0 || ly == 0 || lx == TILE - 1 || ly == TILE - 1;
        write_pixel<MODE>(vd, o, st, rc, rd, rs, !(enc && vd.partial_color && !border),
                          !(vd.partial_quality && !s_hole && !s_deferred));
    }
    if (enc) encode_tile_pixel<TILE>(vd, st.c, px, py, inside, lx, ly);
    if (t3 == 0) ws.tile_enc[T] = (enc ? 1 : 0) | (MODE == MODE_BLEND && (s_hole || s_deferred) ? 2 : 0);
    // the replay reads the per-pixel flags of the listed tiles only
    if (s_deferred) {
        if (inside) ws.deferred[vd.px_base + o] = deferred ? 1 : 0;
        if (t3 == 0) {
            const uint32_t k = atomicAdd(&ws.status[0], 1u);
            ws.deferred_tiles[k] = T;
        }
    }
}

// ------------------------------------------------------ in-order replay

struct ReplayEntry {
    Tri t[4];
    uint32_t alive, prol, q;   // q: quad code (row << 16 | column)
};

// A candidate (triangle covering a deferred pixel) recorded for the in-order
// fold: quad code, slot | prol | next-in-list, depth, validity.
struct Cand {
    uint32_t ij;
    uint32_t meta;   // bits 0-15: next record (0xffff = end), 16-17: slot, 18: prol
    float d, tv;
};
constexpr int kCandPool = kFallbackSortCap * 4 / (int)sizeof(Cand);   // records sharing the sort buffer
static_assert(kCandPool < 0xffff, "candidate links are 16-bit");

// Sorts L[0, n) ascending in place, one block: a bitonic network in its
// all-ascending form (the first step of each merge compares e with its mirror
// e ^ (k - 1), the later ones with e ^ jj, always putting the smaller value at
// the lower index), so the positions from n up to the next power of two act
// as +inf and their comparisons are skipped.  Global memory written by the
// block's threads is read back by other threads of the block after each
// barrier.
__device__ void block_sort_global(uint32_t* L, uint32_t n) {
    uint32_t np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (uint32_t k = 2; k <= np2; k <<= 1)
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            const uint32_t mate = jj == (k >> 1) ? k - 1 : jj;
            for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
                const uint32_t p = e ^ mate;
                if (p > e && p < n) {
                    const uint32_t a = L[e], b = L[p];
                    if (a > b) { L[e] = b; L[p] = a; }
                }
            }
            __syncthreads();
        }
}

// Deferred pixels of one tile and source, in the reference's traversal order.
// Fast path: the block rasterises the source's bin in parallel and appends
// every candidate covering a deferred pixel of the current pixel range to
// that pixel's list in LDS; each pixel then folds its own candidates in
// ascending (quad, slot) order.  When a range's lists overflow the pool, the
// range is split (down to a single pixel); only a single pixel with more
// candidates than the pool falls back to sorting the bin (in LDS, or in place
// in the entry list when it is longer than the LDS buffer; only this block
// reads this bin after the raster) and scanning it in order.
template <class G>
__device__ void replay_source(const G& g, uint32_t* L, uint32_t n, int tx0, int ty0, int out_w, int out_h,
                              bool mine, int px, int py, const uint8_t* dmask, uint32_t* sorted, ReplayEntry* rec,
                              uint32_t* pool_n, FoldState& fs, const RasterWS& ws, const BatchDims& d, uint32_t vs) {
    const int tid = threadIdx.x;
    // vertex (ii, jj) of the neighbourhood of a quad whose work item keeps
    // the vertex window vb (null: none) starting at row r0, column c0: from
    // the window (the same bits as vertex()), else evaluated
    auto kept_or_eval = [&](const float* vb, int r0, int c0, int ii, int jj) -> float3 {
        const int rr = ii - r0, cc = jj - c0;
        if (vb && rr >= 0 && rr <= kBinQY && cc >= 0 && cc <= kBinQX) return kept_vertex(vb, rr, cc);
        return vertex(g, ii, jj);
    };
    Cand* pool = reinterpret_cast<Cand*>(sorted);
    int* head = reinterpret_cast<int*>(rec);
    const int pme = (py - ty0) * kTile + (px - tx0);
    fold_init(fs);
    int lo = 0, span = kTile * kTile;
    bool scan_all = false;
    while (lo < kTile * kTile) {
        const int hi = min(lo + span, kTile * kTile);
        head[tid] = -1;
        if (tid == 0) *pool_n = 0u;
        __syncthreads();
        for (uint32_t e = tid; e < n; e += 256) {
            const uint32_t q = L[e];
            int i, j;
            unpack_ij(q, i, j);
            const float* vb = kept_block(ws, d, vs, i, j);
            const int r0 = i - i % kBinQY, c0 = j - j % kBinQX;
            auto vtx = [&](int ii, int jj) { return kept_or_eval(vb, r0, c0, ii, jj); };
            quad_tris(g, i, j, vtx, out_w, out_h, [&](int k, Tri& Tr, bool prol) {
                bool have_tv = false;   // the validity, evaluated for the first candidate
                int xa = max(Tr.xmin, tx0), xb = min(Tr.xmax, tx0 + kTile - 1);
                int ya = max(Tr.ymin, ty0), yb = min(Tr.ymax, ty0 + kTile - 1);
                if (is_finite(Tr.A0) && is_finite(Tr.A1) && is_finite(Tr.B0) && is_finite(Tr.B1) &&
                    is_finite(Tr.C0) && is_finite(Tr.C1) && is_finite(Tr.den))
                    tri_tight(Tr, xa, xb, ya, yb);
                for (int y = ya; y <= yb; ++y)
                    for (int x = xa; x <= xb; ++x) {
                        const int p = (y - ty0) * kTile + (x - tx0);
                        if (p < lo || p >= hi || !dmask[p]) continue;
                        float n1, n2;
                        tri_num(Tr, x, y, n1, n2);
                        if (tri_surely_outside(Tr, n1, n2)) continue;
                        float l1, l2, l3;
                        if (!tri_cover(Tr, n1, n2, l1, l2, l3)) continue;
                        if (!have_tv) {
                            Tr.tv = tri_validity(Tr, g.rescale);
                            have_tv = true;
                        }
                        const uint32_t r = atomicAdd(pool_n, 1u);
                        if (r >= (uint32_t)kCandPool) continue;
                        const int nx = atomicExch(&head[p], (int)r);
                        Cand& C = pool[r];
                        C.ij = q;
                        C.meta = (uint32_t)(nx < 0 ? 0xffff : nx) | ((uint32_t)k << 16) | ((prol ? 1u : 0u) << 18);
                        C.d = tri_depth(Tr, l1, l2, l3);
                        C.tv = Tr.tv;
                    }
            });
        }
        __syncthreads();
        const bool overflow = *pool_n > (uint32_t)kCandPool;
        __syncthreads();
        if (overflow) {
            if (span > 1) {
                span = max(1, span / 4);   // retry this range in smaller pieces
                continue;
            }
            scan_all = true;               // one pixel with more candidates than the pool
            break;
        }
        if (mine && pme >= lo && pme < hi) {
            bool first = true;
            uint64_t last = 0;
            while (true) {
                uint64_t best = ~0ull;
                int bi = -1;
                for (int r = head[pme]; r >= 0;) {
                    const Cand& C = pool[r];
                    const uint64_t key = ((uint64_t)C.ij << 2) | ((C.meta >> 16) & 3u);
                    if ((first || key > last) && key < best) { best = key; bi = r; }
                    const uint32_t nx = C.meta & 0xffffu;
                    r = nx == 0xffffu ? -1 : (int)nx;
                }
                if (bi < 0) break;
                const Cand& C = pool[bi];
                fold_step(fs, C.d, C.tv, (C.meta >> 18) & 1u, C.ij, (int)((C.meta >> 16) & 3u));
                last = best;
                first = false;
            }
        }
        __syncthreads();
        lo = hi;
        span = min(span * 2, kTile * kTile);
    }
    if (!scan_all) return;
    // the remaining deferred pixels (from lo on) scan the whole bin in order
    mine = mine && pme >= lo;
    if (mine) fold_init(fs);
    // overflow: sort the bin (bitonic, ascending quad code = traversal order) and scan it
    const uint32_t* S = sorted;
    if (n <= (uint32_t)kFallbackSortCap) {
        uint32_t np2 = 1;
        while (np2 < n) np2 <<= 1;
        for (uint32_t e = tid; e < np2; e += 256) sorted[e] = e < n ? L[e] : 0xffffffffu;
        __syncthreads();
        for (uint32_t k = 2; k <= np2; k <<= 1)
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                for (uint32_t e = tid; e < np2; e += 256) {
                    const uint32_t p = e ^ jj;
                    if (p > e) {
                        const uint32_t a = sorted[e], b = sorted[p];
                        const bool asc = (e & k) == 0;
                        if ((a > b) == asc) { sorted[e] = b; sorted[p] = a; }
                    }
                }
                __syncthreads();
            }
    } else {
        block_sort_global(L, n);
        S = L;
    }
    for (uint32_t b = 0; b < n; b += kFallbackChunk) {
        if (tid < kFallbackChunk && b + tid < n) {
            const uint32_t q = S[b + tid];
            int i, j;
            unpack_ij(q, i, j);
            Quad Q;
            const float* vb = kept_block(ws, d, vs, i, j);
            const int r0 = i - i % kBinQY, c0 = j - j % kBinQX;
            auto vtx = [&](int ii, int jj) { return kept_or_eval(vb, r0, c0, ii, jj); };
            quad_build(Q, g, i, j, vtx, out_w, out_h);
            ReplayEntry& R = rec[tid];
            for (int k = 0; k < 4; ++k) R.t[k] = Q.t[k];
            R.alive = Q.alive; R.prol = Q.prol; R.q = q;
        }
        __syncthreads();
        if (mine) {
            const uint32_t m = min((uint32_t)kFallbackChunk, n - b);
            for (uint32_t e = 0; e < m; ++e) {
                const ReplayEntry& R = rec[e];
                for (int k = 0; k < 4; ++k) {
                    if (!(R.alive & (1u << k))) continue;
                    const Tri& Tr = R.t[k];
                    if (px < Tr.xmin || px > Tr.xmax || py < Tr.ymin || py > Tr.ymax) continue;
                    float l1, l2, l3;
                    if (!tri_pixel(Tr, px, py, l1, l2, l3)) continue;
                    fold_step(fs, tri_depth(Tr, l1, l2, l3), Tr.tv, (R.prol >> k) & 1u, R.q, k);
                }
            }
        }
        __syncthreads();
    }
}

// Deferred tiles [first, n_deferred): one block per tile folds every source
// in order and blends (the tiles before `first` take the per-source pair of
// kernels below).
template <class G, int MODE>
__global__ void __launch_bounds__(256) k_fallback(const G* __restrict__ geoms, const ViewDesc* __restrict__ views,
                                                  BatchDims d, RasterWS ws, uint32_t first) {
    __shared__ uint32_t sorted[kFallbackSortCap];
    __shared__ ReplayEntry rec[kFallbackChunk];
    __shared__ uint8_t dmask[kTile * kTile];
    __shared__ uint32_t pool_n[1];   // candidate records
    const int tid = threadIdx.x;
    const uint32_t ndef = ws.status[1] ? 0u : ws.status[0];
    for (uint32_t ti = first + blockIdx.x; ti < ndef; ti += gridDim.x) {
        const uint32_t T = ws.deferred_tiles[ti];
        const int v = find_view(views, d.nviews, T);
        const ViewDesc& vd = views[v];
        const G* gv = geoms + (size_t)v * d.nsrc;
        const int tile = (int)(T - vd.tile_base);
        const int edge = 1 << d.tile_shift;
        // the tile in sub-tiles of kTile x kTile pixels, a thread per pixel
        // (each folds the whole bin of the tile, clipped to its sub-tile)
        const int nsub = edge / kTile;
        for (int sub = 0; sub < nsub * nsub; ++sub) {
        const int tx0 = (tile % vd.tiles_x) * edge + (sub % nsub) * kTile;
        const int ty0 = (tile / vd.tiles_x) * edge + (sub / nsub) * kTile;
        const int px = tx0 + (tid % kTile), py = ty0 + (tid / kTile);
        const bool inside = px < vd.w && py < vd.h;
        const size_t o = inside ? (size_t)py * vd.w + px : 0;
        const bool mine = inside && ws.deferred[vd.px_base + o];
        dmask[tid] = mine ? 1 : 0;
        // a sub-tile of a larger tile may hold no deferred pixel (the barrier
        // also publishes dmask)
        if (!__syncthreads_or(mine ? 1 : 0)) continue;
        BlendState st;
        float3 rc = make_float3(0.f, 0.f, 0.f);
        float rd = __builtin_inff(), rs = 0.f;
        for (int s = 0; s < d.nsrc; ++s) {
            const G& g = gv[s];
            const size_t bin = (size_t)T * d.nsrc + s;
            const uint32_t n = ws.counts[bin];
            FoldState fs;
            replay_source(g, ws.lists + ws.offsets[bin], n, tx0, ty0, vd.w, vd.h, mine, px, py, dmask, sorted, rec,
                          pool_n, fs, ws, d, (uint32_t)(v * d.nsrc + s));
            __syncthreads();
            if (mine) {
                float3 c = make_float3(0.f, 0.f, 0.f);
                float dep = __builtin_inff(), shape = 0.f;
                if (fs.slot >= 0) {
                    float dd, sh;
                    int i, j;
                    unpack_ij(fs.ij, i, j);
                    resolve<false>(g, i, j, fs.slot, px, py, c, dd, sh, vd.w, vd.h, nullptr);
                    dep = fs.new_d;
                    shape = fs.shape;
                }
                if (MODE == MODE_RAW) {
                    rc = c; rd = dep; rs = shape;
                } else if (s == 0) {
                    blend_first(st, c, dep, shape);
                } else {
                    blend_step(st, c, dep, shape, vd.blend_exp, vd.blend_inv);
                }
            }
        }
        if (mine) write_pixel<MODE>(vd, o, st, rc, rd, rs);
        __syncthreads();
        }
    }
}

// The first d.rbuf_tiles deferred tiles, sources in parallel: a block per
// (tile, source) folds that source's bin for the tile's deferred pixels (the
// same replay_source and resolve as k_fallback) and stores each pixel's
// colour, depth and validity in ws.rbuf; k_fallback_blend then blends them in
// source order.  A tile's chain of sources is the replay's critical path when
// few tiles are deferred.
template <class G>
__global__ void __launch_bounds__(256) k_fallback_src(const G* __restrict__ geoms,
                                                      const ViewDesc* __restrict__ views, BatchDims d, RasterWS ws) {
    __shared__ uint32_t sorted[kFallbackSortCap];
    __shared__ ReplayEntry rec[kFallbackChunk];
    __shared__ uint8_t dmask[kTile * kTile];
    __shared__ uint32_t pool_n[1];
    const int tid = threadIdx.x;
    const uint32_t ndef = ws.status[1] ? 0u : min(ws.status[0], d.rbuf_tiles);
    const int edge = 1 << d.tile_shift;
    const int nsub = edge / kTile;
    for (uint32_t job = blockIdx.x; job < ndef * (uint32_t)d.nsrc; job += gridDim.x) {
        const uint32_t ti = job / (uint32_t)d.nsrc;
        const int s = (int)(job % (uint32_t)d.nsrc);
        const uint32_t T = ws.deferred_tiles[ti];
        const int v = find_view(views, d.nviews, T);
        const ViewDesc& vd = views[v];
        const G& g = geoms[(size_t)v * d.nsrc + s];
        const int tile = (int)(T - vd.tile_base);
        float* out = ws.rbuf + ((size_t)ti * d.nsrc + s) * ((size_t)edge * edge) * 5;
        for (int sub = 0; sub < nsub * nsub; ++sub) {
            const int tx0 = (tile % vd.tiles_x) * edge + (sub % nsub) * kTile;
            const int ty0 = (tile / vd.tiles_x) * edge + (sub / nsub) * kTile;
            const int px = tx0 + (tid % kTile), py = ty0 + (tid / kTile);
            const bool inside = px < vd.w && py < vd.h;
            const size_t o = inside ? (size_t)py * vd.w + px : 0;
            const bool mine = inside && ws.deferred[vd.px_base + o];
            dmask[tid] = mine ? 1 : 0;
            if (!__syncthreads_or(mine ? 1 : 0)) continue;
            const size_t bin = (size_t)T * d.nsrc + s;
            FoldState fs;
            replay_source(g, ws.lists + ws.offsets[bin], ws.counts[bin], tx0, ty0, vd.w, vd.h, mine, px, py, dmask,
                          sorted, rec, pool_n, fs, ws, d, (uint32_t)(v * d.nsrc + s));
            if (mine) {
                float3 c = make_float3(0.f, 0.f, 0.f);
                float dep = __builtin_inff(), shape = 0.f;
                if (fs.slot >= 0) {
                    float dd, sh;
                    int i, j;
                    unpack_ij(fs.ij, i, j);
                    resolve<false>(g, i, j, fs.slot, px, py, c, dd, sh, vd.w, vd.h, nullptr);
                    dep = fs.new_d;
                    shape = fs.shape;
                }
                float* r = out + (size_t)(sub * kTile * kTile + tid) * 5;
                r[0] = c.x; r[1] = c.y; r[2] = c.z; r[3] = dep; r[4] = shape;
            }
            __syncthreads();
        }
    }
}

template <class G>
__global__ void __launch_bounds__(256) k_fallback_blend(const ViewDesc* __restrict__ views, BatchDims d,
                                                        RasterWS ws) {
    const int tid = threadIdx.x;
    const uint32_t ndef = ws.status[1] ? 0u : min(ws.status[0], d.rbuf_tiles);
    const int edge = 1 << d.tile_shift;
    const int nsub = edge / kTile;
    for (uint32_t ti = blockIdx.x; ti < ndef; ti += gridDim.x) {
        const uint32_t T = ws.deferred_tiles[ti];
        const int v = find_view(views, d.nviews, T);
        const ViewDesc& vd = views[v];
        const int tile = (int)(T - vd.tile_base);
        for (int sub = 0; sub < nsub * nsub; ++sub) {
            const int px = (tile % vd.tiles_x) * edge + (sub % nsub) * kTile + (tid % kTile);
            const int py = (tile / vd.tiles_x) * edge + (sub / nsub) * kTile + (tid / kTile);
            if (px >= vd.w || py >= vd.h) continue;
            const size_t o = (size_t)py * vd.w + px;
            if (!ws.deferred[vd.px_base + o]) continue;
            BlendState st;
            for (int s = 0; s < d.nsrc; ++s) {
                const float* r = ws.rbuf + (((size_t)ti * d.nsrc + s) * ((size_t)edge * edge) +
                                            (size_t)(sub * kTile * kTile + tid)) * 5;
                const float3 c = make_float3(r[0], r[1], r[2]);
                if (s == 0) blend_first(st, c, r[3], r[4]);
                else blend_step(st, c, r[3], r[4], vd.blend_exp, vd.blend_inv);
            }
            write_pixel<MODE_BLEND>(vd, o, st, make_float3(0.f, 0.f, 0.f), 0.f, 0.f);
        }
    }
}

template <class G, int MODE>
hipError_t launch_raster(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws,
                         hipStream_t s) {
    const unsigned blocks = 8u * ((d.ntiles + 7u) / 8u);
    const bool multi = d.nsrc > kRasterGroup;
    if (d.tile_shift == 5) {
        constexpr int NT = kTileBig * kTileBig;
        if (multi)
            hipLaunchKernelGGL((k_raster<G, MODE, true, kTileBig>), dim3(blocks), dim3(NT), 0, s, geoms, views, d, ws);
        else
            hipLaunchKernelGGL((k_raster<G, MODE, false, kTileBig>), dim3(blocks), dim3(NT), 0, s, geoms, views, d, ws);
    } else {
        constexpr int NT = kTile * kTile;
        if (multi)
            hipLaunchKernelGGL((k_raster<G, MODE, true, kTile>), dim3(blocks), dim3(NT), 0, s, geoms, views, d, ws);
        else
            hipLaunchKernelGGL((k_raster<G, MODE, false, kTile>), dim3(blocks), dim3(NT), 0, s, geoms, views, d, ws);
    }
    return hipGetLastError();
}

template <class G, int MODE>
hipError_t launch_replay(const G* geoms, const ViewDesc* views, const BatchDims& d, const RasterWS& ws,
                         hipStream_t s) {
    const unsigned fb_blocks = d.ntiles < (uint32_t)kFallbackBlocks ? d.ntiles : (uint32_t)kFallbackBlocks;
    uint32_t first = 0;
    if (MODE == MODE_BLEND && d.nsrc > 1 && d.rbuf_tiles > 0) {
        // the first rbuf_tiles deferred tiles: sources in parallel, then the blend
        const uint64_t jobs = (uint64_t)d.rbuf_tiles * (uint64_t)d.nsrc;
        const unsigned sb = (unsigned)(jobs < (uint64_t)kFallbackBlocks * 4 ? jobs : (uint64_t)kFallbackBlocks * 4);
        hipLaunchKernelGGL((k_fallback_src<G>), dim3(sb), dim3(256), 0, s, geoms, views, d, ws);
        const unsigned bb = d.rbuf_tiles < (uint32_t)kFallbackBlocks ? d.rbuf_tiles : (uint32_t)kFallbackBlocks;
        hipLaunchKernelGGL((k_fallback_blend<G>), dim3(bb), dim3(256), 0, s, views, d, ws);
        first = d.rbuf_tiles;
    }
    hipLaunchKernelGGL((k_fallback<G, MODE>), dim3(fb_blocks), dim3(256), 0, s, geoms, views, d, ws, first);
    return hipGetLastError();
}

template hipError_t launch_bin<PairGeom>(const PairGeom*, const ViewDesc*, const BatchDims&, const RasterWS&,
                                         hipStream_t);
template hipError_t launch_bin<ArrayGeom>(const ArrayGeom*, const ViewDesc*, const BatchDims&, const RasterWS&,
                                          hipStream_t);
template hipError_t launch_raster<PairGeom, MODE_BLEND>(const PairGeom*, const ViewDesc*, const BatchDims&,
                                                        const RasterWS&, hipStream_t);
template hipError_t launch_raster<ArrayGeom, MODE_RAW>(const ArrayGeom*, const ViewDesc*, const BatchDims&,
                                                       const RasterWS&, hipStream_t);
template hipError_t launch_replay<PairGeom, MODE_BLEND>(const PairGeom*, const ViewDesc*, const BatchDims&,
                                                        const RasterWS&, hipStream_t);
template hipError_t launch_replay<ArrayGeom, MODE_RAW>(const ArrayGeom*, const ViewDesc*, const BatchDims&,
                                                       const RasterWS&, hipStream_t);

}  // namespace rvs_k
