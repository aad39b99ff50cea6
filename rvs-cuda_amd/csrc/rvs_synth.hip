// rvs_synth.hip -- triangle binning, tile rasterisation and blending for gfx950.
//
// Reference semantics: detail::transform_trianglesMethod (rvs/src/transform.cpp:148-181)
// visits source quads in raster order and folds every covered target pixel
// sequentially: a candidate (depth d, triangle validity v) replaces the state
// (D, V) when fl(fl(fl(r*r)*r)*V) < v with r = fl(d/D), strict, so ties keep the
// earlier triangle.  Mathematically that is the argmax of v/d^3 with the lowest
// traversal index on ties, but the float test is not transitive for
// near-equal scores.  This file computes it in three steps:
//   1. k_bin: every source quad is appended to the 16x16 target tiles its
//      surviving triangles' (reference-exact) bounding boxes touch.
//   2. k_raster: one 256-thread block per tile, up to four sources per phase.
//      Every covered pixel takes two LDS 64-bit atomicMax: key1 = max of
//      (score bits << 32 | ~seq), key2 = max over the candidates that are not
//      the maximum (the loser of each key1 update).  A pixel whose runner-up
//      scores within 2^-18 of the winner, or that saw a candidate with
//      non-finite or extreme values, is deferred.  Every other pixel provably
//      ends the sequential fold on the argmax (the float test errs by less than
//      2^-21 relative), so its winner is re-evaluated and blended.
//   3. k_fallback: deferred pixels replay the reference fold literally, in
//      traversal order, over the tile's bins sorted in LDS.
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

// ------------------------------------------------------------------ quads

// The (up to four) triangles of source quad (i, j), with the reference's
// vertex gating (transform.cpp:161-176).  alive bit k set = colorize_triangle
// is called for slot k and gets past its early returns.
struct Quad {
    Tri t[4];
    uint32_t alive;
    uint32_t special;   // bit k: slot k has non-finite inputs or a seam prol edge case
    uint32_t prol;      // bit k: prol flag of slot k (transform.cpp:121-123)
};

__device__ inline bool finite3(float3 v) { return is_finite(v.x) && is_finite(v.y) && is_finite(v.z); }

__device__ inline void quad_add(Quad& Q, int slot, float3 a, float3 b, float3 c, int out_w, int out_h, float rescale) {
    Tri T;
    const bool ok = tri_setup(T, a, b, c, out_w, out_h, rescale);
    Q.t[slot] = T;
    if (!ok) return;
    Q.alive |= 1u << slot;
    const bool prol = (a.z > 0.f) || (b.z > 0.f) || (c.z > 0.f);
    if (prol) Q.prol |= 1u << slot;
    if (!finite3(a) || !finite3(b) || !finite3(c) || !is_finite(T.den) || !prol)
        Q.special |= 1u << slot;
}

template <class G, class VF>
__device__ inline void quad_build(Quad& Q, const G& g, int i, int j, VF vtx, int out_w, int out_h) {
    Q.alive = 0; Q.special = 0; Q.prol = 0;
    const float3 v00 = vtx(i, j), v01 = vtx(i, j + 1), v10 = vtx(i + 1, j), v11 = vtx(i + 1, j + 1);
    if (!(v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f)) return;
    if (v00.z > 0.f && v00.x > 0.f) quad_add(Q, 0, v00, v01, v10, out_w, out_h, g.rescale);
    if (v11.z > 0.f && v11.x > 0.f) quad_add(Q, 1, v11, v10, v01, out_w, out_h, g.rescale);
    if (g.wrap && j == 0) {
        const float3 l0 = vtx(i, g.sw - 1), l1 = vtx(i + 1, g.sw - 1);
        quad_add(Q, 2, l0, v00, l1, out_w, out_h, g.rescale);
        quad_add(Q, 3, v10, l1, v00, out_w, out_h, g.rescale);
    }
}

// ------------------------------------------------------------------ binning

template <class G>
__global__ void __launch_bounds__(256) k_bin(const G* __restrict__ geoms, RasterDims d, RasterWS ws) {
    const G& g = geoms[blockIdx.z];
    const int i0 = blockIdx.y * kBinQY, j0 = blockIdx.x * kBinQX;
    if (i0 >= g.sh - 1 || j0 >= g.sw - 1) return;
    __shared__ float3 V[kBinQY + 1][kBinQX + 1];
    const int tid = threadIdx.x;
    for (int k = tid; k < (kBinQY + 1) * (kBinQX + 1); k += 256) {
        const int r = k / (kBinQX + 1), c = k % (kBinQX + 1);
        const int i = i0 + r, j = j0 + c;
        if (i < g.sh && j < g.sw) V[r][c] = vertex(g, i, j);
    }
    __syncthreads();
    const int qi = tid / kBinQX, qj = tid % kBinQX;
    const int i = i0 + qi, j = j0 + qj;
    bool more = false;
    int tx0 = 0, tx1 = -1, ty0 = 0, ty1 = -1;
    uint32_t q = 0;
    if (i < g.sh - 1 && j < g.sw - 1) {
        Quad Q;
        auto vtx = [&](int ii, int jj) -> float3 {
            const int r = ii - i0, c = jj - j0;
            if (r >= 0 && r <= kBinQY && c >= 0 && c <= kBinQX) return V[r][c];
            return vertex(g, ii, jj);
        };
        quad_build(Q, g, i, j, vtx, d.out_w, d.out_h);
        if (Q.alive) {
            int xmin = INT_MAX, xmax = INT_MIN, ymin = INT_MAX, ymax = INT_MIN;
            for (int s = 0; s < 4; ++s)
                if (Q.alive & (1u << s)) {
                    xmin = min(xmin, Q.t[s].xmin); xmax = max(xmax, Q.t[s].xmax);
                    ymin = min(ymin, Q.t[s].ymin); ymax = max(ymax, Q.t[s].ymax);
                }
            tx0 = xmin / kTile; tx1 = xmax / kTile; ty0 = ymin / kTile; ty1 = ymax / kTile;
            more = true;
            q = (uint32_t)i * (uint32_t)(g.sw - 1) + (uint32_t)j;
        }
    }
    const uint32_t src = blockIdx.z;
    uint32_t* counts = ws.counts + (size_t)src * d.ntiles;
    uint32_t* lists = ws.lists + (size_t)src * d.ntiles * d.cap;
    const int lane = __lane_id();
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    int tx = tx0, ty = ty0;
    // Wave-aggregated appends: one atomic per distinct tile per round.
    while (__any(more)) {
        const uint32_t t = more ? (uint32_t)(ty * d.tiles_x + tx) : 0xffffffffu;
        const uint64_t act = __ballot(more);
        const int leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t tl = __shfl(t, leader);
        const bool mine = more && t == tl;
        const uint64_t m = __ballot(mine);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counts[tl], (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        if (mine) {
            const uint32_t pos = base + (uint32_t)__popcll(m & lt_mask);
            if (pos < d.cap)
                lists[(size_t)tl * d.cap + pos] = q;
            else
                atomicMax(&ws.status[1], 1u);
            if (++tx > tx1) { tx = tx0; if (++ty > ty1) more = false; }
        }
    }
}

template <class G>
hipError_t launch_bin(const G* geoms, int nsrc, int max_sw, int max_sh, const RasterDims& d,
                      const RasterWS& ws, hipStream_t s) {
    dim3 grid((max_sw - 1 + kBinQX - 1) / kBinQX, (max_sh - 1 + kBinQY - 1) / kBinQY, nsrc);
    hipLaunchKernelGGL(k_bin<G>, grid, dim3(256), 0, s, geoms, d, ws);
    return hipGetLastError();
}

// ------------------------------------------------------------ tile raster

template <class G>
__device__ inline void quad_of(const G& g, uint32_t q, int& i, int& j) {
    const uint32_t w1 = (uint32_t)(g.sw - 1);
    i = (int)(q / w1);
    j = (int)(q - (uint32_t)i * w1);
}

// Winner's colour, depth and validity at pixel (x, y) (the state the fold
// holds after accepting candidate `seq`).
template <class G>
__device__ inline void resolve(const G& g, uint32_t seq, int x, int y, float3& c, float& dep, float& shape,
                               int out_w, int out_h) {
    int i, j;
    quad_of(g, seq >> 2, i, j);
    const int slot = (int)(seq & 3u);
    int r[3], cc[3];
    float3 v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        slot_vertex(slot, k, i, j, g.sw, r[k], cc[k]);
        v[k] = vertex(g, r[k], cc[k]);
    }
    Tri T;
    tri_setup(T, v[0], v[1], v[2], out_w, out_h, g.rescale);
    float l1, l2, l3;
    tri_pixel(T, x, y, l1, l2, l3);
    dep = tri_depth(T, l1, l2, l3);
    shape = T.tv;
    c = tri_color(color(g, r[0], cc[0]), color(g, r[1], cc[1]), color(g, r[2], cc[2]), l1, l2, l3);
}

// Tile index of block b: each XCD (blocks are dealt round-robin over 8 XCDs)
// gets a contiguous band of tiles, so neighbouring tiles share that XCD's L2.
__device__ inline int tile_of_block(int b, int ntiles) {
    const int per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Pass over one triangle of a binned quad for the tile at (tx0, ty0): per
// covered pixel, push the candidate key into the top-2 keys of the pixel.
// Candidates with non-finite inputs or extreme depths mark the pixel for the
// in-order replay instead.
__device__ inline void raster_tri(float3 a, float3 b, float3 c, uint32_t seq, bool prol_false_possible, int tx0,
                                  int ty0, int out_w, int out_h, float rescale, unsigned long long* key1,
                                  unsigned long long* key2, int* flag) {
    Tri T;
    if (!tri_setup(T, a, b, c, out_w, out_h, rescale)) return;
    const bool prol = (a.z > 0.f) || (b.z > 0.f) || (c.z > 0.f);
    const bool special = !finite3(a) || !finite3(b) || !finite3(c) || !is_finite(T.den) ||
                         (prol_false_possible && !prol);
    const int xa = max(T.xmin, tx0), xb = min(T.xmax, tx0 + kTile - 1);
    const int ya = max(T.ymin, ty0), yb = min(T.ymax, ty0 + kTile - 1);
    const unsigned long long lo = (unsigned long long)(~seq);
    for (int y = ya; y <= yb; ++y)
        for (int x = xa; x <= xb; ++x) {
            float l1, l2, l3;
            if (!tri_pixel(T, x, y, l1, l2, l3)) continue;
            const int p = (y - ty0) * kTile + (x - tx0);
            const float dd = tri_depth(T, l1, l2, l3);
            if (special || !depth_regular(dd)) {
                flag[p] = 1;
                continue;
            }
            const float f = T.tv / (dd * dd * dd);
            const unsigned long long k = ((unsigned long long)f2u(f) << 32) | lo;
            const unsigned long long old = atomicMax(&key1[p], k);
            // the candidate that is not (or no longer) the maximum goes to the runner-up slot
            atomicMax(&key2[p], old < k ? old : k);
        }
}

template <class G>
__device__ inline void raster_quad(const G& g, uint32_t q, int tx0, int ty0, int out_w, int out_h,
                                   unsigned long long* key1, unsigned long long* key2, int* flag) {
    int i, j;
    quad_of(g, q, i, j);
    const float3 v00 = vertex(g, i, j), v01 = vertex(g, i, j + 1), v10 = vertex(g, i + 1, j);
    const float3 v11 = vertex(g, i + 1, j + 1);
    if (!(v01.z > 0.f && v10.z > 0.f && v01.x > 0.f && v10.x > 0.f)) return;
    const uint32_t s0 = q << 2;
    if (v00.z > 0.f && v00.x > 0.f)
        raster_tri(v00, v01, v10, s0, false, tx0, ty0, out_w, out_h, g.rescale, key1, key2, flag);
    if (v11.z > 0.f && v11.x > 0.f)
        raster_tri(v11, v10, v01, s0 | 1u, false, tx0, ty0, out_w, out_h, g.rescale, key1, key2, flag);
    if (g.wrap && j == 0) {
        const float3 l0 = vertex(g, i, g.sw - 1), l1 = vertex(g, i + 1, g.sw - 1);
        raster_tri(l0, v00, l1, s0 | 2u, true, tx0, ty0, out_w, out_h, g.rescale, key1, key2, flag);
        raster_tri(v10, l1, v00, s0 | 3u, true, tx0, ty0, out_w, out_h, g.rescale, key1, key2, flag);
    }
}

constexpr int kGroup = 4;   // sources rasterised together per block phase

template <class G, int MODE>
__global__ void __launch_bounds__(256) k_raster(const G* __restrict__ geoms, int nsrc, RasterDims d, RasterWS ws,
                                                float blend_exp, RasterOut out) {
    __shared__ unsigned long long key1[kGroup][kTile * kTile];
    __shared__ unsigned long long key2[kGroup][kTile * kTile];
    __shared__ int flag[kGroup][kTile * kTile];
    __shared__ uint32_t pref[kGroup + 1];
    const int tile = tile_of_block(blockIdx.x, d.ntiles);
    if (tile >= d.ntiles) return;
    const int tid = threadIdx.x;
    const int tx0 = (tile % d.tiles_x) * kTile, ty0 = (tile / d.tiles_x) * kTile;
    const int px = tx0 + (tid % kTile), py = ty0 + (tid / kTile);
    const bool inside = px < d.out_w && py < d.out_h;
    bool deferred = false;
    BlendState st;
    float3 rc = make_float3(0.f, 0.f, 0.f);
    float rd = __builtin_inff(), rs = 0.f;
    for (int s0 = 0; s0 < nsrc; s0 += kGroup) {
        const int gs = min(kGroup, nsrc - s0);
        for (int k = 0; k < gs; ++k) {
            key1[k][tid] = 0ull;
            key2[k][tid] = 0ull;
            flag[k][tid] = 0;
        }
        if (tid == 0) {
            uint32_t acc = 0;
            for (int k = 0; k < gs; ++k) {
                pref[k] = acc;
                acc += min(ws.counts[(size_t)(s0 + k) * d.ntiles + tile], d.cap);
            }
            pref[gs] = acc;
        }
        __syncthreads();
        const uint32_t N = pref[gs];
        for (uint32_t e = tid; e < N; e += 256) {
            int k = 0;
            while (k + 1 < gs && e >= pref[k + 1]) ++k;
            const uint32_t q = ws.lists[((size_t)(s0 + k) * d.ntiles + tile) * d.cap + (e - pref[k])];
            raster_quad(geoms[s0 + k], q, tx0, ty0, d.out_w, d.out_h, key1[k], key2[k], flag[k]);
        }
        __syncthreads();
        if (inside) {
            for (int k = 0; k < gs; ++k) {
                const int s = s0 + k;
                if (deferred) break;
                const unsigned long long kw = key1[k][tid];
                bool amb = flag[k][tid] != 0;
                if (!amb && kw != 0ull) {
                    const unsigned long long k2 = key2[k][tid];
                    if (k2 != 0ull) {
                        const float fw = u2f((uint32_t)(kw >> 32)), f2 = u2f((uint32_t)(k2 >> 32));
                        amb = f2 >= fw - fw * RVS_AMBIGUITY;
                    }
                }
                if (amb) { deferred = true; break; }
                float3 c = make_float3(0.f, 0.f, 0.f);
                float dep = __builtin_inff(), shape = 0.f;
                if (kw != 0ull) resolve(geoms[s], ~(uint32_t)kw, px, py, c, dep, shape, d.out_w, d.out_h);
                if (MODE == MODE_RAW) {
                    rc = c; rd = dep; rs = shape;
                } else if (s == 0) {
                    blend_first(st, c, dep, shape);
                } else {
                    blend_step(st, c, dep, shape, blend_exp);
                }
            }
        }
        __syncthreads();
    }
    if (inside) {
        const size_t o = (size_t)py * d.out_w + px;
        ws.deferred[o] = deferred ? 1 : 0;
        if (!deferred) {
            if (MODE == MODE_RAW) {
                out.color[3 * o] = rc.x; out.color[3 * o + 1] = rc.y; out.color[3 * o + 2] = rc.z;
                out.depth[o] = rd;
                out.validity[o] = rs;
            } else {
                out.color[3 * o] = st.c.x; out.color[3 * o + 1] = st.c.y; out.color[3 * o + 2] = st.c.z;
                if (out.depth) out.depth[o] = st.d;
                if (out.quality) out.quality[o] = st.q;
                if (out.validity) out.validity[o] = st.v;
            }
        }
    }
    if (__syncthreads_or(deferred) && tid == 0) {
        const uint32_t k = atomicAdd(&ws.status[0], 1u);
        ws.deferred_tiles[k] = (uint32_t)tile;
    }
}

// ------------------------------------------------------ in-order replay

struct ReplayEntry {
    Tri t[4];
    uint32_t alive, prol, q;
};

template <class G, int MODE>
__global__ void __launch_bounds__(256) k_fallback(const G* __restrict__ geoms, int nsrc, RasterDims d, RasterWS ws,
                                                  float blend_exp, RasterOut out) {
    __shared__ uint32_t sorted[kFallbackSortCap];
    __shared__ ReplayEntry rec[kFallbackChunk];
    const int tid = threadIdx.x;
    const uint32_t ndef = ws.status[0];
    for (uint32_t ti = blockIdx.x; ti < ndef; ti += gridDim.x) {
        const int tile = (int)ws.deferred_tiles[ti];
        const int tx0 = (tile % d.tiles_x) * kTile, ty0 = (tile / d.tiles_x) * kTile;
        const int px = tx0 + (tid % kTile), py = ty0 + (tid / kTile);
        const bool inside = px < d.out_w && py < d.out_h;
        const size_t o = inside ? (size_t)py * d.out_w + px : 0;
        const bool mine = inside && ws.deferred[o];
        BlendState st;
        float3 rc = make_float3(0.f, 0.f, 0.f);
        float rd = __builtin_inff(), rs = 0.f;
        for (int s = 0; s < nsrc; ++s) {
            const G& g = geoms[s];
            const uint32_t n = min(ws.counts[(size_t)s * d.ntiles + tile], d.cap);
            const uint32_t* L = ws.lists + ((size_t)s * d.ntiles + tile) * d.cap;
            if (n > (uint32_t)kFallbackSortCap) {
                if (tid == 0) atomicMax(&ws.status[3], n);
                __syncthreads();
                continue;
            }
            uint32_t np2 = 1;
            while (np2 < n) np2 <<= 1;
            for (uint32_t e = tid; e < np2; e += 256) sorted[e] = e < n ? L[e] : 0xffffffffu;
            __syncthreads();
            // bitonic sort ascending (quad index order = traversal order)
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
            FoldState fs;
            fold_init(fs);
            for (uint32_t b = 0; b < n; b += kFallbackChunk) {
                if (tid < kFallbackChunk && b + tid < n) {
                    const uint32_t q = sorted[b + tid];
                    int i, j;
                    quad_of(g, q, i, j);
                    Quad Q;
                    auto vtx = [&](int ii, int jj2) -> float3 { return vertex(g, ii, jj2); };
                    quad_build(Q, g, i, j, vtx, d.out_w, d.out_h);
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
                            const Tri& T = R.t[k];
                            if (px < T.xmin || px > T.xmax || py < T.ymin || py > T.ymax) continue;
                            float l1, l2, l3;
                            if (!tri_pixel(T, px, py, l1, l2, l3)) continue;
                            fold_step(fs, tri_depth(T, l1, l2, l3), T.tv, (R.prol >> k) & 1u, (R.q << 2) | (uint32_t)k);
                        }
                    }
                }
                __syncthreads();
            }
            if (mine) {
                float3 c = make_float3(0.f, 0.f, 0.f);
                float dep = __builtin_inff(), shape = 0.f;
                if (fs.seq != 0xffffffffu) {
                    float dd, sh;
                    resolve(g, fs.seq, px, py, c, dd, sh, d.out_w, d.out_h);
                    dep = fs.new_d;
                    shape = fs.shape;
                }
                if (MODE == MODE_RAW) {
                    rc = c; rd = dep; rs = shape;
                } else if (s == 0) {
                    blend_first(st, c, dep, shape);
                } else {
                    blend_step(st, c, dep, shape, blend_exp);
                }
            }
        }
        if (mine) {
            if (MODE == MODE_RAW) {
                out.color[3 * o] = rc.x; out.color[3 * o + 1] = rc.y; out.color[3 * o + 2] = rc.z;
                out.depth[o] = rd;
                out.validity[o] = rs;
            } else {
                out.color[3 * o] = st.c.x; out.color[3 * o + 1] = st.c.y; out.color[3 * o + 2] = st.c.z;
                if (out.depth) out.depth[o] = st.d;
                if (out.quality) out.quality[o] = st.q;
                if (out.validity) out.validity[o] = st.v;
            }
        }
        __syncthreads();
    }
}

template <class G, int MODE>
hipError_t launch_raster(const G* geoms, int nsrc, const RasterDims& d, const RasterWS& ws, float blend_exp,
                         const RasterOut& out, hipStream_t s) {
    const int blocks = 8 * ((d.ntiles + 7) / 8);
    hipLaunchKernelGGL((k_raster<G, MODE>), dim3(blocks), dim3(256), 0, s, geoms, nsrc, d, ws, blend_exp, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int fb_blocks = d.ntiles < 512 ? d.ntiles : 512;
    hipLaunchKernelGGL((k_fallback<G, MODE>), dim3(fb_blocks), dim3(256), 0, s, geoms, nsrc, d, ws, blend_exp, out);
    return hipGetLastError();
}

template hipError_t launch_bin<PairGeom>(const PairGeom*, int, int, int, const RasterDims&, const RasterWS&, hipStream_t);
template hipError_t launch_bin<ArrayGeom>(const ArrayGeom*, int, int, int, const RasterDims&, const RasterWS&, hipStream_t);
template hipError_t launch_raster<PairGeom, MODE_BLEND>(const PairGeom*, int, const RasterDims&, const RasterWS&, float,
                                                        const RasterOut&, hipStream_t);
template hipError_t launch_raster<ArrayGeom, MODE_RAW>(const ArrayGeom*, int, const RasterDims&, const RasterWS&, float,
                                                       const RasterOut&, hipStream_t);

}  // namespace rvs_k
