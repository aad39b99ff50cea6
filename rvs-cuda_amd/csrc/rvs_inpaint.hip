// rvs_inpaint.hip -- exact nearest-known-pixel inpainting for gfx950.
//
// Reference: detail::compute_nearest (rvs/src/inpainting.cpp:104-136) fills a
// map (src x, src y, dist) by alternating full-image sweeps, forward (x outer
// ascending, y inner ascending) then backward, until a sweep changes nothing.
// A hole pixel visited by a sweep takes, among [its current value, left+1,
// up+1, down+1, right+1] (in that order, strict <), the first minimum, where
// left/up are this sweep's values in a forward sweep and down/right in a
// backward sweep.  So one sweep is a pure function of (current value, two
// "new" neighbours, two "old" neighbours): a dynamic program over a DAG, and
// any topological order reproduces it.  This file evaluates it row by row
// (forward: rows top-down, backward: bottom-up) with a block-wide scan along
// each row of the function  S -> (S.v + 1 < C.v + t) ? S + 1 : C,  a family
// closed under composition.  Hole pixels only ever read known pixels or holes of
// their own 4-connected component, so components are solved independently
// (labelled with a lock-free union-find) and each sweeps until it stops
// changing, which is what the global loop gives every component.  Two sweeps
// always suffice: after the forward and the backward sweep every hole holds its
// exact L1 distance to the nearest known pixel (the two-pass city-block
// distance transform of Rosenfeld and Pfaltz: a nearest known pixel in any
// quadrant is reached by a monotone path whose pixels are all holes, one leg
// per sweep), so the reference's third sweep improves nothing -- the distances,
// and the sources, which only change with a strictly smaller distance, are
// final.  The solvers stop after the backward sweep (or after the forward one
// if it changed nothing: then no hole touches a known pixel).  One block
// solves one component: in LDS when its bounding box fits (solve_component),
// otherwise on the global map -- with the previous line in registers and the
// next lines prefetched when its lines fit one block step (solve_lines), else
// with several pixels per thread per step (solve_big).  Large components are
// started first; those whose lines fit a 1024-thread step have their own
// kernel (k_inp_big).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_inpaint.h"

namespace rvs_k {

constexpr int kAbsent = 1 << 29;
constexpr int kSweeps = 2;   // forward, backward: the fixed point of the alternating sweeps

__device__ inline int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Label (union-find parent) of record p.
__device__ inline int* label_of(int4* px, int p) { return &px[p].z; }

__device__ inline int find_root(int4* px, int p) {
    while (true) {
        const int q = ld(label_of(px, p));
        if (q == p) return p;
        p = q;
    }
}

__device__ inline void unite(int4* px, int a, int b) {
    while (true) {
        a = find_root(px, a);
        b = find_root(px, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        const int old = atomicMin(label_of(px, a), b);
        if (old == a) return;
        a = old;
    }
}

// Grid of the per-pixel kernels: x = pixel blocks of the largest view, y = view.
#define RVS_VIEW_PIXEL(vd, p, gp)                                           \
    const ViewDesc& vd = views[blockIdx.y];                                 \
    const int p = blockIdx.x * blockDim.x + threadIdx.x;                    \
    if (p >= vd.w * vd.h) return;                                           \
    const int gp = (int)vd.px_base + p;

constexpr int kCclTile = 32;

__device__ inline int lfind(const int* lab, int p) {
    while (true) {
        const int q = lab[p];
        if (q == p) return p;
        p = q;
    }
}

__device__ inline void lunite(int* lab, int a, int b) {
    while (true) {
        a = lfind(lab, a);
        b = lfind(lab, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        const int old = atomicMin(&lab[a], b);
        if (old == a) return;
        a = old;
    }
}

// Inpainting state of one 32x32 tile and its labels in LDS: known pixels get
// the label -1 (their map value (itself, 0) is implied by it and never
// stored), each hole pixel the map's initial (itself, w + h) (inpaint_color,
// inpainting.cpp:185-191) and the global index of its tile-local root (the
// smallest local index of its local component).  A tile without holes stores
// nothing but its flag (0); a tile with holes sets its flag and joins the list
// of hole tiles that the later passes walk.
__global__ void __launch_bounds__(256) k_ccl_local(const ViewDesc* __restrict__ views, InpaintWS ws) {
    __shared__ int lab[kCclTile * kCclTile];
    const ViewDesc& vd = views[blockIdx.z];
    const int w = vd.w, h = vd.h, base = (int)vd.px_base;
    const int tx0 = blockIdx.x * kCclTile, ty0 = blockIdx.y * kCclTile;
    if (tx0 >= w || ty0 >= h) return;
    const int tid = threadIdx.x;
    const int t = (int)blockIdx.z * ws.tpv + (int)blockIdx.y * ws.ttx + (int)blockIdx.x;
    if (ws.rflags) {
        // the raster tiles covering this tile (one, or up to 2 x 2 of 16 x 16)
        const int sh = vd.tile_shift;
        const int rx0 = tx0 >> sh, rx1 = (min(tx0 + kCclTile, w) - 1) >> sh;
        const int ry0 = ty0 >> sh, ry1 = (min(ty0 + kCclTile, h) - 1) >> sh;
        int f = 0;
        for (int ry = ry0; ry <= ry1; ++ry)
            for (int rx = rx0; rx <= rx1; ++rx) f |= ws.rflags[vd.tile_base + (uint32_t)(ry * vd.tiles_x + rx)];
        if (!(f & 2)) {   // the same in every thread
            if (tid == 0) ws.thole[t] = 0;
            return;
        }
    }
    bool any = false;
    for (int k = tid; k < kCclTile * kCclTile; k += 256) {
        const int x = tx0 + (k % kCclTile), y = ty0 + (k / kCclTile);
        bool hole = false;
        // a pixel of a raster tile without holes is known (its quality may not be stored)
        if (x < w && y < h &&
            (!ws.rflags ||
             (ws.rflags[vd.tile_base + (uint32_t)((y >> vd.tile_shift) * vd.tiles_x + (x >> vd.tile_shift))] & 2))) {
            const int p = y * w + x, gp = base + p;
            hole = !(vd.quality[p] > 0.f);   // View::get_inpaint_mask (View.cpp:198-203)
            if (hole) {
                ws.px[gp].x = w + h;
                ws.px[gp].y = gp;
            }
        }
        lab[k] = hole ? k : -1;
        any |= hole;
    }
    if (!__syncthreads_or(any)) {   // also publishes lab
        if (tid == 0) ws.thole[t] = 0;
        return;
    }
    if (tid == 0) {
        ws.thole[t] = 1;
        ws.htiles[atomicAdd(&ws.counters[8], 1)] = t;
    }
    // a known pixel's map value is (itself, 0), implied by its label
    for (int k = tid; k < kCclTile * kCclTile; k += 256) {
        const int x = tx0 + (k % kCclTile), y = ty0 + (k / kCclTile);
        if (lab[k] < 0 && x < w && y < h) ws.px[base + y * w + x].z = -1;
    }
    for (int k = tid; k < kCclTile * kCclTile; k += 256) {
        if (lab[k] < 0) continue;
        const int lx = k % kCclTile, ly = k / kCclTile;
        if (lx + 1 < kCclTile && lab[k + 1] >= 0) lunite(lab, k, k + 1);
        if (ly + 1 < kCclTile && lab[k + kCclTile] >= 0) lunite(lab, k, k + kCclTile);
    }
    __syncthreads();
    for (int k = tid; k < kCclTile * kCclTile; k += 256) {
        if (lab[k] < 0) continue;
        const int r = lfind(lab, k);
        const int x = tx0 + (k % kCclTile), y = ty0 + (k / kCclTile);
        ws.px[base + y * w + x].z = base + (ty0 + r / kCclTile) * w + (tx0 + r % kCclTile);
    }
}

// A hole tile of the list: its view and first pixel.
struct HoleTile {
    int v, tx0, ty0;
};

__device__ inline HoleTile hole_tile(const InpaintWS& ws, int i) {
    const int t = ws.htiles[i];
    HoleTile H;
    H.v = t / ws.tpv;
    const int r = t - H.v * ws.tpv;
    H.tx0 = (r % ws.ttx) * kCclTile;
    H.ty0 = (r / ws.ttx) * kCclTile;
    return H;
}

// Unions across tile borders on the tile-local roots: a hole tile unites its
// right column and bottom row with the next tile's pixels (a hole pair across
// a border has a hole tile on both sides, so every such pair is seen once).
constexpr int kMergeThreads = 2 * kCclTile;

__global__ void __launch_bounds__(kMergeThreads) k_ccl_merge(const ViewDesc* __restrict__ views, InpaintWS ws) {
    const int nt = ws.counters[8];
    const int tid = threadIdx.x;
    const bool right = tid < kCclTile;
    for (int i = blockIdx.x; i < nt; i += gridDim.x) {
        const HoleTile H = hole_tile(ws, i);
        const ViewDesc& vd = views[H.v];
        const int x = right ? H.tx0 + kCclTile - 1 : H.tx0 + (tid - kCclTile);
        const int y = right ? H.ty0 + tid : H.ty0 + kCclTile - 1;
        const int nx = right ? x + 1 : x, ny = right ? y : y + 1;
        if (nx >= vd.w || ny >= vd.h) continue;
        const int gp = (int)vd.px_base + y * vd.w + x, gq = (int)vd.px_base + ny * vd.w + nx;
        const bool nh = ws.thole[H.v * ws.tpv + (ny / kCclTile) * ws.ttx + nx / kCclTile] != 0;
        if (nh && ws.px[gp].z >= 0 && ws.px[gq].z >= 0) unite(ws.px, gp, gq);
    }
}

__global__ void __launch_bounds__(256) k_ccl_flatten(const ViewDesc* __restrict__ views, InpaintWS ws) {
    const int nt = ws.counters[8];
    for (int i = blockIdx.x; i < nt; i += gridDim.x) {
        const HoleTile H = hole_tile(ws, i);
        const ViewDesc& vd = views[H.v];
        for (int k = threadIdx.x; k < kCclTile * kCclTile; k += 256) {
            const int x = H.tx0 + (k % kCclTile), y = H.ty0 + (k / kCclTile);
            if (x >= vd.w || y >= vd.h) continue;
            const int gp = (int)vd.px_base + y * vd.w + x;
            if (ws.px[gp].z < 0) continue;
            const int r = find_root(ws.px, gp);
            ws.px[gp].z = r;
            if (r == gp) {
                const int id = atomicAdd(&ws.counters[0], 1);
                ws.px[gp].w = id;
                ws.comp[4 * id + 0] = INT_MAX;
                ws.comp[4 * id + 1] = INT_MIN;
                ws.comp[4 * id + 2] = INT_MAX;
                ws.comp[4 * id + 3] = INT_MIN;
                ws.root[id] = gp;
            }
        }
    }
}

// Component bounding boxes, a block per hole tile.  Same-address device
// atomics are serialised at the memory side, so a large hole must not cost
// atomics per pixel: each wave reduces, over its kBboxPx pixels per lane, the
// box of the component of its first hole (usually the whole wave when a large
// hole covers it), the block merges the waves' boxes per component, and one
// thread takes four atomics per distinct component of the tile.  Holes of
// other components use plain atomics.
constexpr int kBboxThreads = 256;
constexpr int kBboxPx = kCclTile * kCclTile / kBboxThreads;   // pixels per thread, kBboxThreads apart

__global__ void __launch_bounds__(kBboxThreads) k_comp_bbox(const ViewDesc* __restrict__ views, InpaintWS ws) {
    __shared__ int s_lid[kBboxThreads / 64];
    __shared__ int s_box[kBboxThreads / 64][4];
    const int nt = ws.counters[8];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (int i = blockIdx.x; i < nt; i += gridDim.x) {
    const HoleTile H = hole_tile(ws, i);
    const ViewDesc& vd = views[H.v];
    int lid = -1;   // wave-uniform: component of the wave's first hole
    int x0 = INT_MAX, x1 = INT_MIN, y0 = INT_MAX, y1 = INT_MIN;
    for (int k = 0; k < kBboxPx; ++k) {
        const int l = k * kBboxThreads + (int)threadIdx.x;
        const int x = H.tx0 + l % kCclTile, y = H.ty0 + l / kCclTile;
        const bool in = x < vd.w && y < vd.h;
        const int gp = (int)vd.px_base + (in ? y * vd.w + x : 0);
        const int r = in ? ws.px[gp].z : -1;
        const int id = r >= 0 ? ws.px[r].w : -1;
        const uint64_t act = __ballot(id >= 0);
        if (act == 0ull) continue;
        if (lid < 0) lid = __shfl(id, __ffsll((unsigned long long)act) - 1);
        if (id == lid) {
            x0 = min(x0, x); x1 = max(x1, x);
            y0 = min(y0, y); y1 = max(y1, y);
        } else if (id >= 0) {
            atomicMin(&ws.comp[4 * id + 0], x);
            atomicMax(&ws.comp[4 * id + 1], x);
            atomicMin(&ws.comp[4 * id + 2], y);
            atomicMax(&ws.comp[4 * id + 3], y);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        x0 = min(x0, __shfl_xor(x0, off));
        x1 = max(x1, __shfl_xor(x1, off));
        y0 = min(y0, __shfl_xor(y0, off));
        y1 = max(y1, __shfl_xor(y1, off));
    }
    if (lane == 0) {
        s_lid[wave] = lid;
        s_box[wave][0] = x0; s_box[wave][1] = x1; s_box[wave][2] = y0; s_box[wave][3] = y1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        constexpr int NW = kBboxThreads / 64;
        for (int w = 0; w < NW; ++w) {
            const int id = s_lid[w];
            if (id < 0) continue;
            int b0 = s_box[w][0], b1 = s_box[w][1], b2 = s_box[w][2], b3 = s_box[w][3];
            for (int u = w + 1; u < NW; ++u)
                if (s_lid[u] == id) {
                    b0 = min(b0, s_box[u][0]); b1 = max(b1, s_box[u][1]);
                    b2 = min(b2, s_box[u][2]); b3 = max(b3, s_box[u][3]);
                    s_lid[u] = -1;
                }
            atomicMin(&ws.comp[4 * id + 0], b0);
            atomicMax(&ws.comp[4 * id + 1], b1);
            atomicMin(&ws.comp[4 * id + 2], b2);
            atomicMax(&ws.comp[4 * id + 3], b3);
        }
    }
    __syncthreads();   // s_lid, s_box free for the next tile
    }
}

// S -> (S.v + k < C.v + t) ? (S.v + k, S.src) : C
struct Fn {
    int cv, cs, t, k;
};

// Selects rather than branches throughout (the solvers evaluate these per
// pixel and line step; divergent branches would cost exec-mask juggling).
__device__ inline Fn compose(const Fn& f1, const Fn& f2) {   // f2 after f1
    Fn r;
    const bool a = f1.cv + f2.k < f2.cv + f2.t;
    r.cv = a ? f1.cv + f2.k : f2.cv;
    r.cs = a ? f1.cs : f2.cs;
    r.t = a ? f1.t : f2.t;
    r.k = f1.k + f2.k;
    return r;
}

// lexicographic (value, rank) minimum, first argument wins ties of both
__device__ inline void cand(int& bv, int& br, int& bs, int v, int r, int s) {
    const bool take = (v < bv) | ((v == bv) & (r < br));
    bv = take ? v : bv;
    br = take ? r : br;
    bs = take ? s : bs;
}

// the same, offered only when `ok`
__device__ inline void cand_if(bool ok, int& bv, int& br, int& bs, int v, int r, int s) {
    const bool take = ok & ((v < bv) | ((v == bv) & (r < br)));
    bv = take ? v : bv;
    br = take ? r : br;
    bs = take ? s : bs;
}

// Component region access: LDS copy of the bounding box (small components) or
// the global map (large ones).  Pixels outside the box that a member can see
// are known pixels; non-member holes inside the box are never a member's
// neighbour (4-connectivity), so they only pass through scans as constants.
constexpr int kLdsCap = 6144;   // pixels of a component box held in LDS (48 KB)
constexpr int kOther = -1;      // dist of a hole of another component (inside the box)

struct Region {
    int x0, y0, bw, bh, w, h, root;
    int base;     // global index of the view's pixel (0, 0)
    int tb;       // the view's first thole entry
    int* ldist;   // LDS mode
    int* lsrc;
};

// The hole flag of the tile of view pixel (x, y).
__device__ inline bool tile_has_hole(const Region& R, const InpaintWS& ws, int x, int y) {
    return ws.thole[R.tb + (y / kCclTile) * ws.ttx + x / kCclTile] != 0;
}

// The label of view pixel (x, y) at global index q: -1 (known) in a tile
// without holes, which stores none; both words are loaded unconditionally.
__device__ inline int label_at(const Region& R, const InpaintWS& ws, int x, int y, int q) {
    const int l = ws.px[q].z;
    return tile_has_hole(R, ws, x, y) ? l : -1;
}

template <bool LDS>
__device__ inline void region_get(const Region& R, const InpaintWS& ws, int x, int y, int& v, int& s) {
    const int p = R.base + y * R.w + x;
    if (LDS) {
        const int bx = x - R.x0, by = y - R.y0;
        if (bx >= 0 && bx < R.bw && by >= 0 && by < R.bh) {
            const int k = by * R.bw + bx;
            v = R.ldist[k]; s = R.lsrc[k];
        } else {
            v = 0; s = p;                     // a known pixel next to the box
        }
    } else {
        const int lab = label_at(R, ws, x, y, p);
        if (lab < 0) { v = 0; s = p; }
        else if (lab != R.root) { v = kOther; s = -1; }
        else { const int4 m = ws.px[p]; v = m.x; s = m.y; }
    }
}

template <bool LDS>
__device__ inline void region_put(const Region& R, const InpaintWS& ws, int x, int y, int v, int s) {
    if (LDS) {
        const int k = (y - R.y0) * R.bw + (x - R.x0);
        R.ldist[k] = v; R.lsrc[k] = s;
    } else {
        const int p = R.base + y * R.w + x;
        *reinterpret_cast<int2*>(&ws.px[p]) = make_int2(v, s);
    }
}

// One component: a forward sweep, then a backward one if the first changed anything.
// Lines run along the box's shorter side and are processed in sweep order; each
// line is a scan along the longer side.  Row lines scan x (forward: the scan
// input is the left neighbour, rank 1; backward: the right one, rank 4); column
// lines scan y (forward: up, rank 2; backward: down, rank 3).  Every other
// neighbour is read from the map, which at that moment holds this sweep's value
// for neighbours earlier in the sweep order and the previous sweep's otherwise.
template <bool LDS>
__device__ void solve_component(const Region& R, const InpaintWS& ws, Fn* wave_tot, int* s_carry) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const bool cols = R.bh > R.bw;
    const int nlines = cols ? R.bw : R.bh;
    const int llen = cols ? R.bh : R.bw;
    for (int sweep = 0; sweep < kSweeps; ++sweep) {
        const bool fwd = (sweep & 1) == 0;
        const int rs = cols ? (fwd ? 2 : 3) : (fwd ? 1 : 4);
        bool changed = false;
        for (int li = 0; li < nlines; ++li) {
            const int line = fwd ? li : nlines - 1 - li;
            // carry in: the pixel before the first element in scan order (outside the box)
            int Lv = kAbsent, Ls = -1;
            {
                const int pb = fwd ? -1 : llen;
                const int bx = cols ? R.x0 + line : R.x0 + pb;
                const int by = cols ? R.y0 + pb : R.y0 + line;
                if (bx >= 0 && bx < R.w && by >= 0 && by < R.h) {
                    int v, sr;
                    region_get<LDS>(R, ws, bx, by, v, sr);
                    if (v == 0) { Lv = 0; Ls = sr; }
                }
            }
            for (int cb = 0; cb < llen; cb += 256) {
                const int e = cb + tid;
                const bool in = e < llen;
                const int pos = fwd ? e : llen - 1 - e;
                const int x = cols ? R.x0 + line : R.x0 + pos;
                const int y = cols ? R.y0 + pos : R.y0 + line;
                Fn f;
                f.k = 1;
                bool member = false;
                int cur_v = 0;
                f.cv = kAbsent; f.cs = -1; f.t = 0;
                if (in) {
                    int v, sr;
                    region_get<LDS>(R, ws, x, y, v, sr);
                    if (v == 0) {
                        f.cv = 0; f.cs = sr;                       // known pixel: (0, itself)
                    } else if (v != kOther) {
                        member = true;
                        cur_v = v;
                        int bv = v, br = 0, bs = sr;
                        // neighbours in the reference order, left(1) up(2) down(3) right(4),
                        // except the scan input
                        int nv, ns;
                        if (rs != 1 && x > 0) { region_get<LDS>(R, ws, x - 1, y, nv, ns); if (nv != kOther) cand(bv, br, bs, nv + 1, 1, ns); }
                        if (rs != 2 && y > 0) { region_get<LDS>(R, ws, x, y - 1, nv, ns); if (nv != kOther) cand(bv, br, bs, nv + 1, 2, ns); }
                        if (rs != 3 && y + 1 < R.h) { region_get<LDS>(R, ws, x, y + 1, nv, ns); if (nv != kOther) cand(bv, br, bs, nv + 1, 3, ns); }
                        if (rs != 4 && x + 1 < R.w) { region_get<LDS>(R, ws, x + 1, y, nv, ns); if (nv != kOther) cand(bv, br, bs, nv + 1, 4, ns); }
                        f.cv = bv; f.cs = bs; f.t = (rs < br) ? 1 : 0;
                    }
                }
                // inclusive scan of compositions in thread (= scan) order
                for (int off = 1; off < 64; off <<= 1) {
                    Fn g;
                    g.cv = __shfl_up(f.cv, off);
                    g.cs = __shfl_up(f.cs, off);
                    g.t = __shfl_up(f.t, off);
                    g.k = __shfl_up(f.k, off);
                    if (lane >= off) f = compose(g, f);
                }
                if (lane == 63) wave_tot[wave] = f;
                __syncthreads();   // all reads of this chunk are done before any write
                Fn pre;
                bool have = false;
                for (int wv = 0; wv < wave; ++wv) {
                    pre = have ? compose(pre, wave_tot[wv]) : wave_tot[wv];
                    have = true;
                }
                if (have) f = compose(pre, f);
                int rv, rsrc;
                if (Lv + f.k < f.cv + f.t) { rv = Lv + f.k; rsrc = Ls; } else { rv = f.cv; rsrc = f.cs; }
                if (member && rv != cur_v) {
                    changed = true;
                    region_put<LDS>(R, ws, x, y, rv, rsrc);
                }
                const int last = min(256, llen - cb) - 1;
                if (tid == last) { s_carry[0] = rv; s_carry[1] = rsrc; }
                __syncthreads();
                Lv = s_carry[0]; Ls = s_carry[1];
                __syncthreads();
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

constexpr int kBigPx = 4;   // consecutive pixels per thread in a line step of a large component

__device__ inline void apply_fn(const Fn& f, int& v, int& s) {
    const bool a = v + f.k < f.cv + f.t;
    s = a ? s : f.cs;
    v = a ? v + f.k : f.cv;
}

// Map value of view pixel (x, y) for component R from its three words, loaded
// unconditionally so that a line step issues all its loads at once.
__device__ inline void map_value(const Region& R, int lab, int d, int sr, int p, int& v, int& s) {
    if (lab < 0) { v = 0; s = p; }
    else if (lab != R.root) { v = kOther; s = -1; }
    else { v = d; s = sr; }
}

// A component whose box exceeds the LDS copy, on the global map: the same
// sweeps and line order as solve_component, but each thread takes kBigPx
// consecutive pixels of a line (composed in registers before the block-wide
// scan) and loads every word it needs (the pixels and their neighbours)
// before using any, so a line step costs about one memory latency.
__device__ void solve_big(const Region& R, const InpaintWS& ws, Fn* wave_tot, int* s_carry) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const bool cols = R.bh > R.bw;
    const int nlines = cols ? R.bw : R.bh;
    const int llen = cols ? R.bh : R.bw;
    const int span = 256 * kBigPx;
    for (int sweep = 0; sweep < kSweeps; ++sweep) {
        const bool fwd = (sweep & 1) == 0;
        const int rs = cols ? (fwd ? 2 : 3) : (fwd ? 1 : 4);
        // neighbour offsets in the reference order left(1) up(2) down(3) right(4)
        const int ndx[4] = {-1, 0, 0, 1}, ndy[4] = {0, -1, 1, 0};
        bool changed = false;
        for (int li = 0; li < nlines; ++li) {
            const int line = fwd ? li : nlines - 1 - li;
            int Lv = kAbsent, Ls = -1;
            {
                const int pb = fwd ? -1 : llen;
                const int bx = cols ? R.x0 + line : R.x0 + pb;
                const int by = cols ? R.y0 + pb : R.y0 + line;
                if (bx >= 0 && bx < R.w && by >= 0 && by < R.h) {
                    int v, sr;
                    region_get<false>(R, ws, bx, by, v, sr);
                    if (v == 0) { Lv = 0; Ls = sr; }
                }
            }
            for (int cb = 0; cb < llen; cb += span) {
                // loads: kBigPx pixels x (itself + 4 neighbours) x one record each
                int lab[kBigPx][5], dd[kBigPx][5], ss[kBigPx][5], pp[kBigPx][5];
                bool ok[kBigPx][5];
#pragma unroll
                for (int k = 0; k < kBigPx; ++k) {
                    const int e = min(cb + tid * kBigPx + k, llen - 1);
                    const int pos = fwd ? e : llen - 1 - e;
                    const int x = cols ? R.x0 + line : R.x0 + pos;
                    const int y = cols ? R.y0 + pos : R.y0 + line;
#pragma unroll
                    for (int n = 0; n < 5; ++n) {
                        const int nx = n ? x + ndx[n - 1] : x, ny = n ? y + ndy[n - 1] : y;
                        ok[k][n] = nx >= 0 && nx < R.w && ny >= 0 && ny < R.h;
                        const int q = R.base + (ok[k][n] ? ny * R.w + nx : y * R.w + x);
                        pp[k][n] = q;
                        const int4 m = ws.px[q];
                        lab[k][n] = tile_has_hole(R, ws, ok[k][n] ? nx : x, ok[k][n] ? ny : y) ? m.z : -1;
                        dd[k][n] = m.x;
                        ss[k][n] = m.y;
                    }
                }
                Fn f[kBigPx];
                int cur[kBigPx];   // > 0: a hole of this component with that distance
#pragma unroll
                for (int k = 0; k < kBigPx; ++k) {
                    const int e = cb + tid * kBigPx + k;
                    f[k].k = 1; f[k].cv = kAbsent; f[k].cs = -1; f[k].t = 0;
                    cur[k] = 0;
                    int v, sr;
                    map_value(R, lab[k][0], dd[k][0], ss[k][0], pp[k][0], v, sr);
                    if (e < llen) {
                        if (v == 0) {
                            f[k].cv = 0; f[k].cs = sr;
                        } else if (v != kOther) {
                            cur[k] = v;
                            int bv = v, br = 0, bs = sr;
#pragma unroll
                            for (int n = 1; n < 5; ++n) {
                                if (n == rs || !ok[k][n]) continue;
                                int nv, ns;
                                map_value(R, lab[k][n], dd[k][n], ss[k][n], pp[k][n], nv, ns);
                                if (nv != kOther) cand(bv, br, bs, nv + 1, n, ns);
                            }
                            f[k].cv = bv; f[k].cs = bs; f[k].t = (rs < br) ? 1 : 0;
                        }
                    }
                }
                Fn A = f[0];
#pragma unroll
                for (int k = 1; k < kBigPx; ++k) A = compose(A, f[k]);
                for (int off = 1; off < 64; off <<= 1) {
                    Fn g;
                    g.cv = __shfl_up(A.cv, off);
                    g.cs = __shfl_up(A.cs, off);
                    g.t = __shfl_up(A.t, off);
                    g.k = __shfl_up(A.k, off);
                    if (lane >= off) A = compose(g, A);
                }
                Fn ex;
                ex.cv = __shfl_up(A.cv, 1);
                ex.cs = __shfl_up(A.cs, 1);
                ex.t = __shfl_up(A.t, 1);
                ex.k = __shfl_up(A.k, 1);
                if (lane == 63) wave_tot[wave] = A;
                __syncthreads();   // all reads of this span are done before any write
                Fn pre;
                bool have = false;
                for (int wv = 0; wv < wave; ++wv) {
                    pre = have ? compose(pre, wave_tot[wv]) : wave_tot[wv];
                    have = true;
                }
                if (lane > 0) {
                    pre = have ? compose(pre, ex) : ex;
                    have = true;
                }
                int Sv = Lv, Ss = Ls;
                if (have) apply_fn(pre, Sv, Ss);
                const int last = min(cb + span, llen) - 1;
#pragma unroll
                for (int k = 0; k < kBigPx; ++k) {
                    apply_fn(f[k], Sv, Ss);
                    const int e = cb + tid * kBigPx + k;
                    if (e <= last) {
                        if (cur[k] > 0 && Sv != cur[k]) {
                            changed = true;
                            *reinterpret_cast<int2*>(&ws.px[pp[k][0]]) = make_int2(Sv, Ss);
                        }
                        if (e == last) { s_carry[0] = Sv; s_carry[1] = Ss; }
                    }
                }
                __syncthreads();
                Lv = s_carry[0]; Ls = s_carry[1];
                __syncthreads();
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

constexpr int kBigThreads = 1024;                 // block of k_inp_big
constexpr int kLinePx = 4;                        // consecutive pixels per thread in its line steps
constexpr int kLineMax = kBigThreads * kLinePx;   // longest line of solve_lines (one block step)

// Large components whose lines fit one block step.  Each thread owns the same
// PX positions on every line, so the previous line's values from this
// sweep stay in registers; a line step loads only its own pixels, the next
// line and one along-line neighbour (all independent loads) and takes one
// barrier.  cols: lines are columns.  Same sweeps, candidate order and line
// order semantics as solve_component.
template <int NT, int PX>
__device__ void solve_lines(const Region& R, const InpaintWS& ws, bool cols, Fn* wave_tot) {
    constexpr int NW = NT / 64;   // waves per block
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int nlines = cols ? R.bw : R.bh;
    const int llen = cols ? R.bh : R.bw;
    for (int sweep = 0; sweep < kSweeps; ++sweep) {
        const bool fwd = (sweep & 1) == 0;
        // ranks (left 1, up 2, down 3, right 4) of the scan input, of the
        // across-line neighbour already visited in this sweep, of the one not yet
        // visited, and of the along-line neighbour after the pixel in scan order
        const int rs = cols ? (fwd ? 2 : 3) : (fwd ? 1 : 4);
        const int r_prev = cols ? (fwd ? 1 : 4) : (fwd ? 2 : 3);
        const int r_next = cols ? (fwd ? 4 : 1) : (fwd ? 3 : 2);
        const int r_along = cols ? (fwd ? 3 : 2) : (fwd ? 4 : 1);
        // view pixel of line coordinate lc (box-relative) and scan index e
        auto at = [&](int lc, int e, int& x, int& y) {
            const int pos = fwd ? e : llen - 1 - e;
            x = cols ? R.x0 + lc : R.x0 + pos;
            y = cols ? R.y0 + pos : R.y0 + lc;
        };
        auto inside = [&](int x, int y) { return x >= 0 && x < R.w && y >= 0 && y < R.h; };
        auto load = [&](int x, int y, int& v, int& sr) {
            const bool in = inside(x, y);
            const int q = R.base + (in ? y * R.w + x : 0);
            const int4 m = ws.px[q];
            map_value(R, tile_has_hole(R, ws, in ? x : 0, in ? y : 0) ? m.z : -1, m.x, m.y, q, v, sr);
        };
        bool changed = false;
        // line coordinate of processing step li (-1 / nlines: outside the box)
        auto lcoord = [&](int li) { return fwd ? li : nlines - 1 - li; };
        // words of line lc at this thread's positions, the along-line
        // neighbour after them and the pixel before the line's first one
        auto load_line = [&](int lc, int* v, int* sr, bool* in) {
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                int x, y;
                at(lc, tid * PX + k, x, y);
                in[k] = inside(x, y);
                load(x, y, v[k], sr[k]);
            }
        };
        auto load_ends = [&](int lc, int& av, int& as_, bool& ain, int& Lv, int& Ls) {
            int x, y;
            at(lc, tid * PX + PX, x, y);
            ain = inside(x, y);
            load(x, y, av, as_);
            at(lc, -1, x, y);
            int v, sr;
            load(x, y, v, sr);
            const bool known = inside(x, y) && v == 0;
            Lv = known ? 0 : kAbsent;
            Ls = known ? sr : -1;
        };
        // the line before the first one (outside the box, constant), the first
        // line and the line after it; later lines are loaded one step ahead
        int pv[PX], ps[PX], sv[PX], ss[PX], nv[PX], ns[PX];
        bool pin[PX], sin[PX], nin[PX];
        int av, as_, Lv, Ls;
        bool ain;
        load_line(lcoord(-1), pv, ps, pin);
        load_line(lcoord(0), sv, ss, sin);
        load_line(lcoord(1), nv, ns, nin);
        load_ends(lcoord(0), av, as_, ain, Lv, Ls);
        for (int li = 0; li < nlines; ++li) {
            const int lc = lcoord(li);
            // prefetch for the next step: the line after its next one and its ends
            int fv[PX], fs[PX], fav, fas, fLv, fLs;
            bool fin[PX], fain;
            load_line(lcoord(li + 2), fv, fs, fin);
            load_ends(lcoord(li + 1), fav, fas, fain, fLv, fLs);
            Fn f[PX];
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                // past the line: (kAbsent, -1, 0); known pixel: (0, itself); a
                // hole of another component: (kAbsent, -1, 0); a member: its
                // best candidate
                const bool valid = tid * PX + k < llen;
                const int v = sv[k];
                const bool known = valid & (v == 0);
                const bool member = valid & (v != 0) & (v != kOther);
                int bv = v, br = 0, bs = ss[k];
                const int alv = k + 1 < PX ? sv[k + 1] : av, als = k + 1 < PX ? ss[k + 1] : as_;
                const bool alin = k + 1 < PX ? sin[k + 1] : ain;
                // cand keeps the (value, rank) minimum, so the calls' order is free
                cand_if(pin[k] & (pv[k] != kOther), bv, br, bs, pv[k] + 1, r_prev, ps[k]);
                cand_if(nin[k] & (nv[k] != kOther), bv, br, bs, nv[k] + 1, r_next, ns[k]);
                cand_if(alin & (alv != kOther), bv, br, bs, alv + 1, r_along, als);
                f[k].k = 1;
                f[k].cv = known ? 0 : (member ? bv : kAbsent);
                f[k].cs = known ? ss[k] : (member ? bs : -1);
                f[k].t = (member & (rs < br)) ? 1 : 0;
            }
            Fn A = f[0];
#pragma unroll
            for (int k = 1; k < PX; ++k) A = compose(A, f[k]);
            for (int off = 1; off < 64; off <<= 1) {
                Fn g;
                g.cv = __shfl_up(A.cv, off);
                g.cs = __shfl_up(A.cs, off);
                g.t = __shfl_up(A.t, off);
                g.k = __shfl_up(A.k, off);
                if (lane >= off) A = compose(g, A);
            }
            Fn ex;
            ex.cv = __shfl_up(A.cv, 1);
            ex.cs = __shfl_up(A.cs, 1);
            ex.t = __shfl_up(A.t, 1);
            ex.k = __shfl_up(A.k, 1);
            Fn* wt = wave_tot + NW * (li & 1);   // alternate buffers: one barrier per line
            if (lane == 63) wt[wave] = A;
            __syncthreads();   // every load of this line is done before any write
            Fn pre;
            bool have = false;
            for (int wv = 0; wv < wave; ++wv) {
                pre = have ? compose(pre, wt[wv]) : wt[wv];
                have = true;
            }
            if (lane > 0) {
                pre = have ? compose(pre, ex) : ex;
                have = true;
            }
            int Sv = Lv, Ss = Ls;
            if (have) apply_fn(pre, Sv, Ss);
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                apply_fn(f[k], Sv, Ss);
                const int e = tid * PX + k;
                const bool member = e < llen && sv[k] != 0 && sv[k] != kOther;
                if (member && Sv != sv[k]) {
                    int x, y;
                    at(lc, e, x, y);
                    const int q = R.base + y * R.w + x;
                    changed = true;
                    *reinterpret_cast<int2*>(&ws.px[q]) = make_int2(Sv, Ss);
                }
                pv[k] = member ? Sv : sv[k];
                ps[k] = member ? Ss : ss[k];
                pin[k] = sin[k];
            }
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                sv[k] = nv[k]; ss[k] = ns[k]; sin[k] = nin[k];
                nv[k] = fv[k]; ns[k] = fs[k]; nin[k] = fin[k];
            }
            av = fav; as_ = fas; ain = fain; Lv = fLv; Ls = fLs;
        }
        if (!__syncthreads_or(changed)) break;
    }
}

// ---------------------------------------------------------- strip solver
//
// Large components whose lines fit kLineMax are solved with each line split
// into strips of kStripPx scan positions, one wave per strip (no barrier: the
// line scan is a wave scan), so that a line step's work is spread over many
// CUs instead of one.  Strip k
// of a line takes as its scan input the state after strip k-1's last pixel of
// the same line (its carry); every other value a strip reads lies in its own
// columns -- the previous line's values of this sweep are kept in registers,
// the next line still holds the previous sweep's -- or is the pixel just past
// its end, which the next strip writes only after taking this strip's carry of
// that line.  Carries travel through one 64-bit slot per (strip, line), zeroed
// before the sweep and written once with a single atomic store, so the reader
// needs no other ordering.  The forward and the backward sweep are separate
// launches (the backward one reads the forward result).  Within a launch the
// tasks (component, strip) are handed out by ticket with the strips of a
// component in scan order, so a strip waits only for a task whose ticket was
// taken before its own, by a block that is already running: no wait depends on
// a block that has not started.  Waits are bounded anyway: after kSpinLimit
// polls a strip gives up, takes an absent carry and sets ws.fail, which the
// host reports as a failed launch set.  Both sweeps always run (a backward
// sweep after a forward one that changed nothing changes nothing either).

constexpr int kStripThreads = 64;          // one wave
constexpr int kStripLanePx = 1;            // scan positions per lane (1: shortest line steps; 2 is 20 % slower)
constexpr int kStripPx = kStripThreads * kStripLanePx;
constexpr int kSpinLimit = 1 << 20;       // polls (each with a short sleep) before a wait gives up

__device__ inline unsigned long long pack_carry(int v, int s) {
    return ((unsigned long long)(uint32_t)(v | (1 << 30)) << 32) | (uint32_t)(s + 1);
}
__device__ inline void unpack_carry(unsigned long long c, int& v, int& s) {
    v = (int)((uint32_t)(c >> 32) & ((1u << 30) - 1u));
    s = (int)(uint32_t)c - 1;
}
__device__ inline unsigned long long load_slot(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The strip solver's wave scan of the line functions (Fn), in two words per
// lane: p = 2 cv + t (t is 0 or 1) and cs; the k of every composite is its
// number of scan positions (each lane's function covers kStripLanePx of them,
// each of k 1), known from the lane, so k is not shuffled.  compose_p is
// compose() on that form: cv1 + k2 < cv2 + t2 holds iff 2 (cv1 + k2) <
// p2 + t2 (the left side is even), and the composite's p is then
// 2 (cv1 + k2) + t1, else p2; apply_p is apply_fn() likewise.  Distances stay
// below kAbsent + 2^20 (line lengths and counts are far below 2^20), so the
// doubled values stay below 2^30 + 2^21, under the identity's p (1.5 * 2^30,
// which leaves room for 2 k below 2^31).  Lanes whose DPP source lies outside
// the shift take the identity, which compose_p leaves alone.
constexpr int kPackIdent = 0x60000000;

__device__ inline void compose_p(int p1, int cs1, int& p2, int& cs2, int k2) {
    const int lhs = (p1 & ~1) + 2 * k2;
    const bool a = lhs < p2 + (p2 & 1);
    p2 = a ? lhs + (p1 & 1) : p2;
    cs2 = a ? cs1 : cs2;
}

__device__ inline void apply_p(int p, int cs, int k, int& v, int& s) {
    const bool a = 2 * (v + k) < p + (p & 1);
    s = a ? s : cs;
    v = a ? v + k : (p >> 1);
}

template <int CTRL, int ROWS>
__device__ inline void scan_level_p(int& p, int& cs, int k2) {
    const int sp = __builtin_amdgcn_update_dpp(kPackIdent, p, CTRL, ROWS, 0xf, false);
    const int scs = __builtin_amdgcn_update_dpp(-1, cs, CTRL, ROWS, 0xf, false);
    compose_p(sp, scs, p, cs, k2);
}

// Inclusive scan of (p, cs) over the wave: DPP row shifts by 1, 2, 4, 8, then
// row broadcasts, each level composing with the number of positions this
// lane's composite covers before it; ep, ecs: the composite of the lanes
// before this one (the identity for lane 0), which covers PX * lane positions.
template <int PX>
__device__ inline void wave_scan_p(int& p, int& cs, int& ep, int& ecs) {
    const int lane = __lane_id(), r = lane & 15, row = lane >> 4;
    scan_level_p<0x111, 0xf>(p, cs, PX);                           // row_shr:1
    scan_level_p<0x112, 0xf>(p, cs, PX * min(2, r + 1));           // row_shr:2
    scan_level_p<0x114, 0xf>(p, cs, PX * min(4, r + 1));           // row_shr:4
    scan_level_p<0x118, 0xf>(p, cs, PX * min(8, r + 1));           // row_shr:8
    scan_level_p<0x142, 0xa>(p, cs, PX * (r + 1));                 // row_bcast:15 into rows 1, 3
    scan_level_p<0x143, 0xc>(p, cs, PX * (r + 1 + 16 * (row & 1)));   // row_bcast:31 into rows 2, 3
    ep = __builtin_amdgcn_update_dpp(kPackIdent, p, 0x138, 0xf, 0xf, false);   // wave_shr:1
    ecs = __builtin_amdgcn_update_dpp(-1, cs, 0x138, 0xf, 0xf, false);
}

// Strip s of component R in one sweep: scan positions [s P, min(llen,
// (s+1) P)) of every line, P = kStripPx, kStripLanePx consecutive ones per
// lane; in_slots / out_slots: the carries of the strip before / this strip,
// one per line (out_slots null for the last).  Same candidate order and line
// order semantics as solve_lines.  The raw map words of a line are loaded
// kRing - 1 steps before their first use (a ring of kRing lines in registers,
// indexed statically by unrolling the line loop kRing times) and interpreted
// only when used, so no step waits on a load issued in the same step.
constexpr int kRing = 4;

struct RawPx {
    int lab, d, s, th;   // th: the hole flag of the pixel's tile (else lab is not stored)
};

template <bool FIRST>
__device__ void solve_strip(const Region& R, const InpaintWS& ws, bool cols, bool fwd, int s,
                            const unsigned long long* in_slots, unsigned long long* out_slots) {
    constexpr int PX = kStripLanePx;
    const int lane = threadIdx.x & 63;
    const int nlines = cols ? R.bw : R.bh;
    const int llen = cols ? R.bh : R.bw;
    const int e0 = s * kStripPx, e1 = min(llen, e0 + kStripPx);
    const int me = e0 + lane * PX;   // this lane's first scan position
    const int rs = cols ? (fwd ? 2 : 3) : (fwd ? 1 : 4);
    const int r_prev = cols ? (fwd ? 1 : 4) : (fwd ? 2 : 3);
    const int r_next = cols ? (fwd ? 4 : 1) : (fwd ? 3 : 2);
    const int r_along = cols ? (fwd ? 3 : 2) : (fwd ? 4 : 1);
    auto lcoord = [&](int li) { return fwd ? li : nlines - 1 - li; };
    // view index of (line coordinate lc, scan index e); false outside the image
    // (ti: its thole entry)
    auto addr = [&](int lc, int e, int& q, int& ti) {
        const int pos = fwd ? e : llen - 1 - e;
        const int x = cols ? R.x0 + lc : R.x0 + pos;
        const int y = cols ? R.y0 + pos : R.y0 + lc;
        const bool in = x >= 0 && x < R.w && y >= 0 && y < R.h;
        q = R.base + (in ? y * R.w + x : 0);
        ti = R.tb + (in ? (y / kCclTile) * ws.ttx + x / kCclTile : 0);
        return in;
    };
    auto fetch = [&](int lc, int e, RawPx& r) {
        int q, ti;
        addr(lc, e, q, ti);
        const int4 m = ws.px[q];
        r.lab = m.z;
        r.d = m.x;
        r.s = m.y;
        r.th = ws.thole[ti];
    };
    auto value = [&](int lc, int e, const RawPx& r, int& v, int& sr) {
        int q, ti;
        const bool in = addr(lc, e, q, ti);
        map_value(R, r.th ? r.lab : -1, r.d, r.s, q, v, sr);
        return in;
    };
    // ring: this lane's pixels, the along-line neighbour after them, (first
    // strip) the pixel before the line, (later strips) the carry of the strip
    // before -- zero while not yet written; every lane reads the same slot
    RawPx rp[kRing][PX], ra[kRing], rl[kRing];
    unsigned long long rc[kRing];
#pragma unroll
    for (int k = 0; k < kRing; ++k) {
#pragma unroll
        for (int j = 0; j < PX; ++j) fetch(lcoord(k), me + j, rp[k][j]);
        fetch(lcoord(k), me + PX, ra[k]);
        if (FIRST) fetch(lcoord(k), -1, rl[k]);
        else rc[k] = load_slot(&in_slots[min(k, nlines - 1)]);
    }
    int pv[PX], ps[PX];   // the previous line (this sweep) at this lane's positions
    bool pin[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
        RawPx r;
        fetch(lcoord(-1), me + j, r);
        pin[j] = value(lcoord(-1), me + j, r, pv[j], ps[j]);
    }
    bool gave_up = false;   // after one wait gave up, later ones do not wait
    auto step = [&](int li, const int k) {
        const int k1 = (k + 1) % kRing;
        const int lc = lcoord(li), lcn = lcoord(li + 1);
        int sv[PX], ss[PX];
        bool sin[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) sin[j] = value(lc, me + j, rp[k][j], sv[j], ss[j]);
        int av, as_;
        const bool ain = value(lc, me + PX, ra[k], av, as_);
        // this lane's functions: past the strip (kAbsent, -1, 0), known (0,
        // itself), a hole of another component (kAbsent, -1, 0), a member: its
        // best candidate among the neighbours other than the scan input
        Fn f[PX];
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            int nv, ns;
            const bool nin = value(lcn, me + j, rp[k1][j], nv, ns);
            const int alv = j + 1 < PX ? sv[j + 1] : av, als = j + 1 < PX ? ss[j + 1] : as_;
            const bool alin = j + 1 < PX ? sin[j + 1] : ain;
            const bool valid = me + j < e1;
            const bool known = valid & (sv[j] == 0);
            const bool member = valid & (sv[j] != 0) & (sv[j] != kOther);
            int bv = sv[j], br = 0, bs = ss[j];
            cand_if(pin[j] & (pv[j] != kOther), bv, br, bs, pv[j] + 1, r_prev, ps[j]);
            cand_if(nin & (nv != kOther), bv, br, bs, nv + 1, r_next, ns);
            cand_if(alin & (alv != kOther), bv, br, bs, alv + 1, r_along, als);
            f[j].k = 1;
            f[j].cv = known ? 0 : (member ? bv : kAbsent);
            f[j].cs = known ? ss[j] : (member ? bs : -1);
            f[j].t = (member & (rs < br)) ? 1 : 0;
        }
        // the scan input of the strip's first position
        int Sv, Ss;
        if (FIRST) {
            int v, sr;
            const bool in = value(lc, -1, rl[k], v, sr);
            const bool known = in && v == 0;
            Sv = known ? 0 : kAbsent;
            Ss = known ? sr : -1;
        } else {
            // the carry of the strip before, loaded kRing steps ahead; polled
            // again (bounded) while it is not yet written
            unsigned long long c = rc[k];
            for (int n = 0; c == 0ull; ++n) {
                if (gave_up || n >= kSpinLimit) {
                    if (ws.fail && lane == 0) atomicOr(ws.fail, 1u);
                    gave_up = true;
                    c = pack_carry(kAbsent, -1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                c = load_slot(&in_slots[li]);
            }
            unpack_carry(c, Sv, Ss);
        }
        // line li + kRing into the slots this line frees
#pragma unroll
        for (int j = 0; j < PX; ++j) fetch(lcoord(li + kRing), me + j, rp[k][j]);
        fetch(lcoord(li + kRing), me + PX, ra[k]);
        if (FIRST) fetch(lcoord(li + kRing), -1, rl[k]);
        else rc[k] = load_slot(&in_slots[min(li + kRing, nlines - 1)]);   // past the end: a slot never read
        Fn A = f[0];
#pragma unroll
        for (int j = 1; j < PX; ++j) A = compose(A, f[j]);
        int ap = 2 * A.cv + A.t, acs = A.cs, ep, ecs;
        wave_scan_p<PX>(ap, acs, ep, ecs);
        apply_p(ep, ecs, PX * lane, Sv, Ss);   // the identity for lane 0
#pragma unroll
        for (int j = 0; j < PX; ++j) {
            apply_fn(f[j], Sv, Ss);
            const int e = me + j;
            const bool member = e < e1 && sv[j] != 0 && sv[j] != kOther;
            if (member && Sv != sv[j]) {
                int q, ti;
                addr(lc, e, q, ti);
                *reinterpret_cast<int2*>(&ws.px[q]) = make_int2(Sv, Ss);
            }
            if (out_slots && e == e1 - 1)
                __hip_atomic_store(&out_slots[li], pack_carry(Sv, Ss), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pv[j] = member ? Sv : sv[j];
            ps[j] = member ? Ss : ss[j];
            pin[j] = sin[j];
        }
    };
    for (int li0 = 0; li0 < nlines; li0 += kRing) {
#pragma unroll
        for (int k = 0; k < kRing; ++k)
            if (li0 + k < nlines) step(li0 + k, k);
    }
}

// Which large components the strip solver takes: tasks (component, strip) in
// component order with the strips of a component consecutive, and their carry
// slots, while both fit the workspace; the rest go to the one-block solver.
// One block: a scan over the components in chunks of 1024.
__global__ void __launch_bounds__(1024) k_inp_plan(InpaintWS ws) {
    __shared__ int wsum[2][16];
    __shared__ int s_base[2];
    __shared__ int s_used[2];
    __shared__ int s_legacy;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncomp = ws.counters[0];
    if (tid == 0) { s_base[0] = s_base[1] = 0; s_used[0] = s_used[1] = 0; s_legacy = 0; }
    __syncthreads();
    for (int c0 = 0; c0 < ncomp; c0 += 1024) {
        const int c = c0 + tid;
        int S = 0, nslots = 0;
        bool big = false, cols = false;
        if (c < ncomp) {
            const int bw = ws.comp[4 * c + 1] - ws.comp[4 * c + 0] + 1;
            const int bh = ws.comp[4 * c + 3] - ws.comp[4 * c + 2] + 1;
            big = !(bw * bh <= kLdsCap || (bw > kLineMax && bh > kLineMax));
            if (big) {
                cols = bh <= kLineMax && (bw > kLineMax || bh > bw);   // the orientation of k_inp_big
                const int nlines = cols ? bw : bh, llen = cols ? bh : bw;
                S = (llen + kStripPx - 1) / kStripPx;
                nslots = S * nlines;
            }
        }
        // inclusive block scan of (S, nslots) in component order
        int a = S, b = nslots;
        for (int off = 1; off < 64; off <<= 1) {
            const int oa = __shfl_up(a, off), ob = __shfl_up(b, off);
            if (lane >= off) { a += oa; b += ob; }
        }
        if (lane == 63) { wsum[0][wave] = a; wsum[1][wave] = b; }
        __syncthreads();
        for (int w = 0; w < wave; ++w) { a += wsum[0][w]; b += wsum[1][w]; }
        const int ta = s_base[0] + a, tb = s_base[1] + b;   // inclusive ends
        // prefix sums only grow, so the accepted components are a prefix of
        // the candidates and their tasks leave no gaps
        const bool fits = big && ta <= ws.task_cap && tb <= ws.slot_cap;
        if (fits) {
            for (int k = 0; k < S; ++k)
                ws.tasks[ta - S + k] = make_int4(c, k | (S << 8) | (cols ? 1 << 16 : 0), tb - nslots, 0);
            atomicMax(&s_used[0], ta);
            atomicMax(&s_used[1], tb);
        } else if (big) {
            ws.legacy[atomicAdd(&s_legacy, 1)] = c;
        }
        __syncthreads();
        // clamped past the capacities: nothing later fits, and the sums cannot overflow
        if (tid == 1023) { s_base[0] = min(ta, ws.task_cap + 1); s_base[1] = min(tb, ws.slot_cap + 1); }
        __syncthreads();
    }
    if (tid == 0) {
        ws.counters[3] = s_used[0];
        ws.counters[6] = s_legacy;
        ws.counters[7] = s_used[1];
    }
}

// Zeroes the carry slots the plan handed out (both sweeps).
__global__ void k_inp_slot_reset(InpaintWS ws) {
    const size_t n = (size_t)ws.counters[7];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        ws.slots[0][i] = 0ull;
        ws.slots[1][i] = 0ull;
    }
}

__device__ inline int view_of_pixel(const ViewDesc* views, int nviews, int gp);

template <bool FWD>
__global__ void __launch_bounds__(kStripThreads) k_inp_strips(const ViewDesc* __restrict__ views, int nviews,
                                                              InpaintWS ws) {
    const int ntasks = ws.counters[3];
    while (true) {
        // one wave: the ticket is taken by lane 0 and broadcast
        int t = 0;
        if (threadIdx.x == 0) t = atomicAdd(&ws.counters[FWD ? 4 : 5], 1);
        t = __shfl(t, 0);
        if (t >= ntasks) break;
        const int4 task = ws.tasks[t];
        const int c = task.x, s = task.y & 0xff, S = (task.y >> 8) & 0xff;
        const bool cols = (task.y >> 16) & 1;
        Region R;
        R.root = ws.root[c];
        const int vi = view_of_pixel(views, nviews, R.root);
        const ViewDesc& vd = views[vi];
        R.base = (int)vd.px_base;
        R.tb = vi * ws.tpv;
        R.x0 = ws.comp[4 * c + 0];
        R.y0 = ws.comp[4 * c + 2];
        R.bw = ws.comp[4 * c + 1] - R.x0 + 1;
        R.bh = ws.comp[4 * c + 3] - R.y0 + 1;
        R.w = vd.w; R.h = vd.h;
        R.ldist = nullptr; R.lsrc = nullptr;
        const int nlines = cols ? R.bw : R.bh;
        unsigned long long* sl = ws.slots[FWD ? 0 : 1] + task.z;
        unsigned long long* out = s + 1 < S ? sl + (size_t)s * nlines : nullptr;
        if (s == 0) solve_strip<true>(R, ws, cols, FWD, s, nullptr, out);
        else solve_strip<false>(R, ws, cols, FWD, s, sl + (size_t)(s - 1) * nlines, out);
    }
}

// View owning global pixel gp (views are sorted by px_base).
__device__ inline int view_of_pixel(const ViewDesc* views, int nviews, int gp) {
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)views[mid].px_base <= gp) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_inp_solve(const ViewDesc* __restrict__ views, int nviews, InpaintWS ws) {
    __shared__ Fn wave_tot[8];
    __shared__ int s_carry[2];
    __shared__ int s_comp;
    __shared__ int ldist[kLdsCap];
    __shared__ int lsrc[kLdsCap];
    const int tid = threadIdx.x;
    const int ncomp = ws.counters[0];
    // two passes over the components: the large ones that k_inp_big does not
    // take (both box sides beyond its line length; they take longest) are all
    // started first -- found by a static stride, so the small ones cost no
    // queue atomics here -- then the small ones are pulled from a queue
    for (int pass = 0; pass < 2; ++pass) {
    int next = blockIdx.x;
    while (true) {
        int c;
        if (pass == 0) {
            c = next;
            next += gridDim.x;
        } else {
            if (tid == 0) s_comp = atomicAdd(&ws.counters[2], 1);
            __syncthreads();
            c = s_comp;
            __syncthreads();
        }
        if (c >= ncomp) break;
        {
            const int bw = ws.comp[4 * c + 1] - ws.comp[4 * c + 0] + 1, bh = ws.comp[4 * c + 3] - ws.comp[4 * c + 2] + 1;
            const bool small = bw * bh <= kLdsCap;
            if (!small && (bw <= kLineMax || bh <= kLineMax)) continue;   // k_inp_big's
            if (small ? pass == 0 : pass == 1) continue;   // the same decision in every thread of the block
        }
        Region R;
        R.root = ws.root[c];
        const int vi = view_of_pixel(views, nviews, R.root);
        const ViewDesc& vd = views[vi];
        R.base = (int)vd.px_base;
        R.tb = vi * ws.tpv;
        R.x0 = ws.comp[4 * c + 0];
        R.y0 = ws.comp[4 * c + 2];
        R.bw = ws.comp[4 * c + 1] - R.x0 + 1;
        R.bh = ws.comp[4 * c + 3] - R.y0 + 1;
        R.w = vd.w; R.h = vd.h;
        R.ldist = ldist; R.lsrc = lsrc;
        const int area = R.bw * R.bh;
        if (area <= kLdsCap) {
            for (int k = tid; k < area; k += 256) {
                const int x = R.x0 + k % R.bw, y = R.y0 + k / R.bw;
                const int p = R.base + y * R.w + x;
                const int lab = label_at(R, ws, x, y, p);
                const int4 m = ws.px[p];
                ldist[k] = lab < 0 ? 0 : (lab == R.root ? m.x : kOther);
                lsrc[k] = lab < 0 ? p : (lab == R.root ? m.y : -1);
            }
            __syncthreads();
            solve_component<true>(R, ws, wave_tot, s_carry);
            __syncthreads();
            for (int k = tid; k < area; k += 256) {
                const int x = R.x0 + k % R.bw, y = R.y0 + k / R.bw;
                const int p = R.base + y * R.w + x;
                if (label_at(R, ws, x, y, p) == R.root)
                    *reinterpret_cast<int2*>(&ws.px[p]) = make_int2(ldist[k], lsrc[k]);
            }
        } else {
            solve_big(R, ws, wave_tot, s_carry);
        }
        __syncthreads();
    }
    }
}

// Components too large for the LDS copy whose lines (along one side of the
// box) fit one step of a 1024-thread block and that the strip plan could not
// place: one block each, static stride.
__global__ void __launch_bounds__(kBigThreads) k_inp_big(const ViewDesc* __restrict__ views, int nviews,
                                                         InpaintWS ws) {
    __shared__ Fn wave_tot[2 * kBigThreads / 64];
    const int nlegacy = ws.counters[6];
    for (int k = blockIdx.x; k < nlegacy; k += gridDim.x) {
        const int c = ws.legacy[k];
        const int bw = ws.comp[4 * c + 1] - ws.comp[4 * c + 0] + 1, bh = ws.comp[4 * c + 3] - ws.comp[4 * c + 2] + 1;
        Region R;
        R.root = ws.root[c];
        const int vi = view_of_pixel(views, nviews, R.root);
        const ViewDesc& vd = views[vi];
        R.base = (int)vd.px_base;
        R.tb = vi * ws.tpv;
        R.x0 = ws.comp[4 * c + 0];
        R.y0 = ws.comp[4 * c + 2];
        R.bw = bw; R.bh = bh;
        R.w = vd.w; R.h = vd.h;
        R.ldist = nullptr; R.lsrc = nullptr;
        // fewer lines when both orientations fit
        const bool cols = bh <= kLineMax && (bw > kLineMax || bh > bw);
        solve_lines<kBigThreads, kLinePx>(R, ws, cols, wave_tot);
        __syncthreads();
    }
}

// Copies the chosen known pixel's colour into every hole of the hole tiles.
__global__ void __launch_bounds__(256) k_inp_apply(const ViewDesc* __restrict__ views, InpaintWS ws) {
    const int nt = ws.counters[8];
    for (int i = blockIdx.x; i < nt; i += gridDim.x) {
        const HoleTile H = hole_tile(ws, i);
        const ViewDesc& vd = views[H.v];
        for (int k = threadIdx.x; k < kCclTile * kCclTile; k += 256) {
            const int x = H.tx0 + (k % kCclTile), y = H.ty0 + (k / kCclTile);
            if (x >= vd.w || y >= vd.h) continue;
            const int p = y * vd.w + x, gp = (int)vd.px_base + p;
            const int4 m = ws.px[gp];
            if (m.z < 0) continue;
            const int s = m.y - (int)vd.px_base;
            if (s != p) {
                float* color = vd.color;
                color[3 * p] = color[3 * s];
                color[3 * p + 1] = color[3 * s + 1];
                color[3 * p + 2] = color[3 * s + 2];
            }
        }
    }
}

// The inpainting map of a single-view launch: source x, source y and
// distance per pixel (a known pixel: itself, 0).
__global__ void k_inp_map(const ViewDesc* __restrict__ views, InpaintWS ws, int32_t* map_out) {
    RVS_VIEW_PIXEL(vd, p, gp)
    const int x = p % vd.w, y = p / vd.w;
    const bool th = ws.thole[blockIdx.y * ws.tpv + (y / kCclTile) * ws.ttx + x / kCclTile] != 0;
    const int4 m = ws.px[gp];
    const bool hole = th && m.z >= 0;
    const int s = hole ? m.y - (int)vd.px_base : p;
    map_out[3 * p] = s % vd.w;
    map_out[3 * p + 1] = s / vd.w;
    map_out[3 * p + 2] = hole ? m.x : 0;
}

namespace {
// Strip tasks: a large component spans at least its box's longer side in
// pixels and has a box over kLdsCap pixels, so npx / 64 tasks cover every
// realistic case; carry slots: one per strip and line.  Components beyond
// either bound go to the one-block solver.
size_t task_cap_of(size_t npx, int nviews) { return npx / 64 + 16 * (size_t)nviews + 16; }
size_t slot_cap_of(size_t npx, int nviews) { return npx / 16 + 4096 * (size_t)nviews; }
size_t legacy_cap_of(size_t npx, int nviews) { return npx / 128 + 2 * (size_t)nviews + 1; }
size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }
int tiles_of(int n) { return (n + kCclTile - 1) / kCclTile; }
size_t grid_tiles(int nviews, int max_w, int max_h) { return (size_t)nviews * tiles_of(max_w) * tiles_of(max_h); }
}  // namespace

size_t inpaint_ws_bytes(size_t npx, int nviews, int max_w, int max_h) {
    const size_t maxc = npx / 2 + (size_t)nviews + 1;
    const size_t nt = grid_tiles(nviews, max_w, max_h);
    return a256(task_cap_of(npx, nviews) * sizeof(int4)) +
           2 * a256(slot_cap_of(npx, nviews) * sizeof(unsigned long long)) +
           a256(4 * legacy_cap_of(npx, nviews)) + a256(nt) + a256(4 * nt) + 4 * (4 * npx) + 4 * (4 * maxc) +
           4 * maxc + 64;
}

InpaintWS inpaint_ws_carve(void* base, size_t npx, int nviews, int max_w, int max_h) {
    const size_t maxc = npx / 2 + (size_t)nviews + 1;
    const size_t nt = grid_tiles(nviews, max_w, max_h);
    char* b = (char*)base;   // 256-byte aligned
    InpaintWS ws;
    ws.task_cap = (int)task_cap_of(npx, nviews);
    ws.slot_cap = (int)slot_cap_of(npx, nviews);
    ws.ttx = tiles_of(max_w);
    ws.tpv = ws.ttx * tiles_of(max_h);
    ws.thole = (uint8_t*)b; b += a256(nt);
    ws.htiles = (int*)b; b += a256(4 * nt);
    ws.tasks = (int4*)b; b += a256(task_cap_of(npx, nviews) * sizeof(int4));
    ws.slots[0] = (unsigned long long*)b; b += a256(slot_cap_of(npx, nviews) * sizeof(unsigned long long));
    ws.slots[1] = (unsigned long long*)b; b += a256(slot_cap_of(npx, nviews) * sizeof(unsigned long long));
    ws.legacy = (int*)b; b += a256(4 * legacy_cap_of(npx, nviews));
    ws.px = (int4*)b; b += 16 * npx;
    int* p = (int*)b;
    ws.comp = p; p += 4 * maxc;
    ws.root = p; p += maxc;
    ws.counters = p;
    ws.fail = nullptr;
    ws.rflags = nullptr;
    return ws;
}

hipError_t launch_inpaint(const ViewDesc* views, int nviews, int max_w, int max_h, const InpaintWS& ws,
                          int32_t* map_out, hipStream_t s) {
    const dim3 b(256);
    hipError_t e = launch_reset_u32((uint32_t*)ws.counters, 16, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ccl_local, dim3(tiles_of(max_w), tiles_of(max_h), nviews), b, 0, s, views, ws);
    // the passes over the hole tiles: persistent blocks striding over the list
    const size_t nt = grid_tiles(nviews, max_w, max_h);
    const dim3 gt((unsigned)(nt < (size_t)kHoleTileBlocks ? nt : (size_t)kHoleTileBlocks));
    hipLaunchKernelGGL(k_ccl_merge, gt, dim3(kMergeThreads), 0, s, views, ws);
    hipLaunchKernelGGL(k_ccl_flatten, gt, b, 0, s, views, ws);
    hipLaunchKernelGGL(k_comp_bbox, gt, dim3(kBboxThreads), 0, s, views, ws);
    hipLaunchKernelGGL(k_inp_plan, dim3(1), dim3(1024), 0, s, ws);
    hipLaunchKernelGGL(k_inp_slot_reset, dim3(256), dim3(256), 0, s, ws);
    hipLaunchKernelGGL(k_inp_strips<true>, dim3(kStripBlocks), dim3(kStripThreads), 0, s, views, nviews, ws);
    hipLaunchKernelGGL(k_inp_strips<false>, dim3(kStripBlocks), dim3(kStripThreads), 0, s, views, nviews, ws);
    hipLaunchKernelGGL(k_inp_big, dim3(kInpaintBigBlocks), dim3(kBigThreads), 0, s, views, nviews, ws);
    hipLaunchKernelGGL(k_inp_solve, dim3(kInpaintBlocks), b, 0, s, views, nviews, ws);
    hipLaunchKernelGGL(k_inp_apply, gt, b, 0, s, views, ws);
    if (map_out)
        hipLaunchKernelGGL(k_inp_map, dim3((unsigned)(((size_t)max_w * max_h + 255) / 256), (unsigned)nviews), b, 0, s,
                           views, ws, map_out);
    return hipGetLastError();
}

}  // namespace rvs_k
