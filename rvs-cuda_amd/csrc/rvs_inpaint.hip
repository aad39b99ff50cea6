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
// changing, which is what the global loop gives every component.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvs_inpaint.h"

namespace rvs_k {

constexpr int kAbsent = 1 << 29;

__device__ inline int ld(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline int find_root(const int* label, int p) {
    while (true) {
        const int q = ld(&label[p]);
        if (q == p) return p;
        p = q;
    }
}

__device__ inline void unite(int* label, int a, int b) {
    while (true) {
        a = find_root(label, a);
        b = find_root(label, b);
        if (a == b) return;
        if (a < b) { const int t = a; a = b; b = t; }
        const int old = atomicMin(&label[a], b);
        if (old == a) return;
        a = old;
    }
}

__global__ void k_inp_init(const float* __restrict__ quality, int w, int h, InpaintWS ws) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const bool hole = !(quality[p] > 0.f);   // View::get_inpaint_mask (View.cpp:198-203)
    ws.label[p] = hole ? p : -1;
    ws.dist[p] = hole ? (w + h) : 0;
    ws.src[p] = p;
    if (p == 0) { ws.counters[0] = 0; ws.counters[1] = 0; }
}

__global__ void k_ccl_merge(int w, int h, InpaintWS ws) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    if (ws.label[p] < 0) return;
    const int x = p % w, y = p / w;
    if (x + 1 < w && ws.label[p + 1] >= 0) unite(ws.label, p, p + 1);
    if (y + 1 < h && ws.label[p + w] >= 0) unite(ws.label, p, p + w);
}

__global__ void k_ccl_flatten(int w, int h, InpaintWS ws) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    if (ws.label[p] < 0) return;
    const int r = find_root(ws.label, p);
    ws.label[p] = r;
    if (r == p) {
        const int id = atomicAdd(&ws.counters[0], 1);
        ws.cid[p] = id;
        ws.comp[4 * id + 0] = INT_MAX;
        ws.comp[4 * id + 1] = INT_MIN;
        ws.comp[4 * id + 2] = INT_MAX;
        ws.comp[4 * id + 3] = INT_MIN;
        ws.root[id] = p;
    }
}

__global__ void k_comp_bbox(int w, int h, InpaintWS ws) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const int r = ws.label[p];
    if (r < 0) return;
    const int id = ws.cid[r];
    const int x = p % w, y = p / w;
    atomicMin(&ws.comp[4 * id + 0], x);
    atomicMax(&ws.comp[4 * id + 1], x);
    atomicMin(&ws.comp[4 * id + 2], y);
    atomicMax(&ws.comp[4 * id + 3], y);
}

// S -> (S.v + k < C.v + t) ? (S.v + k, S.src) : C
struct Fn {
    int cv, cs, t, k;
};

__device__ inline Fn compose(const Fn& f1, const Fn& f2) {   // f2 after f1
    Fn r;
    if (f1.cv + f2.k < f2.cv + f2.t) { r.cv = f1.cv + f2.k; r.cs = f1.cs; r.t = f1.t; }
    else { r.cv = f2.cv; r.cs = f2.cs; r.t = f2.t; }
    r.k = f1.k + f2.k;
    return r;
}

// lexicographic (value, rank) minimum, first argument wins ties of both
__device__ inline void cand(int& bv, int& br, int& bs, int v, int r, int s) {
    if (v < bv || (v == bv && r < br)) { bv = v; br = r; bs = s; }
}

__global__ void __launch_bounds__(256) k_inp_solve(int w, int h, InpaintWS ws) {
    __shared__ Fn wave_tot[4];
    __shared__ int s_comp;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    while (true) {
        if (tid == 0) s_comp = atomicAdd(&ws.counters[1], 1);
        __syncthreads();
        const int c = s_comp;
        __syncthreads();
        if (c >= ws.counters[0]) return;
        const int root = ws.root[c];
        const int x0 = ws.comp[4 * c + 0], x1 = ws.comp[4 * c + 1];
        const int y0 = ws.comp[4 * c + 2], y1 = ws.comp[4 * c + 3];
        const int cw = x1 - x0 + 1;
        // Every sweep that changes something strictly lowers a distance, so the loop
        // ends; the bound only guards against a broken invariant.
        const int max_sweeps = 4 * (w + h) + 8;
        for (int sweep = 0; sweep < max_sweeps; ++sweep) {
            const bool fwd = (sweep & 1) == 0;
            bool changed = false;
            for (int yy = 0; yy <= y1 - y0; ++yy) {
                const int y = fwd ? y0 + yy : y1 - yy;
                // carry: the pixel before the first element in scan order
                int Lv = kAbsent, Ls = -1;
                {
                    const int xb = fwd ? x0 - 1 : x1 + 1;
                    if (xb >= 0 && xb < w) {
                        const int pb = y * w + xb;
                        if (ws.label[pb] < 0) { Lv = 0; Ls = pb; }
                    }
                }
                for (int cb = 0; cb < cw; cb += 256) {
                    const int e = cb + tid;
                    const bool in = e < cw;
                    const int x = fwd ? x0 + e : x1 - e;
                    const int p = y * w + x;
                    Fn f;
                    f.k = 1;
                    bool member = false;
                    int cur_v = 0;
                    if (!in) {
                        f.cv = kAbsent; f.cs = -1; f.t = 0;
                    } else {
                        const int lab = ws.label[p];
                        if (lab < 0) {
                            f.cv = 0; f.cs = p; f.t = 0;           // known pixel: (0, itself)
                        } else if (lab != root) {
                            f.cv = kAbsent; f.cs = -1; f.t = 0;    // another component (never a neighbour)
                        } else {
                            member = true;
                            cur_v = ws.dist[p];
                            int bv = cur_v, br = 0, bs = ws.src[p];
                            // neighbours in the reference order: left(1) up(2) down(3) right(4);
                            // the scan input is left (forward) or right (backward)
                            if (!fwd && x > 0) cand(bv, br, bs, ws.dist[p - 1] + 1, 1, ws.src[p - 1]);
                            if (y > 0) cand(bv, br, bs, ws.dist[p - w] + 1, 2, ws.src[p - w]);
                            if (y + 1 < h) cand(bv, br, bs, ws.dist[p + w] + 1, 3, ws.src[p + w]);
                            if (fwd && x + 1 < w) cand(bv, br, bs, ws.dist[p + 1] + 1, 4, ws.src[p + 1]);
                            const int rs = fwd ? 1 : 4;
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
                    __syncthreads();   // all reads of this row chunk are done before any write
                    Fn pre; pre.cv = kAbsent; pre.cs = -1; pre.t = 0; pre.k = 0;
                    bool have = false;
                    for (int wv = 0; wv < wave; ++wv) {
                        pre = have ? compose(pre, wave_tot[wv]) : wave_tot[wv];
                        have = true;
                    }
                    if (have) f = compose(pre, f);
                    // apply to the carry
                    int rv, rsrc;
                    if (Lv + f.k < f.cv + f.t) { rv = Lv + f.k; rsrc = Ls; } else { rv = f.cv; rsrc = f.cs; }
                    if (member && rv != cur_v) {
                        changed = true;
                        ws.dist[p] = rv;
                        ws.src[p] = rsrc;
                    }
                    // carry out: result at the last element of the chunk
                    __shared__ int s_carry[2];
                    const int last = min(256, cw - cb) - 1;
                    if (tid == last) { s_carry[0] = rv; s_carry[1] = rsrc; }
                    __syncthreads();
                    Lv = s_carry[0]; Ls = s_carry[1];
                    __syncthreads();
                }
            }
            if (!__syncthreads_or(changed)) break;
        }
    }
}

__global__ void k_inp_apply(float* __restrict__ color, int w, int h, InpaintWS ws, int32_t* map_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const int s = ws.src[p];
    if (map_out) {
        map_out[3 * p] = s % w;
        map_out[3 * p + 1] = s / w;
        map_out[3 * p + 2] = ws.dist[p];
    }
    if (ws.label[p] >= 0 && s != p) {
        color[3 * p] = color[3 * s];
        color[3 * p + 1] = color[3 * s + 1];
        color[3 * p + 2] = color[3 * s + 2];
    }
}

size_t inpaint_ws_bytes(int w, int h) {
    const size_t n = (size_t)w * h;
    const size_t maxc = n / 2 + 1;
    return 4 * (4 * n) + 4 * (4 * maxc) + 4 * maxc + 64;
}

InpaintWS inpaint_ws_carve(void* base, int w, int h) {
    const size_t n = (size_t)w * h;
    const size_t maxc = n / 2 + 1;
    int* p = (int*)base;
    InpaintWS ws;
    ws.label = p; p += n;
    ws.dist = p; p += n;
    ws.src = p; p += n;
    ws.cid = p; p += n;
    ws.comp = p; p += 4 * maxc;
    ws.root = p; p += maxc;
    ws.counters = p;
    return ws;
}

hipError_t launch_inpaint(float* color, const float* quality, int w, int h, const InpaintWS& ws, int32_t* map_out,
                          hipStream_t s) {
    const int n = w * h;
    const dim3 b(256), g((n + 255) / 256);
    hipLaunchKernelGGL(k_inp_init, g, b, 0, s, quality, w, h, ws);
    hipLaunchKernelGGL(k_ccl_merge, g, b, 0, s, w, h, ws);
    hipLaunchKernelGGL(k_ccl_flatten, g, b, 0, s, w, h, ws);
    hipLaunchKernelGGL(k_comp_bbox, g, b, 0, s, w, h, ws);
    hipLaunchKernelGGL(k_inp_solve, dim3(1024), b, 0, s, w, h, ws);
    hipLaunchKernelGGL(k_inp_apply, g, b, 0, s, color, w, h, ws, map_out);
    return hipGetLastError();
}

}  // namespace rvs_k
