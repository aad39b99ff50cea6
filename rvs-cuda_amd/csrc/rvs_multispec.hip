// rvs_multispec.hip -- the frequency split of multispectral blending
// (BlendedViewMultiSpec::blend, BlendedView.cpp:70-88; split_frequencies and
// calcBlurring, blending.cpp:62-104,371-394).
//
// The reference blurs a synthesized view with a masked box filter read from a
// double-precision integral image and divides each window sum by the window's
// count of masked pixels: ColorSpace YUV blurs Y only and truncates the sum to
// an integer first; ColorSpace RGB blurs the three channels and rounds the sums
// to float.  Here the integral image holds exact fixed-point sums (2^-40
// units, unsigned 64-bit with wrap-around, so a window sum is exact whenever
// the window itself fits) built in three passes:
//   k_ms_rows   one block per row: inclusive prefix along the row (wave scans
//               with a carry between 256-pixel chunks);
//   k_ms_cols   16 columns per block: 32 row groups sum their rows, one scan
//               over the group totals, then every group rewrites its rows as
//               the column prefix;
//   k_ms_split  per pixel: the window sums from four corners, the reference's
//               conversions and float division, and the low / high bands.
// NC (1 or 3) is the number of blurred channels.  Bytes per pixel for NC = 1:
// 20 read + 12 written (rows), 24 read + 12 written (cols), about 20 read + 24
// written (split, corners mostly from L2).
#include <hip/hip_runtime.h>

#include "rvs_stages.h"

namespace rvs_k {

namespace {

typedef unsigned long long u64;

constexpr float kFixScale = 1099511627776.0f;  // 2^40: one fixed-point unit is 2^-40
constexpr float kFixClamp = 4194304.0f;        // |v| <= 2^22 keeps v * 2^40 inside int64
constexpr int kColW = 16;                      // columns per k_ms_cols block
constexpr int kColG = 32;                      // row groups per k_ms_cols block

// quality > 0 with quality = shape / depth (SynthesizedView.cpp:440)
__device__ inline bool ms_mask(const float* __restrict__ shape, const float* __restrict__ depth, size_t i) {
    return shape[i] / depth[i] > 0.f;
}

// The integral image's terms of one pixel: its channels and a count of 1 when
// it is under the mask and finite (all NC channels).
template <int NC>
__device__ inline void ms_term(const float* __restrict__ color, const float* __restrict__ shape,
                               const float* __restrict__ depth, size_t i, u64 (&s)[NC], uint32_t& c) {
    float v[NC];
    bool fin = true;
    for (int k = 0; k < NC; ++k) { v[k] = color[3 * i + k]; fin = fin && isfinite(v[k]); }
    c = (fin && ms_mask(shape, depth, i)) ? 1u : 0u;
    for (int k = 0; k < NC; ++k) {
        const float cv = fminf(fmaxf(v[k], -kFixClamp), kFixClamp);
        s[k] = c ? (u64)(long long)(cv * kFixScale) : 0ull;   // exact for |v| >= 2^-16; truncated below
    }
}

template <int NC>
__global__ __launch_bounds__(256) void k_ms_rows(const float* __restrict__ color, const float* __restrict__ shape,
                                                 const float* __restrict__ depth, int w, size_t npx,
                                                 u64* __restrict__ isum, uint32_t* __restrict__ icnt) {
    __shared__ u64 wsum[NC][4];
    __shared__ uint32_t wcnt[4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const size_t row = (size_t)blockIdx.x * w;
    u64 carry_s[NC];
    for (int k = 0; k < NC; ++k) carry_s[k] = 0;
    uint32_t carry_c = 0;
    for (int x0 = 0; x0 < w; x0 += 256) {
        const int x = x0 + t;
        u64 s[NC];
        uint32_t c = 0;
        for (int k = 0; k < NC; ++k) s[k] = 0;
        if (x < w) ms_term<NC>(color, shape, depth, row + x, s, c);
        for (int d = 1; d < 64; d <<= 1) {
            u64 os[NC];
            for (int k = 0; k < NC; ++k) os[k] = __shfl_up(s[k], d, 64);
            const uint32_t oc = __shfl_up(c, d, 64);
            if (lane >= d) {
                for (int k = 0; k < NC; ++k) s[k] += os[k];
                c += oc;
            }
        }
        if (lane == 63) {
            for (int k = 0; k < NC; ++k) wsum[k][wv] = s[k];
            wcnt[wv] = c;
        }
        __syncthreads();
        for (int q = 0; q < wv; ++q) {
            for (int k = 0; k < NC; ++k) s[k] += wsum[k][q];
            c += wcnt[q];
        }
        for (int k = 0; k < NC; ++k) s[k] += carry_s[k];
        c += carry_c;
        if (x < w) {
            for (int k = 0; k < NC; ++k) isum[k * npx + row + x] = s[k];
            icnt[row + x] = c;
        }
        for (int k = 0; k < NC; ++k) carry_s[k] += (wsum[k][0] + wsum[k][1]) + (wsum[k][2] + wsum[k][3]);
        carry_c += (wcnt[0] + wcnt[1]) + (wcnt[2] + wcnt[3]);
        __syncthreads();
    }
}

template <int NC>
__global__ __launch_bounds__(kColW * kColG) void k_ms_cols(int w, int h, size_t npx, u64* __restrict__ isum,
                                                           uint32_t* __restrict__ icnt) {
    __shared__ u64 gsum[NC][kColG][kColW];
    __shared__ uint32_t gcnt[kColG][kColW];
    const int c = threadIdx.x % kColW, g = threadIdx.x / kColW;
    const int x = blockIdx.x * kColW + c;
    const int L = (h + kColG - 1) / kColG;
    const int y0 = min(h, g * L), y1 = min(h, y0 + L);
    u64 s[NC];
    for (int k = 0; k < NC; ++k) s[k] = 0;
    uint32_t n = 0;
    if (x < w)
        for (int y = y0; y < y1; ++y) {
            const size_t i = (size_t)y * w + x;
            for (int k = 0; k < NC; ++k) s[k] += isum[k * npx + i];
            n += icnt[i];
        }
    for (int k = 0; k < NC; ++k) gsum[k][g][c] = s[k];
    gcnt[g][c] = n;
    __syncthreads();
    if (g == 0) {   // exclusive scan of the group totals of column c
        u64 a[NC];
        for (int k = 0; k < NC; ++k) a[k] = 0;
        uint32_t b = 0;
        for (int q = 0; q < kColG; ++q) {
            for (int k = 0; k < NC; ++k) {
                const u64 ts = gsum[k][q][c];
                gsum[k][q][c] = a[k];
                a[k] += ts;
            }
            const uint32_t tc = gcnt[q][c];
            gcnt[q][c] = b;
            b += tc;
        }
    }
    __syncthreads();
    for (int k = 0; k < NC; ++k) s[k] = gsum[k][g][c];
    n = gcnt[g][c];
    if (x < w)
        for (int y = y0; y < y1; ++y) {
            const size_t i = (size_t)y * w + x;
            for (int k = 0; k < NC; ++k) {
                s[k] += isum[k * npx + i];
                isum[k * npx + i] = s[k];
            }
            n += icnt[i];
            icnt[i] = n;
        }
}

// window sum over the inclusive 2D prefix: corners (r, c) with row or column
// -1 the empty prefix
__device__ inline u64 ms_at(const u64* __restrict__ p, int w, int r, int c) {
    return (r < 0 || c < 0) ? 0ull : p[(size_t)r * w + c];
}
__device__ inline uint32_t ms_atc(const uint32_t* __restrict__ p, int w, int r, int c) {
    return (r < 0 || c < 0) ? 0u : p[(size_t)r * w + c];
}

template <int NC>
__global__ __launch_bounds__(256) void k_ms_split(const float* __restrict__ color, const float* __restrict__ shape,
                                                  const float* __restrict__ depth, int w, int h, size_t npx, int rad,
                                                  const u64* __restrict__ isum, const uint32_t* __restrict__ icnt,
                                                  float* __restrict__ low, float* __restrict__ high) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    float col[3], lo[3];
    for (int k = 0; k < 3; ++k) { col[k] = color[3 * i + k]; lo[k] = col[k]; }
    for (int k = 0; k < NC; ++k) lo[k] = 0.f;   // calcBlurring leaves 0 outside the mask
    if (ms_mask(shape, depth, i)) {
        // window [sx, ex) x [sy, ey), clipped at the borders (blending.cpp:89-98)
        const int sx = max(0, x - rad), ex = min(w, x + rad + 1);
        const int sy = max(0, y - rad), ey = min(h, y + rad + 1);
        const uint32_t N = ((ms_atc(icnt, w, ey - 1, ex - 1) - ms_atc(icnt, w, sy - 1, ex - 1)) -
                            ms_atc(icnt, w, ey - 1, sx - 1)) + ms_atc(icnt, w, sy - 1, sx - 1);
        for (int k = 0; k < NC; ++k) {
            const u64* p = isum + k * npx;
            const long long S = (long long)(((ms_at(p, w, ey - 1, ex - 1) - ms_at(p, w, sy - 1, ex - 1)) -
                                             ms_at(p, w, ey - 1, sx - 1)) + ms_at(p, w, sy - 1, sx - 1));
            float sum;
            if (NC == 1) {
                // (int) of the window sum truncates toward zero, then the reference's unsigned sum
                const long long q = S >= 0 ? (S >> 40) : -((-S) >> 40);
                sum = (float)(unsigned)(int)q;
            } else {
                sum = (float)S * (1.0f / kFixScale);   // the window's Vec3d sum rounded to Vec3f
            }
            if (N > 0) lo[k] = sum / (float)N;
        }
    }
    for (int k = 0; k < 3; ++k) {
        low[3 * i + k] = lo[k];
        high[3 * i + k] = col[k] - lo[k];
    }
}

__global__ void k_add_color(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out,
                            size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

template <int NC>
void launch_split_nc(const float* color, const float* depth, const float* shape, int w, int h, u64* isum,
                     uint32_t* icnt, float* low, float* high, hipStream_t s) {
    const size_t npx = (size_t)w * h;
    const int rad = max(w, h) / 20;   // split_frequencies (blending.cpp:373)
    hipLaunchKernelGGL(k_ms_rows<NC>, dim3(h), dim3(256), 0, s, color, shape, depth, w, npx, isum, icnt);
    hipLaunchKernelGGL(k_ms_cols<NC>, dim3((w + kColW - 1) / kColW), dim3(kColW * kColG), 0, s, w, h, npx, isum,
                       icnt);
    hipLaunchKernelGGL(k_ms_split<NC>, dim3((w + 63) / 64, (h + 3) / 4), dim3(256), 0, s, color, shape, depth, w, h,
                       npx, rad, (const u64*)isum, (const uint32_t*)icnt, low, high);
}

}  // namespace

size_t split_ws_bytes(size_t npx, int rgb) {
    const int nc = rgb ? 3 : 1;
    return ((nc * npx * sizeof(u64) + 255) & ~(size_t)255) + ((npx * sizeof(uint32_t) + 255) & ~(size_t)255);
}

hipError_t launch_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h,
                                    int rgb, void* ws, float* low, float* high, hipStream_t s) {
    const size_t npx = (size_t)w * h;
    const int nc = rgb ? 3 : 1;
    u64* isum = (u64*)ws;
    uint32_t* icnt = (uint32_t*)((char*)ws + ((nc * npx * sizeof(u64) + 255) & ~(size_t)255));
    if (rgb) launch_split_nc<3>(color, depth, shape, w, h, isum, icnt, low, high, s);
    else launch_split_nc<1>(color, depth, shape, w, h, isum, icnt, low, high, s);
    return hipGetLastError();
}

hipError_t launch_add_color(const float* a, const float* b, float* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_add_color, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

}  // namespace rvs_k
