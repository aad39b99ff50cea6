// rvs_math.h -- float semantics shared by the host library and the gfx950 kernels.
//
// The reference's CPU path (an MSVC x64 build on OpenCV) fixes a few semantics
// that IEEE float code alone does not; they are restated here once and used by
// every kernel, so the kernels produce the reference's bits:
//   * std::min / std::max with NaN: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a.
//   * float -> int is cvttss2si: truncation, INT_MIN when out of range or NaN
//     (transform.cpp:90-93 feeds floor/ceil of projected coordinates through it).
//   * cv::norm accumulates squares in double (transform.cpp:63-65,
//     EquirectangularProjector.cpp:80).  norm_sq_sum_ff() returns the float
//     result of that double computation without any FP64 arithmetic:
//     float-float sums of the squares (fma splits every square exactly), one
//     float square root with an exact fma residual, and an exact integer path
//     for the rare inputs whose result lies too close to a float rounding
//     midpoint.  tri_stretch() does the same for valid_tri's maximum of three
//     edge norms with a single square root.  Host and device run this same
//     code (checked against double arithmetic on the CPU).
// All code here must be compiled with -ffp-contract=off.
#ifndef RVS_MATH_H
#define RVS_MATH_H

#include <stdint.h>
#include <limits.h>

#include <cmath>
#include <cstring>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define RVS_HD __host__ __device__ inline
#else
#define RVS_HD inline
#endif

namespace rvs_math {

RVS_HD float smin(float a, float b) { return (b < a) ? b : a; }
RVS_HD float smax(float a, float b) { return (a < b) ? b : a; }
RVS_HD int imin(int a, int b) { return (b < a) ? b : a; }
RVS_HD int imax(int a, int b) { return (a < b) ? b : a; }

RVS_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t u; std::memcpy(&u, &f, 4); return u;
#endif
}
RVS_HD float u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float f; std::memcpy(&f, &u, 4); return f;
#endif
}

RVS_HD bool is_finite(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }
RVS_HD bool is_nan(float x) { return (f2u(x) & 0x7fffffffu) > 0x7f800000u; }

// cvttss2si
RVS_HD int cvt_i32(float x) {
    if (x >= -2147483648.0f && x < 2147483648.0f) return (int)x;
    return INT_MIN;
}

RVS_HD float ffloor(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_floorf(x);
#else
    return std::floor(x);
#endif
}
RVS_HD float fceil(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ceilf(x);
#else
    return std::ceil(x);
#endif
}
RVS_HD float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrtf(x);   // correctly rounded (-fhip-fp32-correctly-rounded-divide-sqrt)
#else
    return std::sqrt(x);
#endif
}
RVS_HD float ffma(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}
RVS_HD float fabs_(float x) { return u2f(f2u(x) & 0x7fffffffu); }

// Square root and reciprocal for the float-float norms below.  The decision
// there is exact for any root within one ulp (it is corrected from its fma
// residual; a correction of a whole ulp or more is reported undecided), but
// the correctly rounded root keeps undecided cases rare: the device's bare
// v_sqrt_f32 left ~1e-2 of the triangles of a real scene undecided, so both
// sides use the correctly rounded root.  The reciprocal only scales a
// correction compared with a tolerance: v_rcp_f32 (one ulp) on the device.  A
// test build may substitute perturbed versions (RVS_MATH_TEST_ROOT /
// RVS_MATH_TEST_RCP) to check that any faithful root and reciprocal give the
// same results.
RVS_HD float ff_root(float x) {
#if defined(RVS_MATH_TEST_ROOT)
    return RVS_MATH_TEST_ROOT(x);
#else
    return fsqrt(x);
#endif
}
RVS_HD float ff_rcp(float x) {
#if defined(RVS_MATH_TEST_RCP)
    return RVS_MATH_TEST_RCP(x);
#elif defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// float -> uint8/uint16 as MSVC emits it (32-bit cvttss2si then truncation)
RVS_HD uint32_t quantize_u32(float v, unsigned maxv) { return (uint32_t)cvt_i32(v * (float)maxv); }

// ---------------------------------------------------------------------------
// Exact emulation of  (float) sqrt( ((0.0 + (double)a*a) + (double)b*b) + (double)c*c )
// with IEEE double round-to-nearest-even at every step.
// ---------------------------------------------------------------------------
namespace detail {

typedef unsigned __int128 u128;

RVS_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)x);
#else
    return __builtin_clzll(x);
#endif
}
RVS_HD int bitlen128(u128 x) {
    uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
    if (hi) return 128 - clz64(hi);
    if (lo) return 64 - clz64(lo);
    return 0;
}

// value = M * 2^E with M normalised to [2^63, 2^64), or M = 0
struct XF { uint64_t M; int E; };

RVS_HD XF xf_from_float_sq(float x) {
    uint32_t b = f2u(x) & 0x7fffffffu;
    uint32_t ex = b >> 23, mant = b & 0x7fffffu;
    uint64_t m; int e;
    if (ex == 0) { m = mant; e = -149; } else { m = mant | 0x800000u; e = (int)ex - 150; }
    XF r;
    if (m == 0) { r.M = 0; r.E = 0; return r; }
    uint64_t sq = m * m;                 // < 2^48, exact
    int s = clz64(sq);
    r.M = sq << s;
    r.E = 2 * e - s;
    return r;
}

// Round the 128-bit value S * 2^E (plus a sticky bit for nonzero bits below) to
// `prec` significant bits, round-to-nearest-even; returns normalised XF.
RVS_HD XF xf_round(u128 S, int E, bool sticky, int prec) {
    XF r;
    int L = bitlen128(S);
    if (L == 0) { r.M = 0; r.E = 0; return r; }
    if (L > prec) {
        int drop = L - prec;
        u128 kept = S >> drop;
        u128 rem = S - (kept << drop);
        u128 half = (u128)1 << (drop - 1);
        bool up = (rem > half) || (rem == half && (sticky || (kept & 1)));
        if (rem == half && sticky) up = true;
        if (up) kept += 1;
        S = kept;
        E += drop;
        L = bitlen128(S);
    }
    // normalise to bit 63
    r.M = (uint64_t)(S << (64 - L));
    r.E = E - (64 - L);
    return r;
}

// RN53(X + Y), X, Y >= 0
RVS_HD XF xf_add_rn53(XF X, XF Y) {
    if (X.M == 0) return Y;
    if (Y.M == 0) return X;
    if (X.E < Y.E) { XF t = X; X = Y; Y = t; }
    int shift = X.E - Y.E;
    u128 sx = (u128)X.M << 62;
    u128 sy = (u128)Y.M << 62;
    bool sticky = false;
    if (shift >= 126) {
        sticky = true; sy = 0;
    } else if (shift > 0) {
        u128 kept = sy >> shift;
        sticky = (kept << shift) != sy;
        sy = kept;
    }
    return xf_round(sx + sy, X.E - 62, sticky, 53);
}

// RN53(sqrt(X))
RVS_HD XF xf_sqrt_rn53(XF X) {
    if (X.M == 0) return X;
    int k = (X.E & 1) ? 49 : 48;          // X.E - k even
    u128 N = (u128)X.M << k;              // < 2^113
    int Eh = (X.E - k) / 2;
    // digit-by-digit integer square root
    u128 op = N, res = 0, one = (u128)1 << 112;
    while (one > op) one >>= 2;
    while (one != 0) {
        if (op >= res + one) { op -= res + one; res = (res >> 1) + one; }
        else res >>= 1;
        one >>= 2;
    }
    return xf_round(res, Eh, op != 0, 53);
}

// RN24(X) as a float (with subnormal and overflow handling)
RVS_HD float xf_to_float(XF X) {
    if (X.M == 0) return 0.0f;
    // value = M * 2^E, M in [2^63, 2^64): exponent of leading bit = E + 63
    int lead = X.E + 63;
    int prec = 24;
    if (lead < -126) prec = 24 - (-126 - lead);   // subnormal precision
    if (prec <= 0) {
        // below half the smallest subnormal (or exactly, handled by RNE on 1 bit)
        if (prec == 0) {
            // value in [2^-150, 2^-149): rounds to 2^-149 unless exactly 2^-150
            bool exact_half = (X.M == 0x8000000000000000ull);
            return exact_half ? 0.0f : u2f(1u);
        }
        return 0.0f;
    }
    XF r = xf_round((u128)X.M, X.E, false, prec);
    int rlead = r.E + 63;
    if (rlead > 127) return u2f(0x7f800000u);
    // assemble float bits
    uint32_t bits;
    if (rlead >= -126) {
        uint32_t mant = (uint32_t)(r.M >> 40) & 0x7fffffu;   // 24 bits incl. hidden
        bits = ((uint32_t)(rlead + 127) << 23) | mant;
    } else {
        // subnormal: value = m * 2^-149
        int sh = -149 - r.E;                  // r.M >> sh gives m
        bits = (sh >= 64) ? 0u : (uint32_t)(r.M >> sh);
    }
    return u2f(bits);
}

// Rare path (inputs near a float rounding midpoint, or outside the
// float-float range); kept out of line on the device so that its 128-bit
// arithmetic does not weigh on the callers' registers.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline)) float norm_slow(float a, float b, float c);
__device__ __attribute__((noinline)) float norm_slow(float a, float b, float c) {
#else
inline float norm_slow(float a, float b, float c) {
#endif
    XF A = xf_from_float_sq(a), B = xf_from_float_sq(b), Cc = xf_from_float_sq(c);
    XF s = xf_add_rn53(xf_add_rn53(A, B), Cc);
    return xf_to_float(xf_sqrt_rn53(s));
}

// error-free transforms
RVS_HD void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

}  // namespace detail

namespace detail {

// Components for which the float-float path is exact: zero or in [2^-40, 2^40)
// (the squares' fma error terms stay representable, nothing overflows).
RVS_HD bool ff_range(float ax) {
    const float lo_lim = 9.094947e-13f;   // 2^-40
    const float hi_lim = 1.0995116e12f;   // 2^40
    return ax < hi_lim && (ax == 0.0f || ax >= lo_lim);
}

// shi + slo = a^2 + b^2 + c^2 (|a|, |b|, |c| in ff_range) within 2^-46
// relative: every square split exactly by fma, two error-free additions.
RVS_HD void sq_sum3_ff(float aa, float ab, float ac, float& shi, float& slo) {
    const float pa = aa * aa, ea = ffma(aa, aa, -pa);
    const float pb = ab * ab, eb = ffma(ab, ab, -pb);
    const float pc = ac * ac, ec = ffma(ac, ac, -pc);
    float s1, t1;
    two_sum(pa, pb, s1, t1);
    float lo = t1 + (ea + eb);
    float s2, t2;
    two_sum(s1, pc, s2, t2);
    lo = lo + (t2 + ec);
    shi = s2 + lo;
    slo = lo - (shi - s2);
}

// The same for two components (what sq_sum3_ff gives with c = 0: its c terms
// add exact zeros).
RVS_HD void sq_sum2_ff(float aa, float ab, float& shi, float& slo) {
    const float pa = aa * aa, ea = ffma(aa, aa, -pa);
    const float pb = ab * ab, eb = ffma(ab, ab, -pb);
    float s1, t1;
    two_sum(pa, pb, s1, t1);
    const float lo = t1 + (ea + eb);
    shi = s1 + lo;
    slo = lo - (shi - s1);
}

// RN24(sqrt(shi + slo)) for shi > 0 (normal), decided when the root lies more
// than 2^-44 (relative) from a float rounding midpoint -- far beyond the error
// of shi + slo and of the reference's two double roundings (sum and root), so
// the reference's (float) sqrt of the double sum is that same float.  r is any
// root within one ulp; rlo ~ (sqrt - r) from the exact fma residual.  Returns
// false (undecided) otherwise, or when the correction reaches a whole ulp.
RVS_HD bool sqrt_ff_rn24(float shi, float slo, float& out) {
    const float r = ff_root(shi);
    const float resid = ffma(-r, r, shi);       // exact
    const float rlo = (resid + slo) * (0.5f * ff_rcp(r));
    const float tol = r * 5.684342e-14f;        // 2^-44 relative
    const uint32_t rb = f2u(r);
    if (rlo > 0.0f) {
        const float up = u2f(rb + 1u);
        const float ulp = up - r;
        const float d = rlo - ulp * 0.5f;
        if (rlo >= ulp) return false;
        if (d > tol) { out = up; return true; }
        if (d < -tol) { out = r; return true; }
    } else if (rlo < 0.0f) {
        const float dn = u2f(rb - 1u);
        const float ulp = r - dn;
        const float d = -rlo - ulp * 0.5f;
        if (-rlo >= ulp) return false;
        if (d > tol) { out = dn; return true; }
        if (d < -tol) { out = r; return true; }
    } else {
        out = r;
        return true;
    }
    return false;
}

}  // namespace detail

// (float)sqrt(((0.0 + (double)a*a) + (double)b*b) + (double)c*c) without the
// exact integer path: false (undecided, out = the float root of the float-float
// sum) for components outside the float-float range or a root within 2^-44 of
// a float rounding midpoint.
RVS_HD bool norm_sq_sum_fast(float a, float b, float c, float& out) {
    using namespace detail;
    const float aa = fabs_(a), ab = fabs_(b), ac = fabs_(c);
    if (is_nan(a) || is_nan(b) || is_nan(c)) { out = a + b + c; return true; }  // NaN
    if (!is_finite(aa) || !is_finite(ab) || !is_finite(ac)) { out = u2f(0x7f800000u); return true; }
    const float mx = smax(aa, smax(ab, ac));
    if (mx == 0.0f) { out = 0.0f; return true; }
    if (!(ff_range(aa) && ff_range(ab) && ff_range(ac))) {
        out = fsqrt(aa * aa + ab * ab + ac * ac);
        return false;
    }
    float shi, slo;
    sq_sum3_ff(aa, ab, ac, shi, slo);
    if (sqrt_ff_rn24(shi, slo, out)) return true;
    out = fsqrt(shi);
    return false;
}

// (float)sqrt(((0.0 + (double)a*a) + (double)b*b) + (double)c*c), exactly,
// in float arithmetic.
RVS_HD float norm_sq_sum_ff(float a, float b, float c) {
    float r;
    if (norm_sq_sum_fast(a, b, c, r)) return r;
    return detail::norm_slow(a, b, c);
}

// The float result of the reference's double norm of a 3-vector (the radius
// of an equirectangular projection, EquirectangularProjector.cpp:80).
RVS_HD float norm_sq_sum_d(float a, float b, float c) { return norm_sq_sum_ff(a, b, c); }

// valid_tri's stretch before the rescale (transform.cpp:62-68):
// (float) std::max(dbc, std::max(dab, dac)) of the double norms |A - B|,
// |A - C|, |B - C| (differences taken in float).  Fast form: the three squared
// lengths as float-float sums, the largest chosen with the reference's
// comparisons (so a NaN length is passed over or kept exactly as the double
// std::max does), one square root.  Two lengths whose float-float values
// compare in the wrong order differ by less than 2^-45 relative, so their
// floats differ only across a rounding midpoint within that distance, which
// the root's 2^-44 test reports as undecided.  Returns false (undecided) for
// non-finite or out-of-range components and near midpoints.
RVS_HD bool tri_stretch_fast(float A0, float A1, float B0, float B1, float C0, float C1, float& out) {
    using namespace detail;
    const float x0 = fabs_(A0 - B0), y0 = fabs_(A1 - B1);
    const float x1 = fabs_(A0 - C0), y1 = fabs_(A1 - C1);
    const float x2 = fabs_(B0 - C0), y2 = fabs_(B1 - C1);
    if (!(ff_range(x0) && ff_range(y0) && ff_range(x1) && ff_range(y1) && ff_range(x2) && ff_range(y2)))
        return false;   // also NaN (every comparison false)
    float h0, l0, h1, l1, h2, l2;
    sq_sum2_ff(x0, y0, h0, l0);
    sq_sum2_ff(x1, y1, h1, l1);
    sq_sum2_ff(x2, y2, h2, l2);
    // m = max(dab, dac) = (dab < dac) ? dac : dab; then (dbc < m) ? m : dbc
    float mh = h0, ml = l0;
    if (h0 < h1 || (h0 == h1 && l0 < l1)) { mh = h1; ml = l1; }
    if (!(h2 < mh || (h2 == mh && l2 < ml))) { mh = h2; ml = l2; }
    if (mh == 0.0f) { out = 0.0f; return true; }
    return sqrt_ff_rn24(mh, ml, out);
}

namespace detail {
// The three exact norms (max of the floats = float of the max of the doubles:
// the float conversion is monotone and keeps NaN, so the comparisons pick the
// same value).  Out of line on the device, like norm_slow.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline)) float tri_stretch_slow(float A0, float A1, float B0, float B1, float C0,
                                                            float C1);
__device__ __attribute__((noinline)) float tri_stretch_slow(float A0, float A1, float B0, float B1, float C0,
                                                            float C1) {
#else
inline float tri_stretch_slow(float A0, float A1, float B0, float B1, float C0, float C1) {
#endif
    const float dab = norm_sq_sum_ff(A0 - B0, A1 - B1, 0.0f);
    const float dac = norm_sq_sum_ff(A0 - C0, A1 - C1, 0.0f);
    const float dbc = norm_sq_sum_ff(B0 - C0, B1 - C1, 0.0f);
    return smax(dbc, smax(dab, dac));
}
}  // namespace detail

// The same, always decided: undecided cases take the three exact norms.
RVS_HD float tri_stretch(float A0, float A1, float B0, float B1, float C0, float C1) {
    float s;
    if (tri_stretch_fast(A0, A1, B0, B1, C0, C1, s)) return s;
    return detail::tri_stretch_slow(A0, A1, B0, B1, C0, C1);
}

// valid_tri's quality from the stretch (transform.cpp:68-74).  The clamp's
// max(1, q) maps NaN to 1, so the result is never 0.
RVS_HD float tri_quality(float stretch, float rescale) {
    if (rescale != 1.0f) stretch /= rescale;   // x / 1 = x for every x: the division is skipped, not changed
    float quality = 10000.f - 1000.f * stretch;
    quality = smax(1.f, quality);
    quality = smin(10000.f, quality);
    return quality;
}

}  // namespace rvs_math

#endif
