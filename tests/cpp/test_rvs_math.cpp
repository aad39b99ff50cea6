// CPU check of rvs_math.h: the exact emulation of OpenCV's double-accumulated
// norm must equal the double computation bit for bit, on random inputs, on
// inputs steered toward float rounding midpoints, and on special values.
// With RVS_FAITHFUL defined, the float-float norms run with a square root and
// a reciprocal rounded to either float neighbour of the exact value (within one
// ulp, what the device's v_sqrt_f32 / v_rcp_f32 guarantee): the results must
// stay exact.
#ifdef RVS_FAITHFUL
#include <random>
static std::mt19937_64 g_perturb(99);
// either float neighbour of the exact value v (a faithful rounding), at random
static float faithful(double v) {
    const float r = (float)v;
    if ((double)r == v || (g_perturb() & 1)) return r;
    return (double)r > v ? std::nextafter(r, 0.f) : std::nextafter(r, 1e30f);
}
static float faithful_sqrt(float x) { return faithful(std::sqrt((double)x)); }
static float faithful_rcp(float x) { return faithful(1.0 / (double)x); }
#define RVS_MATH_TEST_ROOT faithful_sqrt
#define RVS_MATH_TEST_RCP faithful_rcp
#endif
#include "../../rvs-cuda_amd/csrc/rvs_math.h"

#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>

static float ref_norm(float a, float b, float c) {
    double s = 0.0;
    double v;
    v = a; s += v * v;
    v = b; s += v * v;
    v = c; s += v * v;
    return (float)std::sqrt(s);
}

static bool same(float x, float y) {
    if (std::isnan(x) && std::isnan(y)) return true;
    uint32_t a, b; std::memcpy(&a, &x, 4); std::memcpy(&b, &y, 4);
    return a == b;
}

// valid_tri's stretch as the reference computes it (transform.cpp:62-68):
// cv::norm of Vec2f differences in double, std::max of the doubles, then float.
static double ref_norm2(float a0, float a1, float b0, float b1) {
    double s = 0.0;
    double v = (double)(a0 - b0);
    s += v * v;
    v = (double)(a1 - b1);
    s += v * v;
    return std::sqrt(s);
}

static float ref_stretch(float A0, float A1, float B0, float B1, float C0, float C1) {
    const double dab = ref_norm2(A0, A1, B0, B1), dac = ref_norm2(A0, A1, C0, C1), dbc = ref_norm2(B0, B1, C0, C1);
    const double m = (dab < dac) ? dac : dab;
    return (float)((dbc < m) ? m : dbc);
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 2000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<float> U(-3000.f, 3000.f);
    std::uniform_real_distribution<float> S(-2.f, 2.f);
    std::uniform_int_distribution<int> E(-60, 60);
    long bad = 0, slow_bad = 0;
    auto check = [&](float a, float b, float c) {
        float r = ref_norm(a, b, c);
        float f = rvs_math::norm_sq_sum_d(a, b, c);
        float g = rvs_math::detail::norm_slow(a, b, c);
        if (!same(r, f)) { if (bad < 5) printf("fast mismatch %a %a %a: ref %a got %a\n", a, b, c, r, f); ++bad; }
        if (!std::isnan(r) && !std::isinf(r) && !same(r, g)) { if (slow_bad < 5) printf("slow mismatch %a %a %a: ref %a got %a\n", a, b, c, r, g); ++slow_bad; }
    };
    for (long i = 0; i < n; ++i) {
        check(U(rng), U(rng), (i & 1) ? 0.f : U(rng));
        check(S(rng) * std::ldexp(1.f, E(rng)), S(rng) * std::ldexp(1.f, E(rng)), S(rng) * std::ldexp(1.f, E(rng)));
        // integer pairs: sqrt of integers lands near float midpoints far more often
        int ia = (int)(rng() % 20000000), ib = (int)(rng() % 20000000);
        check((float)ia, (float)ib, 0.f);
        // tiny differences like neighbouring projected coordinates
        float base = U(rng);
        float nb = std::nextafter(base, 1e9f);
        check(base - nb, (float)(i % 7) * 0.25f, 0.f);
    }
    // steer toward midpoints: choose m = f + ulp/2 and a = m rounded, b small so that
    // sqrt(a^2 + b^2) ~ a + b^2/(2a) lands exactly around the midpoint
    long near = 0;
    for (long i = 0; i < n; ++i) {
        float a = (float)(1 + rng() % 16777215);
        float ulp = std::nextafter(a, 1e30f) - a;
        // want b^2 / (2a) ~ ulp/2  ->  b ~ sqrt(a * ulp)
        float b = std::sqrt(a * ulp) * (1.0f + (S(rng) * 1e-6f));
        check(a, b, 0.f);
        check(a * 0.75f, b, a * 0.5f);
        ++near;
    }
    const float specials[] = {0.f, -0.f, INFINITY, -INFINITY, NAN, 1e-45f, 1e-40f, 3e38f, 1e-30f, 1.f};
    for (float a : specials) for (float b : specials) for (float c : specials) check(a, b, c);
    // triangle stretch: the fast path (float-float lengths, one root) where it
    // decides, the three exact norms otherwise, against double arithmetic
    long sbad = 0, undecided = 0, tris = 0;
    auto tri = [&](float A0, float A1, float B0, float B1, float C0, float C1) {
        ++tris;
        const float r = ref_stretch(A0, A1, B0, B1, C0, C1);
        float f;
        if (rvs_math::tri_stretch_fast(A0, A1, B0, B1, C0, C1, f)) {
            if (!same(r, f)) { if (sbad < 5) printf("stretch fast mismatch %a %a %a %a %a %a: ref %a got %a\n", A0, A1, B0, B1, C0, C1, r, f); ++sbad; }
        } else {
            ++undecided;
        }
        const float g = rvs_math::tri_stretch(A0, A1, B0, B1, C0, C1);
        if (!same(r, g)) { if (sbad < 5) printf("stretch mismatch %a %a %a %a %a %a: ref %a got %a\n", A0, A1, B0, B1, C0, C1, r, g); ++sbad; }
        const float qr = rvs_math::tri_quality(r, 1.0f), qg = rvs_math::tri_quality(g, 1.0f);
        if (!same(qr, qg)) ++sbad;
    };
    std::uniform_real_distribution<float> P(0.f, 2000.f);
    std::uniform_real_distribution<float> J(-3.f, 3.f);
    for (long i = 0; i < n; ++i) {
        // neighbouring projected vertices: small triangles at image coordinates
        const float x = P(rng), y = P(rng);
        tri(x, y, x + J(rng), y + J(rng), x + J(rng), y + J(rng));
        // grid-aligned corners (exact halves and integers: equal edges, ties in the max)
        const float gx = (float)(rng() % 4096) + 0.5f, gy = (float)(rng() % 2048) + 0.5f;
        tri(gx, gy, gx + 1.f, gy, gx, gy + 1.f);
        tri(gx + (float)(rng() % 5), gy, gx, gy + (float)(rng() % 5), gx + (float)(rng() % 3), gy + (float)(rng() % 3));
        // wide range of scales and equal-length edges in different orders
        const float e1 = S(rng) * std::ldexp(1.f, E(rng) / 3), e2 = S(rng) * std::ldexp(1.f, E(rng) / 3);
        tri(0.f, 0.f, e1, e2, e2, e1);
        tri(e1, e2, 0.f, 0.f, -e2, e1);
    }
    // integer legs whose hypotenuse lands near a float midpoint, as an edge of a triangle
    for (long i = 0; i < n / 4; ++i) {
        const float a = (float)(1 + rng() % 16777215);
        const float ulp = std::nextafter(a, 1e30f) - a;
        const float b = std::sqrt(a * ulp) * (1.0f + (S(rng) * 1e-6f));
        tri(0.f, 0.f, a, b, a * 0.5f, b * 0.25f);
        tri(a * 0.5f, b * 0.25f, 0.f, 0.f, a, b);
    }
    // non-finite and extreme vertices (NaN lengths are passed over or kept as std::max does)
    const float sv[] = {0.f, 1.f, -2.5f, 1e-30f, 1e30f, INFINITY, -INFINITY, NAN, 3e38f};
    for (float a : sv) for (float b : sv) for (float c : sv) for (float d : sv) {
        tri(a, 1.f, b, 2.f, c, d);
        tri(1.f, a, 2.f, b, d, c);
    }
    printf("checked %ld random + %ld steered: fast mismatches %ld, slow mismatches %ld\n", 4 * n, 2 * near, bad, slow_bad);
    printf("stretch: %ld triangles, %ld mismatches, %ld undecided by the fast path\n", tris, sbad, undecided);
    return (bad || slow_bad || sbad) ? 1 : 0;
}
