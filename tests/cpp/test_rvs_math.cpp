// CPU check of rvs_math.h: the exact emulation of OpenCV's double-accumulated
// norm must equal the double computation bit for bit, on random inputs, on
// inputs steered toward float rounding midpoints, and on special values.
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
    printf("checked %ld random + %ld steered: fast mismatches %ld, slow mismatches %ld\n", 4 * n, 2 * near, bad, slow_bad);
    return (bad || slow_bad) ? 1 : 0;
}
