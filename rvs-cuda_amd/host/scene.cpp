// scene.cpp -- seeded synthetic scenes rendered to raw YUV420 colour and
// inverse-depth planes, in the reference's camera model (OMAF axes: x forward,
// y left, z up; Rotation = yaw, pitch, roll composed as Rz*Ry*Rx;
// perspective and equirectangular projections as rvs/src/*Unprojector.cpp).
//
// Scenes (the workloads the view synthesizer is measured on):
//   kind 0 "rects":   6 camera-facing textured rectangles at depths U(1, 10) m in
//                     front of a textured background plane at 10 m.
//   kind 1 "spheres": 24 textured spheres, radius U(0.2, 1.5) m, centres at
//                     distance U(1.5, 20) m in uniform directions, inside a
//                     textured axis-aligned box of half-extent 14 m.
// Texture: 3-octave value noise (lattice 32/16/8 px at the reference scale,
// weights 1/2, 1/4, 1/8, smoothstep interpolation, integer hash of the seed).
// Depth is stored as raw = round(maxv * (1/z - 1/far) / (1/near - 1/far)),
// the inverse-depth quantisation decoded by process_input_frame.cu:358-381.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "rvs_host.h"

namespace {

struct V3 { double x, y, z; };
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x8da6b343u ^ b * 0xd8163841u ^ c * 0xcb1ab31fu;
    h ^= h >> 15; h *= 0x2c1b3c6du; h ^= h >> 12; h *= 0x297a2d39u; h ^= h >> 15;
    return h;
}

double lattice(int ix, int iy, uint32_t seed) {
    return (hash3((uint32_t)ix, (uint32_t)iy, seed) & 0xffffff) / double(0xffffff);
}

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double vnoise1(double x, double y, uint32_t seed) {
    const double fx = std::floor(x), fy = std::floor(y);
    const int ix = (int)fx, iy = (int)fy;
    const double tx = smooth(x - fx), ty = smooth(y - fy);
    const double a = lattice(ix, iy, seed), b = lattice(ix + 1, iy, seed);
    const double c = lattice(ix, iy + 1, seed), d = lattice(ix + 1, iy + 1, seed);
    return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

// 3 octaves, lattice 32/16/8 "pixels", normalised to [0, 1]
double vnoise(double x, double y, uint32_t seed) {
    return (0.5 * vnoise1(x / 32.0, y / 32.0, seed) + 0.25 * vnoise1(x / 16.0, y / 16.0, seed + 101u) +
            0.125 * vnoise1(x / 8.0, y / 8.0, seed + 202u)) / 0.875;
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E5ull) {}
    double uni() {   // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (z >> 11) * (1.0 / 9007199254740992.0);
    }
    double uni(double a, double b) { return a + (b - a) * uni(); }
};

struct Sphere { V3 c; double r; uint32_t tex; };
struct Rect { double x; double y0, y1, z0, z1; uint32_t tex; };   // plane x = const, facing -x
struct Plane { int axis; double pos; uint32_t tex; };             // axis-aligned plane

struct Scene {
    int kind;
    std::vector<Sphere> spheres;
    std::vector<Rect> rects;
    std::vector<Plane> planes;
};

Scene build(int kind, uint64_t seed) {
    Scene S;
    S.kind = kind;
    Rng rng(seed);
    if (kind == 0) {
        // the reference rig looks along +x with f = 500 on 640x480: |y| < 0.64 x, |z| < 0.48 x
        for (int k = 0; k < 6; ++k) {
            const double x = rng.uni(1.0, 10.0);
            const double cy = rng.uni(-0.5, 0.5) * x, cz = rng.uni(-0.4, 0.4) * x;
            const double hy = rng.uni(0.08, 0.3) * x, hz = rng.uni(0.08, 0.3) * x;
            S.rects.push_back({x, cy - hy, cy + hy, cz - hz, cz + hz, (uint32_t)(1000 + k)});
        }
        S.planes.push_back({0, 10.0, 999u});
    } else {
        for (int k = 0; k < 24; ++k) {
            const double u = rng.uni(-1.0, 1.0), a = rng.uni(0.0, 2.0 * M_PI);
            const double s = std::sqrt(1.0 - u * u);
            const double d = rng.uni(1.5, 20.0);
            const double r = rng.uni(0.2, 1.5);
            S.spheres.push_back({{d * s * std::cos(a), d * s * std::sin(a), d * u}, r, (uint32_t)(2000 + k)});
        }
        const double H = 14.0;
        for (int ax = 0; ax < 3; ++ax) {
            S.planes.push_back({ax, -H, (uint32_t)(3000 + 2 * ax)});
            S.planes.push_back({ax, H, (uint32_t)(3001 + 2 * ax)});
        }
    }
    return S;
}

// nearest hit: distance t along dir (dir need not be unit), texture id, texture coords
bool trace(const Scene& S, V3 o, V3 dir, double& t_best, uint32_t& tex, double& tu, double& tv) {
    t_best = 1e30;
    bool hit = false;
    const double dd = dot(dir, dir);
    for (const Sphere& sp : S.spheres) {
        const V3 oc = o - sp.c;
        const double b = dot(oc, dir), c = dot(oc, oc) - sp.r * sp.r;
        const double disc = b * b - dd * c;
        if (disc < 0) continue;
        const double sq = std::sqrt(disc);
        double t = (-b - sq) / dd;
        if (t <= 1e-6) t = (-b + sq) / dd;
        if (t > 1e-6 && t < t_best) {
            t_best = t; hit = true; tex = sp.tex;
            const V3 p = (o + t * dir) - sp.c;
            const double th = std::atan2(p.z, std::sqrt(p.x * p.x + p.y * p.y));
            const double ph = std::atan2(p.y, p.x);
            tu = ph * sp.r * 120.0; tv = th * sp.r * 120.0;   // ~120 px per metre of arc
        }
    }
    for (const Rect& R : S.rects) {
        if (std::fabs(dir.x) < 1e-12) continue;
        const double t = (R.x - o.x) / dir.x;
        if (t <= 1e-6 || t >= t_best) continue;
        const double y = o.y + t * dir.y, z = o.z + t * dir.z;
        if (y < R.y0 || y > R.y1 || z < R.z0 || z > R.z1) continue;
        t_best = t; hit = true; tex = R.tex;
        tu = y / R.x * 500.0; tv = z / R.x * 500.0;           // texture at the rig's pixel scale
    }
    for (const Plane& P : S.planes) {
        const double dc = P.axis == 0 ? dir.x : (P.axis == 1 ? dir.y : dir.z);
        const double oc = P.axis == 0 ? o.x : (P.axis == 1 ? o.y : o.z);
        if (std::fabs(dc) < 1e-12) continue;
        const double t = (P.pos - oc) / dc;
        if (t <= 1e-6 || t >= t_best) continue;
        const V3 p = o + t * dir;
        t_best = t; hit = true; tex = P.tex;
        const double scale = S.kind == 0 ? 50.0 : 40.0;
        if (P.axis == 0) { tu = p.y * scale; tv = p.z * scale; }
        else if (P.axis == 1) { tu = p.x * scale; tv = p.z * scale; }
        else { tu = p.x * scale; tv = p.y * scale; }
    }
    return hit;
}

void rotation(const float rot[3], double R[9]) {
    const double d2r = M_PI / 180.0;
    const double cz = std::cos(rot[0] * d2r), sz = std::sin(rot[0] * d2r);
    const double cy = std::cos(rot[1] * d2r), sy = std::sin(rot[1] * d2r);
    const double cx = std::cos(rot[2] * d2r), sx = std::sin(rot[2] * d2r);
    const double Rz[9] = {cz, -sz, 0, sz, cz, 0, 0, 0, 1};
    const double Ry[9] = {cy, 0, sy, 0, 1, 0, -sy, 0, cy};
    const double Rx[9] = {1, 0, 0, 0, cx, -sx, 0, sx, cx};
    double T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i * 3 + j] = 0;
            for (int k = 0; k < 3; ++k) T[i * 3 + j] += Rz[i * 3 + k] * Ry[k * 3 + j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i * 3 + j] = 0;
            for (int k = 0; k < 3; ++k) R[i * 3 + j] += T[i * 3 + k] * Rx[k * 3 + j];
        }
}

template <typename T>
void put(void* p, size_t i, uint32_t v) { ((T*)p)[i] = (T)v; }

}  // namespace

extern "C" int rvs_scene_render(int kind, uint64_t seed, const rvs_cam* cam, const float pos[3], const float rot[3],
                                int bitdepth_color, int bitdepth_depth, float znear, float zfar, void* y, void* u,
                                void* v, void* depth, int threads) {
    if (!cam || cam->w < 2 || cam->h < 2 || (cam->w & 1) || (cam->h & 1) || !y || !u || !v || !depth) return -1;
    const Scene S = build(kind, seed);
    double R[9];
    rotation(rot, R);
    const V3 o = {pos[0], pos[1], pos[2]};
    const int W = cam->w, H = cam->h;
    const uint32_t cmax = (1u << bitdepth_color) - 1u, dmax = (1u << bitdepth_depth) - 1u;
    const double cscale = cmax / 255.0;
    std::vector<float> uf((size_t)W * H), vf((size_t)W * H);
    auto row_job = [&](int r0, int r1) {
        for (int i = r0; i < r1; ++i)
            for (int j = 0; j < W; ++j) {
                const double uu = j + 0.5, vv = i + 0.5;
                V3 dc;
                if (cam->proj == 1) {
                    const double d2r = M_PI / 180.0;
                    const double phi = cam->hor[1] * d2r - (cam->hor[1] - cam->hor[0]) * d2r * uu / W;
                    const double th = cam->ver[1] * d2r - (cam->ver[1] - cam->ver[0]) * d2r * vv / H;
                    dc = {std::cos(th) * std::cos(phi), std::cos(th) * std::sin(phi), std::sin(th)};
                } else {
                    dc = {1.0, -(uu - cam->p[0]) / cam->f[0], -(vv - cam->p[1]) / cam->f[1]};
                }
                const V3 dw = {R[0] * dc.x + R[1] * dc.y + R[2] * dc.z, R[3] * dc.x + R[4] * dc.y + R[5] * dc.z,
                               R[6] * dc.x + R[7] * dc.y + R[8] * dc.z};
                double t, tu = 0, tv = 0;
                uint32_t tex = 0;
                double n0 = 0.5, n1 = 0.5, n2 = 0.5;
                if (!trace(S, o, dw, t, tex, tu, tv)) t = zfar;
                else {
                    n0 = vnoise(tu, tv, tex * 3u);
                    n1 = vnoise(tu + 17.0, tv - 5.0, tex * 3u + 1u);
                    n2 = vnoise(tu - 9.0, tv + 23.0, tex * 3u + 2u);
                }
                const size_t k = (size_t)i * W + j;
                const double Y = (16.0 + 219.0 * n0) * cscale;
                if (bitdepth_color <= 8) put<uint8_t>(y, k, (uint32_t)std::lround(Y));
                else put<uint16_t>(y, k, (uint32_t)std::lround(Y));
                uf[k] = (float)((128.0 + 80.0 * (n1 - 0.5)) * cscale);
                vf[k] = (float)((128.0 + 80.0 * (n2 - 0.5)) * cscale);
                const double z = std::min(std::max(t, (double)znear), (double)zfar);
                double q = dmax * (1.0 / z - 1.0 / zfar) / (1.0 / znear - 1.0 / zfar);
                q = std::min(std::max(q, 0.0), (double)dmax);
                if (bitdepth_depth <= 8) put<uint8_t>(depth, k, (uint32_t)std::lround(q));
                else put<uint16_t>(depth, k, (uint32_t)std::lround(q));
            }
    };
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, 64);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(row_job, H * t / nt, H * (t + 1) / nt);
    for (auto& th : pool) th.join();
    // chroma: 2x2 mean of the full-resolution chroma
    const int hw = W / 2, hh = H / 2;
    for (int i = 0; i < hh; ++i)
        for (int j = 0; j < hw; ++j) {
            const size_t a = (size_t)(2 * i) * W + 2 * j, b = a + W;
            const uint32_t cu = (uint32_t)std::lround((uf[a] + uf[a + 1] + uf[b] + uf[b + 1]) * 0.25);
            const uint32_t cv = (uint32_t)std::lround((vf[a] + vf[a + 1] + vf[b] + vf[b + 1]) * 0.25);
            const size_t o2 = (size_t)i * hw + j;
            if (bitdepth_color <= 8) { put<uint8_t>(u, o2, cu); put<uint8_t>(v, o2, cv); }
            else { put<uint16_t>(u, o2, cu); put<uint16_t>(v, o2, cv); }
        }
    return 0;
}
