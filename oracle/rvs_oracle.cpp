/*
 * rvs_oracle.cpp -- CPU restatement of the reference CPU synthesis path.
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path); see rvs_oracle.h.
 *
 * Build: -O2 -ffp-contract=off, no fast-math (oracle/Makefile).  Every float
 * expression keeps the reference's operand order; no operation is fused unless
 * the reference's own compiler fuses it (the decode kernels, see below).
 *
 * OpenCV conventions restated here (OpenCV 4.9.0 is the reference's dependency,
 * README.md:83; it is not installed, so these are restated from its public
 * source and are "parity unpinned" against a binary):
 *   - Matx*Matx and Matx*Vec accumulate  s = 0; s += a(i,k)*b(k)  in k order.
 *   - Vec / float is Vec * (1.f/float)   (matx.hpp operator/, Matx_ScaleOp).
 *   - cv::norm(Vec3f) and cv::norm(Vec2f, Vec2f) accumulate squares in double:
 *     s = 0.0; s += (double)(a_i - b_i)^2 ... ; sqrt(s).
 *   - cv::transform with a diagonal 2x2 matrix uses the diagonal kernel:
 *     u' = sx*u + 0, v' = sy*v + 0.
 *   - resize(0.5, 0.5, INTER_LINEAR) on 8U/16U with exact scale 2 takes the
 *     area-fast path: (a + b + c + d + 2) >> 2.
 *   - std::min(a,b) = (b < a) ? b : a ; std::max(a,b) = (a < b) ? b : a.
 * Platform conventions (the reference is an MSVC x64 build):
 *   - float -> int conversion out of range or NaN yields INT_MIN (cvttss2si).
 *   - float -> uint8/uint16 goes through that 32-bit conversion, then truncates.
 * Decode kernels ran through nvcc (default -fmad=true) even on the CPU path, so
 * the contraction nvcc applies to  near + v*(far-near)  is restated as fmaf.
 */
#include "rvs_oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

const float kNaN = std::numeric_limits<float>::quiet_NaN();
const float kInf = std::numeric_limits<float>::infinity();

inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }
inline double smaxd(double a, double b) { return (a < b) ? b : a; }
inline int imin(int a, int b) { return (b < a) ? b : a; }
inline int imax(int a, int b) { return (a < b) ? b : a; }

/* x86 cvttss2si: truncation, INT_MIN when out of range or NaN. */
inline int cvt_i32(float x) {
    if (x >= -2147483648.0f && x < 2147483648.0f) return (int)x;
    return INT_MIN;
}

/* Matx33f * Matx33f (OpenCV Matx_MatMulOp) */
void matmul33(const float* A, const float* B, float* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.f;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
}

/* Matx33f * Vec3f */
void matvec33(const float* A, const float* x, float* y) {
    for (int i = 0; i < 3; ++i) {
        float s = 0.f;
        for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * x[k];
        y[i] = s;
    }
}

/* cv::norm(Vec2f A, Vec2f B) -> normDiffL2_32f, double accumulation */
inline double norm2d(float a0, float a1, float b0, float b1) {
    double s = 0.0;
    double v0 = (double)(a0 - b0);
    s += v0 * v0;
    double v1 = (double)(a1 - b1);
    s += v1 * v1;
    return std::sqrt(s);
}

/* cv::norm(Vec3f) -> normL2Sqr<float,double> */
inline float norm3f(float x, float y, float z) {
    double s = 0.0;
    double v;
    v = x; s += v * v;
    v = y; s += v * v;
    v = z; s += v * v;
    return (float)std::sqrt(s);
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* Definitions below keep the C linkage declared in rvs_oracle.h. */

unsigned orc_max_level(int bitdepth) { return (1u << bitdepth) - 1u; }

/* ------------------------------------------------------------------ geometry */

void orc_euler_to_R(const float rot_deg[3], float R[9]) {
    const float radperdeg = 0.01745329252f;
    float rz = radperdeg * rot_deg[0];
    float ry = radperdeg * rot_deg[1];
    float rx = radperdeg * rot_deg[2];
    const float Rz[9] = {std::cos(rz), -std::sin(rz), 0.f,
                         std::sin(rz), std::cos(rz), 0.f,
                         0.f, 0.f, 1.f};
    const float Ry[9] = {std::cos(ry), 0.f, std::sin(ry),
                         0.f, 1.f, 0.f,
                         -std::sin(ry), 0.f, std::cos(ry)};
    const float Rx[9] = {1.f, 0.f, 0.f,
                         0.f, std::cos(rx), -std::sin(rx),
                         0.f, std::sin(rx), std::cos(rx)};
    float T[9];
    matmul33(Rz, Ry, T);
    matmul33(T, Rx, R);
}

void orc_relative_pose(const float pos_from[3], const float rot_from[3],
                       const float pos_to[3], const float rot_to[3],
                       float R[9], float t[3]) {
    float Rf[9], Rt[9], RtT[9], nRtT[9];
    orc_euler_to_R(rot_from, Rf);
    orc_euler_to_R(rot_to, Rt);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) RtT[i * 3 + j] = Rt[j * 3 + i];
    matmul33(RtT, Rf, R);
    for (int i = 0; i < 9; ++i) nRtT[i] = RtT[i] * -1.f;
    float d[3] = {pos_to[0] - pos_from[0], pos_to[1] - pos_from[1], pos_to[2] - pos_from[2]};
    matvec33(nRtT, d, t);
}

void orc_generate_image_pos(const orc_cam* cam, float* uv) {
    for (int i = 0; i != cam->h; ++i)
        for (int j = 0; j != cam->w; ++j) {
            uv[((size_t)i * cam->w + j) * 2 + 0] = j + 0.5f;
            uv[((size_t)i * cam->w + j) * 2 + 1] = i + 0.5f;
        }
    if (cam->proj == 1) {
        const float eps = 1e-3f;
        for (int j = 0; j != cam->w; ++j) {
            uv[(size_t)j * 2 + 1] = eps;
            uv[((size_t)(cam->h - 1) * cam->w + j) * 2 + 1] = cam->h - eps;
        }
    }
}

void orc_unproject(const orc_cam* cam, const float* uv, const float* depth, size_t n, float* xyz) {
    if (cam->proj == 1) {
        const float radperdeg = 0.01745329252f;
        float phi0 = radperdeg * cam->hor[1];
        float theta0 = radperdeg * cam->ver[1];
        float dphi_du = -radperdeg * (cam->hor[1] - cam->hor[0]) / cam->w;
        float dtheta_dv = -radperdeg * (cam->ver[1] - cam->ver[0]) / cam->h;
        for (size_t k = 0; k < n; ++k) {
            float phi = phi0 + dphi_du * uv[2 * k];
            float theta = theta0 + dtheta_dv * uv[2 * k + 1];
            float a = std::cos(theta) * std::cos(phi);
            float b = std::cos(theta) * std::sin(phi);
            float c = std::sin(theta);
            float d = depth[k];
            xyz[3 * k + 0] = a * d;
            xyz[3 * k + 1] = b * d;
            xyz[3 * k + 2] = c * d;
        }
    } else {
        for (size_t k = 0; k < n; ++k) {
            float d = depth[k];
            if (d > 0.f) {
                xyz[3 * k + 0] = d;
                xyz[3 * k + 1] = -(d / cam->f[0]) * (uv[2 * k] - cam->p[0]);
                xyz[3 * k + 2] = -(d / cam->f[1]) * (uv[2 * k + 1] - cam->p[1]);
            } else {
                xyz[3 * k + 0] = xyz[3 * k + 1] = xyz[3 * k + 2] = kNaN;
            }
        }
    }
}

void orc_affine(const float R[9], const float t[3], const float* xyz_in, size_t n, float* xyz_out) {
    for (size_t k = 0; k < n; ++k) {
        float y[3];
        matvec33(R, xyz_in + 3 * k, y);
        xyz_out[3 * k + 0] = y[0] + t[0];
        xyz_out[3 * k + 1] = y[1] + t[1];
        xyz_out[3 * k + 2] = y[2] + t[2];
    }
}

int orc_project(const orc_cam* cam, const float* xyz, size_t n, float* uv, float* depth) {
    if (cam->proj == 1) {
        const float degperrad = 57.295779513f;
        float u0 = cam->w * cam->hor[1] / (cam->hor[1] - cam->hor[0]);
        float v0 = cam->h * cam->ver[1] / (cam->ver[1] - cam->ver[0]);
        float du_dphi = -degperrad * cam->w / (cam->hor[1] - cam->hor[0]);
        float dv_dtheta = -degperrad * cam->h / (cam->ver[1] - cam->ver[0]);
        for (size_t k = 0; k < n; ++k) {
            float x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
            float radius = norm3f(x, y, z);
            depth[k] = radius;
            float phi = std::atan2(y, x);
            float theta = std::asin(z / radius);
            uv[2 * k] = u0 + du_dphi * phi;
            uv[2 * k + 1] = v0 + dv_dtheta * theta;
        }
        return cam->full_hor ? 1 : 0;
    }
    for (size_t k = 0; k < n; ++k) {
        float x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
        if (x > 0.f) {
            uv[2 * k] = -cam->f[0] * y / x + cam->p[0];
            uv[2 * k + 1] = -cam->f[1] * z / x + cam->p[1];
            depth[k] = x;
        } else {
            uv[2 * k] = uv[2 * k + 1] = kNaN;
            depth[k] = kNaN;
        }
    }
    return 0;
}

void orc_scale_uv(const float* uv, size_t n, float sx, float sy, float* out) {
    for (size_t k = 0; k < n; ++k) {
        out[2 * k] = sx * uv[2 * k] + 0.f;
        out[2 * k + 1] = sy * uv[2 * k + 1] + 0.f;
    }
}

int orc_reproject(const orc_cam* src, const orc_cam* dst, const float R[9], const float t[3],
                  const float* src_depth, float rescale, float* uv, float* depth) {
    size_t n = (size_t)src->w * src->h;
    std::vector<float> pos(2 * n), xyz(3 * n), vxyz(3 * n), vuv(2 * n);
    orc_generate_image_pos(src, pos.data());
    orc_unproject(src, pos.data(), src_depth, n, xyz.data());
    orc_affine(R, t, xyz.data(), n, vxyz.data());
    int wrap = orc_project(dst, vxyz.data(), n, vuv.data(), depth);
    int ow = (int)(0.5f + dst->w * rescale);
    int oh = (int)(0.5f + dst->h * rescale);
    orc_scale_uv(vuv.data(), n, float(ow) / dst->w, float(oh) / dst->h, uv);
    return wrap;
}

/* -------------------------------------------------------------------- raster */

namespace {

struct RasterCtx {
    const float* color;   // in_w*in_h*3
    const uint8_t* mask;  // input_depth > 0
    const float* depth;
    const float* pos;     // uv
    int in_w;
    int out_w, out_h;
    float rescale;
    float* res;           // out color
    float* new_depth;
    uint8_t* new_prol;
    float* shape;
};

float valid_tri(float A0, float A1, float B0, float B1, float C0, float C1, float rescale) {
    double dab = norm2d(A0, A1, B0, B1);
    double dac = norm2d(A0, A1, C0, C1);
    double dbc = norm2d(B0, B1, C0, C1);
    float stretch = static_cast<float>(smaxd(dbc, smaxd(dab, dac)));
    stretch /= rescale;
    float quality = 10000.f - 1000.f * stretch;
    quality = smax(1.f, quality);
    quality = smin(10000.f, quality);
    return quality;
}

void colorize_triangle(RasterCtx& c, int ax, int ay, int bx, int by, int cx, int cy) {
    size_t ia = (size_t)ay * c.in_w + ax, ib = (size_t)by * c.in_w + bx, ic = (size_t)cy * c.in_w + cx;
    float A0 = c.pos[2 * ia], A1 = c.pos[2 * ia + 1];
    float B0 = c.pos[2 * ib], B1 = c.pos[2 * ib + 1];
    float C0 = c.pos[2 * ic], C1 = c.pos[2 * ic + 1];

    float den = (B1 - C1) * (A0 - C0) + (C0 - B0) * (A1 - C1);
    if (den <= 0.f) return;

    float tv = valid_tri(A0, A1, B0, B1, C0, C1, c.rescale);
    if (tv == 0.f) return;

    int Xmin = imax(0, cvt_i32(std::floor(smin(smin(A0, B0), C0))));
    int Ymin = imax(0, cvt_i32(std::floor(smin(smin(A1, B1), C1))));
    int Xmax = imin(c.out_w - 1, cvt_i32(std::ceil(smax(smax(A0, B0), C0))));
    int Ymax = imin(c.out_h - 1, cvt_i32(std::ceil(smax(smax(A1, B1), C1))));
    if (Ymin > Ymax || Xmin > Xmax) return;

    const float* colA = c.color + 3 * ia;
    const float* colB = c.color + 3 * ib;
    const float* colC = c.color + 3 * ic;
    float dA = c.depth[ia], dB = c.depth[ib], dC = c.depth[ic];
    bool prol = c.mask[ia] || c.mask[ib] || c.mask[ic];

    for (int y = Ymin; y <= Ymax; ++y)
        for (int x = Xmin; x <= Xmax; ++x) {
            const float offset = 0.5f;
            float l1 = ((B1 - C1) * ((float)x + offset - C0) + (C0 - B0) * ((float)y + offset - C1)) / den;
            float l2 = ((C1 - A1) * ((float)x + offset - C0) + (A0 - C0) * ((float)y + offset - C1)) / den;
            float l3 = 1.f - l1 - l2;
            const float eps = 1e-6f;
            if (l1 >= -eps && l2 >= -eps && l3 >= -eps) {
                float col[3];
                for (int k = 0; k < 3; ++k) col[k] = colA[k] * l1 + colB[k] * l2 + colC[k] * l3;
                float d = dA * l1 + dB * l2 + dC * l3;
                size_t o = (size_t)y * c.out_w + x;
                float& nd = c.new_depth[o];
                float& shape = c.shape[o];
                float ratio = d / nd;
                bool is_valid = ratio * ratio * ratio * shape < tv;
                uint8_t& np = c.new_prol[o];
                if ((is_valid || np) && !prol) {
                    nd = d; np = 0; shape = tv;
                    c.res[3 * o] = col[0]; c.res[3 * o + 1] = col[1]; c.res[3 * o + 2] = col[2];
                } else if (is_valid && np && prol) {
                    nd = d; shape = tv;
                    c.res[3 * o] = col[0]; c.res[3 * o + 1] = col[1]; c.res[3 * o + 2] = col[2];
                }
            }
        }
}

} // namespace

void orc_transform_triangles(const float* color, const float* depth, const float* uv,
                             int in_w, int in_h, int out_w, int out_h, int wrap, float rescale,
                             float* out_color, float* out_depth, float* out_shape) {
    size_t nin = (size_t)in_w * in_h, nout = (size_t)out_w * out_h;
    std::vector<uint8_t> mask(nin), prol(nout, 1);
    for (size_t k = 0; k < nin; ++k) mask[k] = depth[k] > 0.f;
    for (size_t k = 0; k < nout; ++k) {
        out_color[3 * k] = out_color[3 * k + 1] = out_color[3 * k + 2] = 0.f;
        out_depth[k] = kInf;
        out_shape[k] = 0.f;
    }
    RasterCtx c{color, mask.data(), depth, uv, in_w, out_w, out_h, rescale,
                out_color, out_depth, prol.data(), out_shape};
    const int w = in_w;
    auto D = [&](int i, int j) { return depth[(size_t)i * in_w + j]; };
    auto U = [&](int i, int j) { return uv[((size_t)i * in_w + j) * 2]; };
    for (int i = 0; i < in_h - 1; ++i)
        for (int j = 0; j < in_w - 1; ++j) {
            if (D(i, j + 1) > 0.f && D(i + 1, j) > 0.f && U(i, j + 1) > 0.f && U(i + 1, j) > 0.f) {
                if (D(i, j) > 0.f && U(i, j) > 0.f)
                    colorize_triangle(c, j, i, j + 1, i, j, i + 1);
                if (D(i + 1, j + 1) > 0.f && U(i + 1, j + 1) > 0.f)
                    colorize_triangle(c, j + 1, i + 1, j, i + 1, j + 1, i);
                if (wrap && j == 0) {
                    colorize_triangle(c, w - 1, i, 0, i, w - 1, i + 1);
                    colorize_triangle(c, 0, i + 1, w - 1, i + 1, 0, i);
                }
            }
        }
}

/* --------------------------------------------------------------------- blend */

void orc_blend_step(orc_blender* b, const float* color, const float* depth, const float* shape) {
    const size_t n = (size_t)b->w * b->h;
    const float empty[3] = {0.f, 1.f, 0.f};
    if (b->empty) {
        b->empty = 0;
        for (size_t k = 0; k < n; ++k) {
            for (int c = 0; c < 3; ++c) b->color[3 * k + c] = color[3 * k + c];
            b->depth[k] = depth[k];
            b->quality[k] = shape[k] / depth[k];
            b->validity[k] = shape[k];
            b->depth_mask[k] = depth[k] > 0.f ? 255 : 0;
        }
        return;
    }
    const float a = b->exp;
    for (size_t k = 0; k < n; ++k) {
        const float q_new = shape[k] / depth[k];
        const float q_acc = b->quality[k];
        const float* cols[2] = {b->color + 3 * k, color + 3 * k};
        const float qs[2] = {q_acc, q_new};
        const bool masks[2] = {b->depth_mask[k] != 0, depth[k] > 0.f};
        float out[3];
        float outQ;
        uint8_t outMask;
        if (a < 0) {
            /* blend_img_by_max (blending.cpp:293-308) */
            out[0] = empty[0]; out[1] = empty[1]; out[2] = empty[2];
            outQ = 0.f;
            outMask = 0;
            for (int i = 0; i < 2; ++i) {
                if (qs[i] > outQ) {
                    out[0] = cols[i][0]; out[1] = cols[i][1]; out[2] = cols[i][2];
                    outQ = qs[i];
                    outMask = i == 0 ? b->depth_mask[k] : (masks[1] ? 255 : 0);
                }
            }
        } else {
            /* blend_img (blending.cpp:310-369) */
            float sumWeights = 0.0f;
            float fin[3] = {0.f, 0.f, 0.f};
            float inp[3] = {0.f, 0.f, 0.f};
            float inpSum = 0.0f;
            for (int i = 0; i < 2; ++i) {
                float quality = qs[i];
                bool isValidDepth = qs[i] > 0 && !masks[i];
                if (quality > 0 && isValidDepth) {
                    quality = powf(quality, a);
                    sumWeights += quality;
                    for (int c = 0; c < 3; ++c) fin[c] += cols[i][c] * quality;
                } else if (quality > 0 && qs[i] > 0) {
                    quality = powf(quality, a);
                    inpSum += quality;
                    for (int c = 0; c < 3; ++c) inp[c] += cols[i][c] * quality;
                }
            }
            if (sumWeights == 0) {
                if (inpSum == 0.0) {
                    out[0] = empty[0]; out[1] = empty[1]; out[2] = empty[2];
                    outMask = 1;
                    outQ = 0.0f;
                } else {
                    float r = 1.f / inpSum;
                    for (int c = 0; c < 3; ++c) out[c] = inp[c] * r;
                    outMask = 1;
                    outQ = powf(inpSum, 1.0f / a);
                }
            } else {
                float r = 1.f / sumWeights;
                for (int c = 0; c < 3; ++c) out[c] = fin[c] * r;
                outMask = 0;
                outQ = powf(sumWeights, 1.0f / a);
            }
        }
        /* BlendedView.cpp:206-209: depth where q_new > q_acc (old), validity max */
        if (q_new > q_acc) b->depth[k] = depth[k];
        b->validity[k] = (b->validity[k] < shape[k]) ? shape[k] : b->validity[k];
        for (int c = 0; c < 3; ++c) b->color[3 * k + c] = out[c];
        b->quality[k] = outQ;
        b->depth_mask[k] = outMask;
    }
}

/* ------------------------------------------------------------------- inpaint */

int orc_inpaint(float* color, const float* quality, int w, int h, int32_t* map_out) {
    const size_t n = (size_t)w * h;
    std::vector<int32_t> map(3 * n);
    std::vector<uint8_t> mask(n);
    for (size_t k = 0; k < n; ++k) mask[k] = !(quality[k] > 0.f);
    for (int x = 0; x < w; x++)
        for (int y = 0; y < h; y++) {
            size_t o = 3 * ((size_t)y * w + x);
            map[o] = x; map[o + 1] = y; map[o + 2] = 0;
            if (mask[(size_t)y * w + x]) map[o + 2] = w + h;
        }
    /* compute_nearest (inpainting.cpp:104-136) */
    bool change = true, odd = true;
    int sweeps = 0;
    while (change) {
        odd = !odd;
        change = false;
        ++sweeps;
        int A = 0, B = 1, C = 0, D = 1;
        if (odd) { A = w - 1; B = -1; C = h - 1; D = -1; }
        for (int x = A; x < w && x > -1; x += B)
            for (int y = C; y < h && y > -1; y += D) {
                int32_t* cur = &map[3 * ((size_t)y * w + x)];
                if (cur[2] > 0) {
                    for (int i = std::max(x - 1, 0); i < std::min(x + 2, w); ++i)
                        for (int j = std::max(y - 1, 0); j < std::min(y + 2, h); ++j) {
                            if (std::abs(x - i) + std::abs(y - j) == 1) {
                                const int32_t* path = &map[3 * ((size_t)j * w + i)];
                                if (path[2] + 1 < cur[2]) {
                                    int32_t p0 = path[0], p1 = path[1], p2 = path[2] + 1;
                                    cur[0] = p0; cur[1] = p1; cur[2] = p2;
                                    change = true;
                                }
                            }
                        }
                }
            }
    }
    /* inpaint_color (inpainting.cpp:179-198): dst aliases src */
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            if (mask[(size_t)y * w + x]) {
                const int32_t* pix = &map[3 * ((size_t)y * w + x)];
                size_t s = (size_t)pix[1] * w + pix[0];
                size_t d = (size_t)y * w + x;
                float c0 = color[3 * s], c1 = color[3 * s + 1], c2 = color[3 * s + 2];
                color[3 * d] = c0; color[3 * d + 1] = c1; color[3 * d + 2] = c2;
            }
    if (map_out) std::memcpy(map_out, map.data(), map.size() * sizeof(int32_t));
    return sweeps;
}

/* ------------------------------------------------------------ decode/encode */

namespace {
template <typename T>
void decode_color_t(const T* y, const T* u, const T* v, int w, int h, float cs, float* out) {
    const int inW = w / 2, inH = h / 2;
    const float scaleX = (float)(inW) / w;
    const float scaleY = (float)(inH) / h;
    for (int yy = 0; yy < h; ++yy)
        for (int xx = 0; xx < w; ++xx) {
            float srcX = xx * scaleX, srcY = yy * scaleY;
            int x1 = (int)srcX, y1 = (int)srcY;
            int x2 = (int)fminf((float)(x1 + 1), (float)(inW - 1));
            int y2 = (int)fminf((float)(y1 + 1), (float)(inH - 1));
            float dx = srcX - x1, dy = srcY - y1;
            float up[2];
            const T* planes[2] = {u, v};
            for (int p = 0; p < 2; ++p) {
                const T* in = planes[p];
                float val1 = fmaf((float)in[y1 * inW + x1], (1 - dx), (float)in[y1 * inW + x2] * dx);
                float val2 = fmaf((float)in[y2 * inW + x1], (1 - dx), (float)in[y2 * inW + x2] * dx);
                up[p] = fmaf(val1, (1 - dy), val2 * dy);
            }
            size_t o = (size_t)yy * w + xx;
            out[3 * o] = (float)y[o] * cs;
            out[3 * o + 1] = up[0] * cs;
            out[3 * o + 2] = up[1] * cs;
        }
}

template <typename T>
void decode_depth_t(const T* raw, int w, int h, float scale, float znear, float zfar, int has_invalid, float* out) {
    const size_t n = (size_t)w * h;
    for (size_t k = 0; k < n; ++k) {
        float depth = (float)raw[k] * scale;
        if (zfar >= 1000.f)
            depth = znear / depth;
        else
            depth = zfar * znear / fmaf(depth, (zfar - znear), znear);
        if (has_invalid && depth == 0) depth = kNaN;
        out[k] = depth;
    }
}

template <typename T>
inline T quantize(float v, unsigned maxv) {
    return (T)(uint32_t)cvt_i32(v * (float)maxv);
}

template <typename T>
void encode_t(const float* color, int w, int h, unsigned maxv, T* Y, T* U, T* V) {
    std::vector<T> uf((size_t)w * h), vf((size_t)w * h);
    for (size_t k = 0; k < (size_t)w * h; ++k) {
        Y[k] = quantize<T>(color[3 * k], maxv);
        uf[k] = quantize<T>(color[3 * k + 1], maxv);
        vf[k] = quantize<T>(color[3 * k + 2], maxv);
    }
    const int hw = w / 2, hh = h / 2;
    for (int y = 0; y < hh; ++y)
        for (int x = 0; x < hw; ++x) {
            size_t a = (size_t)(2 * y) * w + 2 * x, b = a + w;
            U[(size_t)y * hw + x] = (T)((uf[a] + uf[a + 1] + uf[b] + uf[b + 1] + 2) >> 2);
            V[(size_t)y * hw + x] = (T)((vf[a] + vf[a + 1] + vf[b] + vf[b + 1] + 2) >> 2);
        }
}
} // namespace

void orc_decode_color(const void* y, const void* u, const void* v, int bitdepth,
                      int w, int h, float color_scale, float* out) {
    if (bitdepth <= 8)
        decode_color_t((const uint8_t*)y, (const uint8_t*)u, (const uint8_t*)v, w, h, color_scale, out);
    else
        decode_color_t((const uint16_t*)y, (const uint16_t*)u, (const uint16_t*)v, w, h, color_scale, out);
}

void orc_decode_depth(const void* raw, int bitdepth, int w, int h, float depth_scale,
                      float znear, float zfar, int has_invalid, float* out) {
    if (bitdepth <= 8)
        decode_depth_t((const uint8_t*)raw, w, h, depth_scale, znear, zfar, has_invalid, out);
    else
        decode_depth_t((const uint16_t*)raw, w, h, depth_scale, znear, zfar, has_invalid, out);
}

void orc_encode_yuv420(const float* color, int w, int h, int bitdepth, void* y, void* u, void* v) {
    unsigned maxv = orc_max_level(bitdepth);
    if (bitdepth <= 8)
        encode_t(color, w, h, maxv, (uint8_t*)y, (uint8_t*)u, (uint8_t*)v);
    else
        encode_t(color, w, h, maxv, (uint16_t*)y, (uint16_t*)u, (uint16_t*)v);
}

/* ------------------------------------------------------------- resize */

namespace {
/* one axis of the INTER_LINEAR rule (imgproc/resize.cpp, resize_ / HResizeLinear) */
void lin_axis(int d, int n_src, int n_dst, int& s0, int& s1, float& f) {
    const double scale = (double)n_src / (double)n_dst;
    f = (float)((d + 0.5) * scale - 0.5);
    s0 = (int)std::floor(f);
    f -= (float)s0;
    if (s0 < 0) { f = 0.f; s0 = 0; }
    if (s0 >= n_src - 1) { f = 0.f; s0 = n_src - 1; }
    s1 = s0 + 1 < n_src ? s0 + 1 : s0;
}
}  // namespace

void orc_resize_color(const float* src, int sw, int sh, float* dst, int dw, int dh) {
    const bool area2 = sw == 2 * dw && sh == 2 * dh;
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x)
            for (int c = 0; c < 3; ++c) {
                float v;
                if (area2) {
                    const float* r0 = src + 3 * ((size_t)(2 * y) * sw + 2 * x) + c;
                    const float* r1 = r0 + 3 * (size_t)sw;
                    v = (((r0[0] + r0[3]) + r1[0]) + r1[3]) * 0.25f;
                } else {
                    int x0, x1, y0, y1;
                    float fx, fy;
                    lin_axis(x, sw, dw, x0, x1, fx);
                    lin_axis(y, sh, dh, y0, y1, fy);
                    const float gx = 1.f - fx, gy = 1.f - fy;
                    auto at = [&](int r, int q) { return src[3 * ((size_t)r * sw + q) + c]; };
                    const float h0 = at(y0, x0) * gx + at(y0, x1) * fx;
                    const float h1 = at(y1, x0) * gx + at(y1, x1) * fx;
                    v = h0 * gy + h1 * fy;
                }
                dst[3 * ((size_t)y * dw + x) + c] = v;
            }
}

/* ------------------------------------------------------ multispectral */

void orc_blur_masked(const float* y, int stride, const uint8_t* mask, int w, int h, int rad, float* out) {
    /* IntegralImage2D<float, 1>::computeIntegralImages with a mask
     * (IntegralImage2D.h:738-770): double sums, finite values under mask == 255,
     * I[r+1][c+1] = ((I[r][c+1] + I[r+1][c]) - I[r][c]) (+ value) */
    const size_t W1 = (size_t)w + 1;
    std::vector<double> I(W1 * (h + 1), 0.0);
    std::vector<unsigned> N(W1 * (h + 1), 0u);
    for (int r = 0; r < h; ++r)
        for (int c = 0; c < w; ++c) {
            const size_t a = (size_t)r * W1, b = (size_t)(r + 1) * W1;
            I[b + c + 1] = I[a + c + 1] + I[b + c] - I[a + c];
            N[b + c + 1] = N[a + c + 1] + N[b + c] - N[a + c];
            const float v = y[((size_t)r * w + c) * stride];
            if (mask[(size_t)r * w + c] == 255 && std::isfinite(v)) {
                I[b + c + 1] += v;
                N[b + c + 1] += 1;
            }
        }
    /* calcBlurring, one channel (blending.cpp:62-104) */
    const int rad_1 = 2 * rad + 1;
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            float& o = out[(size_t)i * w + j];
            if (mask[(size_t)i * w + j] == 0) { o = 0.f; continue; }
            int sx = j - rad, sy = i - rad, ey = sy + rad_1, ex = sx + rad_1;
            sx = std::max(0, sx); sy = std::max(0, sy);
            ex = std::min(ex, w); ey = std::min(ey, h);
            const size_t ul = (size_t)sy * W1 + sx, ur = (size_t)sy * W1 + ex;
            const size_t ll = (size_t)ey * W1 + sx, lr = (size_t)ey * W1 + ex;
            const unsigned cnt = N[lr] + N[ul] - N[ur] - N[ll];
            const unsigned sum = (unsigned)(int)(I[lr] + I[ul] - I[ur] - I[ll]);
            /* the reference leaves blr at OpenCV's uninitialised value when cnt == 0;
             * cnt >= 1 for every pixel under the mask (it counts itself when finite) */
            o = cnt > 0 ? float(float(sum) / float(cnt)) : 0.f;
        }
}

void orc_blur_masked_rgb(const float* color, const uint8_t* mask, int w, int h, int rad, float* out) {
    /* IntegralImage2D<float, 3> with a mask (IntegralImage2D.h:385-420): one
     * double integral per channel in the same recurrence, a pixel counted when
     * mask == 255 and cv::sum of its three channels (a double) is finite */
    const size_t W1 = (size_t)w + 1;
    std::vector<double> I(3 * W1 * (h + 1), 0.0);
    std::vector<unsigned> N(W1 * (h + 1), 0u);
    for (int r = 0; r < h; ++r)
        for (int c = 0; c < w; ++c) {
            const size_t a = (size_t)r * W1, b = (size_t)(r + 1) * W1;
            for (int ch = 0; ch < 3; ++ch)
                I[3 * (b + c + 1) + ch] = I[3 * (a + c + 1) + ch] + I[3 * (b + c) + ch] - I[3 * (a + c) + ch];
            N[b + c + 1] = N[a + c + 1] + N[b + c] - N[a + c];
            const float* v = color + 3 * ((size_t)r * w + c);
            const double vs = ((double)v[0] + (double)v[1]) + (double)v[2];
            if (mask[(size_t)r * w + c] == 255 && std::isfinite(vs)) {
                for (int ch = 0; ch < 3; ++ch) I[3 * (b + c + 1) + ch] += v[ch];
                N[b + c + 1] += 1;
            }
        }
    /* calcBlurring, three channels (blending.cpp:109-159): the window's Vec3d
     * sum converted to Vec3f (rounded, no integer truncation), then divided */
    const int rad_1 = 2 * rad + 1;
    for (int i = 0; i < h; ++i)
        for (int j = 0; j < w; ++j) {
            float* o = out + 3 * ((size_t)i * w + j);
            if (mask[(size_t)i * w + j] == 0) { o[0] = o[1] = o[2] = 0.f; continue; }
            int sx = j - rad, sy = i - rad, ey = sy + rad_1, ex = sx + rad_1;
            sx = std::max(0, sx); sy = std::max(0, sy);
            ex = std::min(ex, w); ey = std::min(ey, h);
            const size_t ul = (size_t)sy * W1 + sx, ur = (size_t)sy * W1 + ex;
            const size_t ll = (size_t)ey * W1 + sx, lr = (size_t)ey * W1 + ex;
            const unsigned cnt = N[lr] + N[ul] - N[ur] - N[ll];
            for (int ch = 0; ch < 3; ++ch) {
                const float sum = (float)(I[3 * lr + ch] + I[3 * ul + ch] - I[3 * ur + ch] - I[3 * ll + ch]);
                o[ch] = cnt > 0 ? float(sum / float(cnt)) : 0.f;
            }
        }
}

void orc_split_frequencies(const float* color, const float* depth, const float* shape, int w, int h, int rgb,
                           float* low, float* high) {
    /* BlendedViewMultiSpec::blend (BlendedView.cpp:70-88): mask = quality > 0
     * with quality = shape / depth (SynthesizedView.cpp:440); split_frequencies
     * (blending.cpp:371-394), radius max(rows, cols) / 20; ColorSpace YUV blurs
     * Y only, RGB blurs all three channels */
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> mask(n);
    for (size_t k = 0; k < n; ++k) mask[k] = (shape[k] / depth[k] > 0.f) ? 255 : 0;
    const int rad = std::max(w, h) / 20;
    if (rgb) {
        orc_blur_masked_rgb(color, mask.data(), w, h, rad, low);
    } else {
        std::vector<float> yb(n);
        orc_blur_masked(color, 3, mask.data(), w, h, rad, yb.data());
        for (size_t k = 0; k < n; ++k) {
            low[3 * k] = yb[k];
            low[3 * k + 1] = color[3 * k + 1];
            low[3 * k + 2] = color[3 * k + 2];
        }
    }
    for (size_t k = 0; k < 3 * n; ++k) high[k] = color[k] - low[k];
}

/* ------------------------------------------------------------- whole frame */

namespace {
int synthesize_frame(const orc_source* srcs, int n_src, const orc_cam* tgt, const float tgt_pos[3],
                     const float tgt_rot[3], int tgt_bitdepth, float blending_exp, float rescale, int multispectral,
                     float exp_low, float exp_high, float* out_color, float* out_quality, float* out_depth,
                     float* out_validity, void* out_y, void* out_u, void* out_v, orc_timing* timing);
}  // namespace

int orc_synthesize(const orc_source* srcs, int n_src, const orc_cam* tgt,
                   const float tgt_pos[3], const float tgt_rot[3], int tgt_bitdepth,
                   float blending_exp, float rescale,
                   float* out_color, float* out_quality, float* out_depth, float* out_validity,
                   void* out_y, void* out_u, void* out_v, orc_timing* timing) {
    return synthesize_frame(srcs, n_src, tgt, tgt_pos, tgt_rot, tgt_bitdepth, blending_exp, rescale, 0, 0.f, 0.f,
                            out_color, out_quality, out_depth, out_validity, out_y, out_u, out_v, timing);
}

int orc_synthesize_multispectral(const orc_source* srcs, int n_src, const orc_cam* tgt, const float tgt_pos[3],
                                 const float tgt_rot[3], int tgt_bitdepth, float exp_low, float exp_high, int rgb,
                                 float rescale, float* out_color, float* out_quality, float* out_depth,
                                 float* out_validity, void* out_y, void* out_u, void* out_v) {
    return synthesize_frame(srcs, n_src, tgt, tgt_pos, tgt_rot, tgt_bitdepth, 0.f, rescale, rgb ? 2 : 1, exp_low,
                            exp_high,
                            out_color, out_quality, out_depth, out_validity, out_y, out_u, out_v, nullptr);
}

namespace {
int synthesize_frame(const orc_source* srcs, int n_src, const orc_cam* tgt, const float tgt_pos[3],
                     const float tgt_rot[3], int tgt_bitdepth, float blending_exp, float rescale, int multispectral,
                     float exp_low, float exp_high, float* out_color, float* out_quality, float* out_depth,
                     float* out_validity, void* out_y, void* out_u, void* out_v, orc_timing* timing) {
    if (n_src <= 0) return -1;
    orc_timing tm;
    std::memset(&tm, 0, sizeof(tm));
    const int ow = (int)(0.5f + tgt->w * rescale);
    const int oh = (int)(0.5f + tgt->h * rescale);
    const size_t nout = (size_t)ow * oh;
    /* function-static scales of read_color_YUV / read_depth_YUV
     * (image_loading.cpp:134,193): initialised from the first view loaded */
    const float color_scale = 1.0f / orc_max_level(srcs[0].bitdepth_color);
    const float depth_scale = 1.0f / orc_max_level(srcs[0].bitdepth_depth);

    std::vector<float> bcol(3 * nout), bdep(nout), bq(nout), bval(nout);
    std::vector<uint8_t> bmask(nout);
    orc_blender bl{ow, oh, 1, blending_exp, bcol.data(), bdep.data(), bq.data(), bval.data(), bmask.data()};
    /* multispectral: the low band in bl (its depth, quality and validity are the
     * result's), the high band in bh, each with its own exponent */
    std::vector<float> hcol, hdep, hq, hval, lowc, highc;
    std::vector<uint8_t> hmask;
    orc_blender bh{};
    if (multispectral) {
        bl.exp = exp_low;
        hcol.resize(3 * nout); hdep.resize(nout); hq.resize(nout); hval.resize(nout); hmask.resize(nout);
        lowc.resize(3 * nout); highc.resize(3 * nout);
        bh = orc_blender{ow, oh, 1, exp_high, hcol.data(), hdep.data(), hq.data(), hval.data(), hmask.data()};
    }
    std::vector<float> vcol(3 * nout), vdep(nout), vshape(nout);

    for (int s = 0; s < n_src; ++s) {
        const orc_source& S = srcs[s];
        const size_t nin = (size_t)S.cam.w * S.cam.h;
        double t0 = now_s();
        std::vector<float> col(3 * nin), dep(nin);
        orc_decode_color(S.y, S.u, S.v, S.bitdepth_color, S.cam.w, S.cam.h, color_scale, col.data());
        orc_decode_depth(S.depth, S.bitdepth_depth, S.cam.w, S.cam.h, depth_scale, S.znear, S.zfar,
                         S.has_invalid, dep.data());
        double t1 = now_s();
        float R[9], t[3];
        orc_relative_pose(S.pos, S.rot, tgt_pos, tgt_rot, R, t);
        std::vector<float> uv(2 * nin), vd(nin);
        int wrap = orc_reproject(&S.cam, tgt, R, t, dep.data(), rescale, uv.data(), vd.data());
        double t2 = now_s();
        orc_transform_triangles(col.data(), vd.data(), uv.data(), S.cam.w, S.cam.h, ow, oh, wrap, rescale,
                                vcol.data(), vdep.data(), vshape.data());
        double t3 = now_s();
        if (multispectral) {
            orc_split_frequencies(vcol.data(), vdep.data(), vshape.data(), ow, oh, multispectral == 2 ? 1 : 0,
                                  lowc.data(), highc.data());
            orc_blend_step(&bl, lowc.data(), vdep.data(), vshape.data());
            orc_blend_step(&bh, highc.data(), vdep.data(), vshape.data());
        } else {
            orc_blend_step(&bl, vcol.data(), vdep.data(), vshape.data());
        }
        double t4 = now_s();
        tm.t_decode += t1 - t0;
        tm.t_reproject += t2 - t1;
        tm.t_raster += t3 - t2;
        tm.t_blend += t4 - t3;
    }
    /* multispectral result: low + high colour, the low band's maps (BlendedView.cpp:86) */
    if (multispectral)
        for (size_t k = 0; k < 3 * nout; ++k) bcol[k] = bcol[k] + hcol[k];
    if (out_quality) std::memcpy(out_quality, bq.data(), nout * sizeof(float));
    if (out_depth) std::memcpy(out_depth, bdep.data(), nout * sizeof(float));
    if (out_validity) std::memcpy(out_validity, bval.data(), nout * sizeof(float));
    double t5 = now_s();
    tm.inpaint_sweeps = orc_inpaint(bcol.data(), bq.data(), ow, oh, nullptr);
    double t6 = now_s();
    tm.t_inpaint = t6 - t5;
    if (out_color) std::memcpy(out_color, bcol.data(), 3 * nout * sizeof(float));
    if (out_y) {
        if (ow != tgt->w || oh != tgt->h) {
            /* write_color resizes the oversampled colour to the view's size first */
            std::vector<float> rcol(3 * (size_t)tgt->w * tgt->h);
            orc_resize_color(bcol.data(), ow, oh, rcol.data(), tgt->w, tgt->h);
            orc_encode_yuv420(rcol.data(), tgt->w, tgt->h, tgt_bitdepth, out_y, out_u, out_v);
        } else {
            orc_encode_yuv420(bcol.data(), ow, oh, tgt_bitdepth, out_y, out_u, out_v);
        }
    }
    tm.t_encode = now_s() - t6;
    if (timing) *timing = tm;
    return 0;
}
}  // namespace
