// pipeline.cpp -- the per-frame loop and the view classes on top of the C ABI
// (rvs/src/Pipeline.cpp, View.cpp, SynthesizedView.cpp, BlendedView.cpp,
// Application.cpp, image_loading.cpp, image_writing.cpp).
#include "rvs/pipeline.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <typeinfo>

namespace rvs {

// Raster size at Precision `rescale` (SynthesizedView.cpp:326-329).
static Size raster_size(Size s, float rescale) {
    return Size{(int)(0.5f + s.width * rescale), (int)(0.5f + s.height * rescale)};
}

namespace {

void check(int rc, const char* what) {
    if (rc != RVS_OK) {
        std::ostringstream m;
        m << what << ": " << rvs_last_error();
        throw std::runtime_error(m.str());
    }
}

// Entry-list capacity for a re-run after RVS_E_CAPACITY (needed + 1/8, saturated).
uint32_t grow_entry_cap(uint32_t needed) {
    const uint64_t c = (uint64_t)needed + needed / 8 + 1024;
    return c > 0xffffffffull ? 0xffffffffu : (uint32_t)c;
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::ostringstream m;
        m << what << ": " << hipGetErrorString(e);
        throw std::runtime_error(m.str());
    }
}

unsigned max_level(int bitdepth) { return (1u << bitdepth) - 1u; }
int bytes_per(int bitdepth) { return bitdepth <= 8 ? 1 : 2; }

rvs_cam cam_of(Parameters const& p) {
    rvs_cam c;
    std::memset(&c, 0, sizeof c);
    const bool erp = p.getProjectionType() == ProjectionType::equirectangular;
    if (!erp && p.getProjectionType() != ProjectionType::perspective) {
        std::ostringstream m;
        m << "Unknown projection type \"" << p.getProjectionType() << "\"";
        throw std::runtime_error(m.str());
    }
    c.proj = erp ? 1 : 0;
    c.w = p.getSize().width;
    c.h = p.getSize().height;
    if (erp) {
        c.hor[0] = p.getHorRange()[0]; c.hor[1] = p.getHorRange()[1];
        c.ver[0] = p.getVerRange()[0]; c.ver[1] = p.getVerRange()[1];
        c.full_hor = p.isFullHorRange() ? 1 : 0;
    } else {
        c.f[0] = p.getFocal()[0]; c.f[1] = p.getFocal()[1];
        c.p[0] = p.getPrinciplePoint()[0]; c.p[1] = p.getPrinciplePoint()[1];
    }
    return c;
}

void read_at(std::string const& path, long offset_per_frame, int frame, std::vector<std::vector<uint8_t>*> planes,
             const char* what) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        std::ostringstream m;
        m << "Failed to read raw YUV " << what << " file \"" << path << "\"";
        throw std::runtime_error(m.str());
    }
    f.seekg(0, std::ios::end);
    const long size = (long)f.tellg();
    long offset = offset_per_frame * frame;
    if (offset >= size) offset = 0;    // past the last frame: back to frame 0 (image_loading.cpp:97-100)
    f.seekg(offset, std::ios::beg);
    for (auto* p : planes) {
        f.read(reinterpret_cast<char*>(p->data()), (std::streamsize)p->size());
        if ((std::size_t)f.gcount() != p->size()) throw std::runtime_error("Failed to read the expected amount of data");
    }
}

}  // namespace

// ------------------------------------------------------------------ I/O

void read_yuv420(std::string const& path, int frame, Size size, int bitdepth, std::vector<uint8_t>& y,
                 std::vector<uint8_t>& u, std::vector<uint8_t>& v) {
    const int e = bytes_per(bitdepth);
    y.resize((std::size_t)size.area() * e);
    u.resize((std::size_t)(size.width / 2) * (size.height / 2) * e);
    v.resize(u.size());
    read_at(path, size.area() * e * 3 / 2, frame, {&y, &u, &v}, "color");
}

void read_depth_plane(std::string const& path, int frame, Size size, int bitdepth, ColorFormat fmt,
                      std::vector<uint8_t>& y) {
    const int e = bytes_per(bitdepth);
    y.resize((std::size_t)size.area() * e);
    const long per = fmt == ColorFormat::YUV420 ? size.area() * e * 3 / 2 : size.area() * e;
    read_at(path, per, frame, {&y}, "depth");
}

void write_frame(std::string const& path, int frame, std::vector<uint8_t> const& bytes) {
    std::ofstream f(path, frame ? std::ios::binary | std::ios::app : std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Failed to open YUV output image");
    f.write(reinterpret_cast<const char*>(bytes.data()), (std::streamsize)bytes.size());
}

int getExtendedIndex(int outputFrameIndex, int numberOfInputFrames) {
    if (numberOfInputFrames <= 0) throw std::runtime_error("Cannot extend frame index with zero input frames");
    const int group = outputFrameIndex / numberOfInputFrames;
    const int rel = outputFrameIndex % numberOfInputFrames;
    return group % 2 != 0 ? numberOfInputFrames - rel - 1 : rel;
}

// ------------------------------------------------------------ DeviceBuffer

DeviceBuffer::DeviceBuffer(std::size_t bytes) { resize(bytes); }
DeviceBuffer::~DeviceBuffer() {
    if (m_ptr) (void)hipFree(m_ptr);
}
DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept : m_ptr(o.m_ptr), m_bytes(o.m_bytes) {
    o.m_ptr = nullptr;
    o.m_bytes = 0;
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (m_ptr) (void)hipFree(m_ptr);
        m_ptr = o.m_ptr; m_bytes = o.m_bytes;
        o.m_ptr = nullptr; o.m_bytes = 0;
    }
    return *this;
}
void DeviceBuffer::resize(std::size_t bytes) {
    if (bytes <= m_bytes && m_ptr) return;
    if (m_ptr) (void)hipFree(m_ptr);
    m_ptr = nullptr;
    m_bytes = 0;
    hip_check(hipMalloc(&m_ptr, bytes ? bytes : 1), "hipMalloc");
    m_bytes = bytes;
}
void DeviceBuffer::upload(const void* host, std::size_t bytes, rvs_stream_t s) {
    resize(bytes);
    hip_check(hipMemcpyAsync(m_ptr, host, bytes, hipMemcpyHostToDevice, (hipStream_t)s), "upload");
}
void DeviceBuffer::download(void* host, std::size_t bytes, rvs_stream_t s) const {
    hip_check(hipMemcpyAsync(host, m_ptr, bytes, hipMemcpyDeviceToHost, (hipStream_t)s), "download");
    hip_check(hipStreamSynchronize((hipStream_t)s), "download sync");
}

// ------------------------------------------------------------------- View

void View::allocate(Size s, bool quality) {
    m_size = s;
    const std::size_t n = (std::size_t)s.area();
    m_color.resize(3 * n * sizeof(float));
    m_depth.resize(n * sizeof(float));
    m_validity.resize(n * sizeof(float));
    if (quality) m_quality.resize(n * sizeof(float));
}

namespace {
std::vector<float> host_copy(DeviceBuffer const& b, std::size_t n) {
    std::vector<float> out(n);
    if (n) b.download(out.data(), n * sizeof(float), nullptr);
    return out;
}
}  // namespace

std::vector<float> View::get_color() const { return host_copy(m_color, 3 * (std::size_t)m_size.area()); }
std::vector<float> View::get_depth() const { return host_copy(m_depth, (std::size_t)m_size.area()); }
std::vector<float> View::get_quality() const { return host_copy(m_quality, (std::size_t)m_size.area()); }
std::vector<float> View::get_validity() const { return host_copy(m_validity, (std::size_t)m_size.area()); }

// -------------------------------------------------------------- InputView

InputView::InputView(std::string const& color_path, std::string const& depth_path, int frame,
                     Parameters const& params, float color_scale, float depth_scale, rvs_stream_t stream)
    : m_params(params), m_color_scale(color_scale), m_depth_scale(depth_scale) {
    if (params.getDisplacementMethod() != DisplacementMethod::depth)
        throw std::runtime_error("Only the depth displacement method is supported by this build");
    if (params.getPaddedSize() != params.getSize())
        throw std::runtime_error("Crop_region is not supported by this build");
    if (color_path.size() < 4 || color_path.substr(color_path.size() - 4) != ".yuv" || depth_path.size() < 4 ||
        depth_path.substr(depth_path.size() - 4) != ".yuv")
        throw std::runtime_error("Only raw .yuv input files are supported by this build");
    m_size = params.getSize();
    std::vector<uint8_t> y, u, v, d;
    read_yuv420(color_path, frame, m_size, params.getColorBitDepth(), y, u, v);
    read_depth_plane(depth_path, frame, params.getPaddedSize(), params.getDepthBitDepth(),
                     params.getDepthColorFormat(), d);
    m_y.upload(y.data(), y.size(), stream);
    m_u.upload(u.data(), u.size(), stream);
    m_v.upload(v.data(), v.size(), stream);
    m_raw_depth.upload(d.data(), d.size(), stream);
    const rvs_cam c = cam_of(params);
    if (c.proj == 1) {
        std::vector<float> tab(rvs_erp_table_len(&c));
        check(rvs_erp_table_fill(&c, tab.data()), "rvs_erp_table_fill");
        m_table.upload(tab.data(), tab.size() * sizeof(float), stream);
    }
    // culling table of this frame's depth, kept with the view
    rvs_source s = source();
    m_prep.resize(rvs_source_prep_size(&s.cam));
    check(rvs_source_prep(&s, m_depth_scale, m_prep.get(), m_prep.bytes(), stream), "rvs_source_prep");
    m_prep_scale = m_depth_scale;
    hip_check(hipStreamSynchronize((hipStream_t)stream), "input upload");
}

rvs_source InputView::source() const {
    rvs_source s;
    std::memset(&s, 0, sizeof s);
    s.cam = cam_of(m_params);
    s.bitdepth_color = m_params.getColorBitDepth();
    s.bitdepth_depth = m_params.getDepthBitDepth();
    s.znear = m_params.getDepthRange()[0];
    s.zfar = m_params.getDepthRange()[1];
    s.has_invalid = m_params.hasInvalidDepth() ? 1 : 0;
    for (int k = 0; k < 3; ++k) { s.pos[k] = m_params.getPosition()[k]; s.rot[k] = m_params.getRotation()[k]; }
    s.y = m_y.get(); s.u = m_u.get(); s.v = m_v.get(); s.depth = m_raw_depth.get();
    s.erp_table = s.cam.proj == 1 ? m_table.as<float>() : nullptr;
    s.prep = m_prep.bytes() ? m_prep.as<float>() : nullptr;
    s.prep_depth_scale = m_prep_scale;
    return s;
}

void InputView::decode(rvs_stream_t stream) {
    if (m_decoded) return;
    allocate(m_size, false);
    check(rvs_decode_yuv420(m_y.get(), m_u.get(), m_v.get(), m_params.getColorBitDepth(), m_size.width,
                            m_size.height, m_color_scale, getDevColor(), stream),
          "rvs_decode_yuv420");
    check(rvs_decode_depth(m_raw_depth.get(), m_params.getDepthBitDepth(), m_size.width, m_size.height,
                           m_depth_scale, m_params.getDepthRange()[0], m_params.getDepthRange()[1],
                           m_params.hasInvalidDepth() ? 1 : 0, getDevDepth(), stream),
          "rvs_decode_depth");
    m_decoded = true;
}

// ------------------------------------------------------- SpaceTransformer

std::array<float, 9> SpaceTransformer::get_rotation() const {
    rvs_xform x;
    const Vec3f pf = m_input->getPosition(), rf = m_input->getRotation();
    const Vec3f pt = m_target->getPosition(), rt = m_target->getRotation();
    check(rvs_relative_pose(pf.data(), rf.data(), pt.data(), rt.data(), &x), "rvs_relative_pose");
    std::array<float, 9> R;
    std::memcpy(R.data(), x.R, sizeof x.R);
    return R;
}

Vec3f SpaceTransformer::get_translation() const {
    rvs_xform x;
    const Vec3f pf = m_input->getPosition(), rf = m_input->getRotation();
    const Vec3f pt = m_target->getPosition(), rt = m_target->getRotation();
    check(rvs_relative_pose(pf.data(), rf.data(), pt.data(), rt.data(), &x), "rvs_relative_pose");
    return {x.t[0], x.t[1], x.t[2]};
}

// ------------------------------------------------ SynthetisedViewTriangle

void SynthetisedViewTriangle::compute(InputView& input, rvs_stream_t stream) {
    if (!m_transformer) throw std::logic_error("SynthesizedView: no space transformer");
    input.decode(stream);
    Parameters const& vp = m_transformer->getVirtualParameters();
    const rvs_cam src = cam_of(input.parameters()), dst = cam_of(vp);
    rvs_xform x;
    const auto R = m_transformer->get_rotation();
    const auto t = m_transformer->get_translation();
    std::memcpy(x.R, R.data(), sizeof x.R);
    std::memcpy(x.t, t.data(), sizeof x.t);
    const std::size_t nin = (std::size_t)src.w * src.h;
    m_uv.resize(2 * nin * sizeof(float));
    m_vdepth.resize(nin * sizeof(float));
    rvs_source s = input.source();
    // rasterised with the oversampling of detail::g_rescale (SynthesizedView.cpp:325-334)
    const float rescale = detail::g_rescale;
    check(rvs_reproject(&src, &dst, &x, input.getDevDepth(), s.erp_table, rescale, m_uv.as<float>(),
                        m_vdepth.as<float>(), stream),
          "rvs_reproject");
    const Size rs = raster_size(vp.getSize(), rescale);
    allocate(rs, false);
    const int wrap = (dst.proj == 1 && dst.full_hor) ? 1 : 0;
    uint32_t cap = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        const std::size_t need = rvs_raster_workspace_size(src.w, src.h, rs.width, rs.height, cap);
        m_ws.resize(need);
        uint32_t needed = 0;
        const int rc = rvs_raster(input.getDevColor(), m_vdepth.as<float>(), m_uv.as<float>(), src.w, src.h,
                                  rs.width, rs.height, wrap, rescale, m_ws.get(), need, cap, getDevColor(),
                                  getDevDepth(), getDevValidity(), &needed, stream);
        if (rc == RVS_E_CAPACITY && needed > 0) {
            cap = grow_entry_cap(needed);
            continue;
        }
        check(rc, "rvs_raster");
        return;
    }
    throw std::runtime_error("rvs_raster: entry capacity could not be satisfied");
}

// ------------------------------------------------------ BlendedViewSimple

void BlendedViewSimple::blend(View const& view, rvs_stream_t stream) {
    blend_maps(view.getDevColor(), view.getDevDepth(), view.getDevValidity(), view.get_size(), stream);
}

void BlendedViewSimple::blend_maps(const float* color, const float* depth, const float* shape, Size s,
                                   rvs_stream_t stream) {
    if (m_empty) {
        allocate(s, true);
        m_mask.resize((std::size_t)s.area());
    } else if (s.width != m_size.width || s.height != m_size.height) {
        throw std::runtime_error("BlendedViewSimple: views of different sizes");
    }
    check(rvs_blend_step(getDevColor(), getDevQuality(), getDevDepth(), getDevValidity(), m_mask.as<uint8_t>(), color,
                         depth, shape, s.width, s.height, m_exp, m_empty ? 1 : 0, stream),
          "rvs_blend_step");
    m_empty = false;
}

// --------------------------------------------------- BlendedViewMultiSpec

void BlendedViewMultiSpec::blend(View const& view, rvs_stream_t stream) {
    const Size s = view.get_size();
    const std::size_t n = (std::size_t)s.area();
    const int rgb = m_rgb ? 1 : 0;
    if (m_empty) allocate(s, true);
    else if (s.width != m_size.width || s.height != m_size.height)
        throw std::runtime_error("BlendedViewMultiSpec: views of different sizes");
    const std::size_t ws = rvs_split_frequencies_workspace_size(s.width, s.height, rgb);
    if (ws == 0) check(RVS_E_INVALID, "rvs_split_frequencies_workspace_size");
    m_ws.resize(ws);
    m_low.resize(3 * n * sizeof(float));
    m_high.resize(3 * n * sizeof(float));
    check(rvs_split_frequencies(view.getDevColor(), view.getDevDepth(), view.getDevValidity(), s.width, s.height, rgb,
                                m_ws.get(), m_ws.bytes(), m_low.as<float>(), m_high.as<float>(), stream),
          "rvs_split_frequencies");
    // the bands as views with the synthesized view's depth and validity (BlendedView.cpp:79-84)
    m_low_freq.blend_maps(m_low.as<float>(), view.getDevDepth(), view.getDevValidity(), s, stream);
    m_high_freq.blend_maps(m_high.as<float>(), view.getDevDepth(), view.getDevValidity(), s, stream);
    // assign(low colour + high colour, the low band's maps) (BlendedView.cpp:86)
    check(rvs_add_color(m_low_freq.getDevColor(), m_high_freq.getDevColor(), getDevColor(), s.width, s.height,
                        stream),
          "rvs_add_color");
    const hipStream_t hs = (hipStream_t)stream;
    hip_check(hipMemcpyAsync(getDevDepth(), m_low_freq.getDevDepth(), n * sizeof(float), hipMemcpyDeviceToDevice, hs),
              "copy");
    hip_check(hipMemcpyAsync(getDevQuality(), m_low_freq.getDevQuality(), n * sizeof(float),
                             hipMemcpyDeviceToDevice, hs),
              "copy");
    hip_check(hipMemcpyAsync(getDevValidity(), m_low_freq.getDevValidity(), n * sizeof(float),
                             hipMemcpyDeviceToDevice, hs),
              "copy");
    m_empty = false;
}

// --------------------------------------------------------------- Pipeline

Pipeline::Pipeline() {
    hip_check(hipGetDevice(&m_device), "hipGetDevice");
    hip_check(hipStreamCreate((hipStream_t*)&m_stream), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags((hipStream_t*)&m_save_stream, hipStreamNonBlocking), "hipStreamCreate");
    for (FrameOut& f : m_out)
        hip_check(hipEventCreateWithFlags((hipEvent_t*)&f.done, hipEventDisableTiming), "hipEventCreate");
}

Pipeline::~Pipeline() {
    if (m_save.valid()) m_save.wait();
    for (FrameOut& f : m_out)
        if (f.done) (void)hipEventDestroy((hipEvent_t)f.done);
    if (m_save_stream) (void)hipStreamDestroy((hipStream_t)m_save_stream);
    if (m_stream) (void)hipStreamDestroy((hipStream_t)m_stream);
}

Pipeline::FrameOut* Pipeline::frameOf(const float3* color) {
    for (FrameOut& f : m_out)
        if (f.color.get() == (const void*)color) return &f;
    return nullptr;
}

float Pipeline::colorScale(int bitdepth) {
    if (m_color_scale < 0) m_color_scale = 1.0f / max_level(bitdepth);
    return m_color_scale;
}

float Pipeline::depthScale(int bitdepth) {
    if (m_depth_scale < 0) m_depth_scale = 1.0f / max_level(bitdepth);
    return m_depth_scale;
}

int Pipeline::execute() {
    Config const& c = getConfig();
    if (c.VirtualCameraNames.size() != 1) throw std::runtime_error("Exactly one virtual camera is supported");
    for (int vf = 0; vf < c.number_of_output_frames; ++vf) computeView(c.start_frame + vf, vf, 0);
    if (m_save.valid()) m_save.get();
    return c.number_of_output_frames;
}

std::shared_ptr<InputView> Pipeline::loadInputView(int, int, Parameters const&) {
    throw std::logic_error("Pipeline::loadInputView not implemented");
}

std::unique_ptr<SynthesizedView> Pipeline::createSynthesizer(int, int) {
    if (getConfig().vs_method == "Triangles") return std::unique_ptr<SynthesizedView>(new SynthetisedViewTriangle);
    std::ostringstream m;
    m << "Unknown view synthesis method \"" << getConfig().vs_method << "\"";
    throw std::runtime_error(m.str());
}

std::unique_ptr<BlendedView> Pipeline::createBlender(int) {
    if (getConfig().blending_method == BlendingMethod::simple)
        return std::unique_ptr<BlendedView>(new BlendedViewSimple(getConfig().blending_factor));
    if (getConfig().blending_method == BlendingMethod::multispectral)
        return std::unique_ptr<BlendedView>(new BlendedViewMultiSpec(getConfig().blending_low_freq_factor,
                                                                     getConfig().blending_high_freq_factor,
                                                                     getConfig().color_space == "RGB"));
    std::ostringstream m;
    m << "Unknown view blending method \"" << getConfig().blending_method << "\"";
    throw std::runtime_error(m.str());
}

std::unique_ptr<SpaceTransformer> Pipeline::createSpaceTransformer(int) {
    return std::unique_ptr<SpaceTransformer>(new GenericTransformer);
}

void Pipeline::saveColor(std::vector<uint8_t> const&, int, int, Parameters const&) {
    throw std::logic_error("Pipeline::saveColor not implemented");
}
void Pipeline::saveMaskedColor(std::vector<uint8_t> const&, int, int, Parameters const&) {
    throw std::logic_error("Pipeline::saveMaskedColor not implemented");
}
void Pipeline::saveMask(std::vector<uint8_t> const&, int, int, Parameters const&) {
    throw std::logic_error("Pipeline::saveMask not implemented");
}
void Pipeline::saveDepth(std::vector<float> const&, int, int, Parameters const&) {
    throw std::logic_error("Pipeline::saveDepth not implemented");
}
void Pipeline::saveMaskedDepth(std::vector<float> const&, std::vector<uint8_t> const&, int, int,
                               Parameters const&) {
    throw std::logic_error("Pipeline::saveMaskedDepth not implemented");
}

namespace {

// OpenCV's INTER_NEAREST resize (imgproc/resize.cpp, resizeNN): source index
// floor(d * src / dst), clamped.
std::vector<uint8_t> resize_nearest(std::vector<uint8_t> const& src, Size s, Size d) {
    std::vector<uint8_t> out((std::size_t)d.area());
    const double fx = (double)s.width / d.width, fy = (double)s.height / d.height;
    for (int y = 0; y < d.height; ++y) {
        const int sy = std::min((int)std::floor(y * fy), s.height - 1);
        for (int x = 0; x < d.width; ++x) {
            const int sx = std::min((int)std::floor(x * fx), s.width - 1);
            out[(std::size_t)y * d.width + x] = src[(std::size_t)sy * s.width + sx];
        }
    }
    return out;
}

// OpenCV's INTER_LINEAR resize of a float map, the rule of rvs_resize_color
// (exact 2x downscale: 2x2 block means; else half-pixel-centred bilinear).
void lin_axis(int d, int n_src, int n_dst, int& s0, int& s1, float& f) {
    const double scale = (double)n_src / (double)n_dst;
    f = (float)((d + 0.5) * scale - 0.5);
    s0 = (int)std::floor(f);
    f -= (float)s0;
    if (s0 < 0) { f = 0.f; s0 = 0; }
    if (s0 >= n_src - 1) { f = 0.f; s0 = n_src - 1; }
    s1 = s0 + 1 < n_src ? s0 + 1 : s0;
}

std::vector<float> resize_linear(std::vector<float> const& src, Size s, Size d) {
    std::vector<float> out((std::size_t)d.area());
    auto at = [&](int r, int c) { return src[(std::size_t)r * s.width + c]; };
    const bool area2 = s.width == 2 * d.width && s.height == 2 * d.height;
    for (int y = 0; y < d.height; ++y)
        for (int x = 0; x < d.width; ++x) {
            float v;
            if (area2) {
                v = (((at(2 * y, 2 * x) + at(2 * y, 2 * x + 1)) + at(2 * y + 1, 2 * x)) + at(2 * y + 1, 2 * x + 1)) *
                    0.25f;
            } else {
                int x0, x1, y0, y1;
                float fx, fy;
                lin_axis(x, s.width, d.width, x0, x1, fx);
                lin_axis(y, s.height, d.height, y0, y1, fy);
                const float gx = 1.f - fx, gy = 1.f - fy;
                const float h0 = at(y0, x0) * gx + at(y0, x1) * fx, h1 = at(y1, x0) * gx + at(y1, x1) * fx;
                v = h0 * gy + h1 * fy;
            }
            out[(std::size_t)y * d.width + x] = v;
        }
    return out;
}

std::vector<uint8_t> concat_planes(DeviceBuffer const& y, DeviceBuffer const& u, DeviceBuffer const& v,
                                   std::size_t ny, std::size_t nuv, rvs_stream_t s) {
    std::vector<uint8_t> out(ny + 2 * nuv);
    y.download(out.data(), ny, s);
    u.download(out.data() + ny, nuv, s);
    v.download(out.data() + ny + nuv, nuv, s);
    return out;
}

}  // namespace

void Pipeline::saveColor(float3* color, int virtualFrame, int virtualView, Parameters const& params, Size size) {
    const Size vs = params.getSize();
    const int bd = params.getColorBitDepth();
    const std::size_t npx = (std::size_t)vs.area(), e = bytes_per(bd);
    FrameOut* f = frameOf(color);
    if (f && f->view == vs && f->raster == size && f->bitdepth == bd) {
        // planes encoded by the frame's own launch sequence
        hip_check(hipEventSynchronize((hipEvent_t)f->done), "frame wait");
        saveColor(concat_planes(f->y, f->u, f->v, npx * e, npx / 4 * e, m_save_stream), virtualFrame, virtualView,
                  params);
        return;
    }
    // a colour of the caller's: resize to the view's size, encode (write_color)
    DeviceBuffer resized, y(npx * e), u(npx / 4 * e), v(npx / 4 * e);
    const float* src = reinterpret_cast<const float*>(color);
    if (size != vs) {
        resized.resize(3 * npx * sizeof(float));
        DeviceBuffer rws(rvs_resize_color_workspace_size(size.width, size.height, vs.width, vs.height));
        check(rvs_resize_color(src, size.width, size.height, resized.as<float>(), vs.width, vs.height, rws.get(),
                               rws.bytes(), m_save_stream),
              "rvs_resize_color");
        src = resized.as<float>();
    }
    check(rvs_encode_yuv420(src, vs.width, vs.height, bd, y.get(), u.get(), v.get(), m_save_stream),
          "rvs_encode_yuv420");
    saveColor(concat_planes(y, u, v, npx * e, npx / 4 * e, m_save_stream), virtualFrame, virtualView, params);
}

void Pipeline::computeView(int inputFrame, int virtualFrame, int virtualView) {
    Config const& c = getConfig();
    detail::g_rescale = c.precision;
    Parameters params_virtual = c.params_virtual[virtualView];
    if (!c.pose_trace.empty()) {   // Pipeline.cpp:173-181
        const Pose& pose = c.pose_trace.at(inputFrame);
        Vec3f p = params_virtual.getPosition();
        for (int k = 0; k < 3; ++k) p[k] = p[k] + pose.position[k];
        params_virtual.setPosition(p);
        params_virtual.setRotation(pose.rotation);
    }
    std::unique_ptr<BlendedView> blender = createBlender(virtualView);
    std::unique_ptr<SpaceTransformer> transformer = createSpaceTransformer(virtualView);
    transformer->set_targetPosition(&params_virtual);
    const int n = (int)c.InputCameraNames.size();
    const int frame_to_load = getExtendedIndex(inputFrame, c.number_of_frames);
    // The first loaded view fixes the colour / depth scales (function-static in the reference).
    colorScale(c.params_real[0].getColorBitDepth());
    depthScale(c.params_real[0].getDepthBitDepth());
    std::vector<std::future<std::shared_ptr<InputView>>> loads;
    for (int k = 0; k < n; ++k)
        loads.push_back(std::async(std::launch::async, [this, frame_to_load, k, &c]() {
            return loadInputView(frame_to_load, k, c.params_real[k]);
        }));
    std::vector<std::shared_ptr<InputView>> inputs;
    for (auto& f : loads) inputs.push_back(f.get());
    std::vector<std::unique_ptr<SynthesizedView>> synths;
    BlendedView& blender_ref = *blender;
    bool fused = !m_stagewise && typeid(blender_ref) == typeid(BlendedViewSimple);
    for (int k = 0; k < n; ++k) {
        synths.push_back(createSynthesizer(k, virtualView));
        SynthesizedView& sref = *synths.back();
        fused = fused && typeid(sref) == typeid(SynthetisedViewTriangle);
    }
    const Size vs = params_virtual.getSize();
    const std::size_t npx = (std::size_t)vs.area();
    // rasteriser, blender and inpainting work at the oversampled size; the
    // output is resized to the view's size on write (image_writing.cpp:189-190)
    const Size rs = raster_size(vs, c.precision);
    const std::size_t nr = (std::size_t)rs.area();
    const bool rescaled = rs.width != vs.width || rs.height != vs.height;
    if (rescaled && wantMaskedColor())
        throw std::runtime_error("MaskedOutputFiles with Precision != 1: the reference masks the oversampled colour "
                                 "with the resized mask (Pipeline.cpp:307-318), which OpenCV rejects");
    const int bd = params_virtual.getColorBitDepth();
    const std::size_t e = bytes_per(bd);
    // this frame's results; the save of the frame that used them last has
    // finished (it is waited for before the next frame's save starts)
    FrameOut& fo = m_out[m_frames++ & 1];
    fo.color.resize(3 * nr * sizeof(float));
    fo.quality.resize(nr * sizeof(float));
    fo.depth.resize(nr * sizeof(float));
    fo.validity.resize(nr * sizeof(float));
    fo.y.resize(npx * e);
    fo.u.resize(npx / 4 * e);
    fo.v.resize(npx / 4 * e);
    fo.raster = rs;
    fo.view = vs;
    fo.bitdepth = bd;
    DeviceBuffer &color = fo.color, &quality = fo.quality, &depth = fo.depth, &validity = fo.validity;
    DeviceBuffer &y = fo.y, &u = fo.u, &v = fo.v;
    const rvs_cam tcam = cam_of(params_virtual);
    if (fused) {
        std::vector<rvs_source> srcs;
        for (auto& in : inputs) srcs.push_back(in->source());
        rvs_target t;
        std::memset(&t, 0, sizeof t);
        t.cam = tcam;
        for (int k = 0; k < 3; ++k) { t.pos[k] = params_virtual.getPosition()[k]; t.rot[k] = params_virtual.getRotation()[k]; }
        t.bitdepth = bd;
        t.blending_exp = static_cast<BlendedViewSimple&>(*blender).exponent();
        t.rescale = c.precision;
        t.color_scale = m_color_scale;
        t.depth_scale = m_depth_scale;
        rvs_view_out o;
        o.color = color.as<float>(); o.quality = quality.as<float>(); o.depth = depth.as<float>();
        o.validity = validity.as<float>(); o.y = y.get(); o.u = u.get(); o.v = v.get();
        for (int attempt = 0;; ++attempt) {
            const std::size_t need = rvs_synth_workspace_size(srcs.data(), n, &t, m_entry_cap);
            if (need == 0) check(RVS_E_INVALID, "rvs_synth_workspace_size");
            m_ws.resize(need);
            uint32_t needed = 0;
            const int rc = rvs_synthesize(srcs.data(), n, &t, m_ws.get(), need, m_entry_cap, &o, &needed, m_stream);
            if (rc == RVS_E_CAPACITY && needed > 0 && attempt < 8) {
                m_entry_cap = grow_entry_cap(needed);
                continue;
            }
            check(rc, "rvs_synthesize");
            break;
        }
    } else {
        for (int k = 0; k < n; ++k) {
            transformer->set_inputPosition(&c.params_real[k]);
            synths[k]->setSpaceTransformer(transformer.get());
            synths[k]->compute(*inputs[k], m_stream);
            onIntermediateSynthesisResult(inputFrame, k, virtualFrame, virtualView, *synths[k]);
            blender->blend(*synths[k], m_stream);
            onIntermediateBlendingResult(inputFrame, k, virtualFrame, virtualView, *blender);
        }
        onFinalBlendingResult(inputFrame, virtualFrame, virtualView, *blender);
        hip_check(hipMemcpyAsync(color.get(), blender->getDevColor(), 3 * nr * sizeof(float),
                                 hipMemcpyDeviceToDevice, (hipStream_t)m_stream), "copy");
        hip_check(hipMemcpyAsync(quality.get(), blender->getDevQuality(), nr * sizeof(float),
                                 hipMemcpyDeviceToDevice, (hipStream_t)m_stream), "copy");
        hip_check(hipMemcpyAsync(depth.get(), blender->getDevDepth(), nr * sizeof(float), hipMemcpyDeviceToDevice,
                                 (hipStream_t)m_stream), "copy");
        hip_check(hipMemcpyAsync(validity.get(), blender->getDevValidity(), nr * sizeof(float),
                                 hipMemcpyDeviceToDevice, (hipStream_t)m_stream), "copy");
        DeviceBuffer iws(rvs_inpaint_workspace_size(rs.width, rs.height));
        check(rvs_inpaint(color.as<float>(), quality.as<float>(), rs.width, rs.height, iws.get(), iws.bytes(),
                          nullptr, m_stream),
              "rvs_inpaint");
        if (rescaled) {
            fo.resized.resize(3 * npx * sizeof(float));
            DeviceBuffer rws(rvs_resize_color_workspace_size(rs.width, rs.height, vs.width, vs.height));
            check(rvs_resize_color(color.as<float>(), rs.width, rs.height, fo.resized.as<float>(), vs.width,
                                   vs.height, rws.get(), rws.bytes(), m_stream),
                  "rvs_resize_color");
        }
        check(rvs_encode_yuv420(rescaled ? fo.resized.as<float>() : color.as<float>(), vs.width, vs.height, bd,
                                y.get(), u.get(), v.get(), m_stream),
              "rvs_encode_yuv420");
    }
    hip_check(hipEventRecord((hipEvent_t)fo.done, (hipStream_t)m_stream), "hipEventRecord");
    inputs.clear();
    if (m_save.valid()) m_save.get();
    if (wantColor()) {
        // the reference's GPU path hands the inpainted device colour to saveColor
        // on a worker thread (Pipeline.cpp:284-296)
        float3* dev = reinterpret_cast<float3*>(color.get());
        void* done = fo.done;
        m_save = std::async(std::launch::async, [this, dev, done, virtualFrame, virtualView, params_virtual, rs]() {
            hip_check(hipSetDevice(m_device), "hipSetDevice");
            hip_check(hipEventSynchronize((hipEvent_t)done), "frame wait");
            saveColor(dev, virtualFrame, virtualView, params_virtual, rs);
        });
    }
    // masks and depth maps (Pipeline.cpp:303-336)
    if (wantMask() || wantMaskedColor() || wantMaskedDepth() || wantDepth()) {
        std::vector<float> val(nr), dep(nr);
        validity.download(val.data(), nr * sizeof(float), m_stream);
        depth.download(dep.data(), nr * sizeof(float), m_stream);
        std::vector<uint8_t> rmask(nr);
        for (std::size_t k = 0; k < nr; ++k) rmask[k] = (val[k] > c.validity_threshold) ? 0 : 255;
        // mask: resize(INTER_NEAREST); depth: resize() (INTER_LINEAR) to the view's size
        std::vector<uint8_t> mask = rescaled ? resize_nearest(rmask, rs, vs) : rmask;
        if (rescaled) dep = resize_linear(dep, rs, vs);
        if (wantMask()) saveMask(mask, virtualFrame, virtualView, params_virtual);
        if (wantMaskedColor()) {
            std::vector<float> col(3 * npx);
            color.download(col.data(), col.size() * sizeof(float), m_stream);
            for (std::size_t k = 0; k < npx; ++k)
                if (mask[k]) col[3 * k] = col[3 * k + 1] = col[3 * k + 2] = 0.5f;
            DeviceBuffer mc;
            mc.upload(col.data(), col.size() * sizeof(float), m_stream);
            DeviceBuffer my(npx * e), mu(npx / 4 * e), mv(npx / 4 * e);
            check(rvs_encode_yuv420(mc.as<float>(), vs.width, vs.height, bd, my.get(), mu.get(), mv.get(), m_stream),
                  "rvs_encode_yuv420");
            saveMaskedColor(concat_planes(my, mu, mv, npx * e, npx / 4 * e, m_stream), virtualFrame, virtualView,
                            params_virtual);
        }
        if (wantDepth()) saveDepth(dep, virtualFrame, virtualView, params_virtual);
        if (wantMaskedDepth()) saveMaskedDepth(dep, mask, virtualFrame, virtualView, params_virtual);
    }
}

// ------------------------------------------------------------ Application

Application::Application(std::string const& config_path, std::string const& source_path)
    : m_config(Config::loadFromFile(config_path)) {
    if (!source_path.empty()) {
        for (auto& n : m_config.texture_names) n = source_path + "/" + n;
        for (auto& n : m_config.depth_names) n = source_path + "/" + n;
    }
}

namespace {

// write_depth / write_maskedDepth (image_writing.cpp:227-279): metric depth back to
// the file's inverse-depth code, YUV420 files get neutral chroma.
std::vector<uint8_t> encode_depth(std::vector<float> const& depth, std::vector<uint8_t> const* mask,
                                  Parameters const& p) {
    const int bd = p.getDepthBitDepth();
    const Size s = p.getSize();
    const std::size_t n = (std::size_t)s.area();
    const float zn = p.getDepthRange()[0], zf = p.getDepthRange()[1];
    const unsigned mx = max_level(bd);
    const int e = bytes_per(bd);
    const unsigned type_max = e == 1 ? 255u : 65535u;
    std::vector<uint8_t> out(n * e);
    for (std::size_t k = 0; k < n; ++k) {
        float v;
        if (zf >= 1000.f) v = zn / depth[k];
        else v = ((float)((double)(zf * zn) / depth[k]) - zn) * (float)(1.0 / (zf - zn));
        if (mask && (*mask)[k]) v = 0.f;
        const float sv = v * (float)mx;
        long r = std::isnan(sv) ? 0 : std::lrint(sv);
        r = r < 0 ? 0 : (r > (long)type_max ? (long)type_max : r);
        if (e == 1) out[k] = (uint8_t)r;
        else { out[2 * k] = (uint8_t)(r & 0xff); out[2 * k + 1] = (uint8_t)(r >> 8); }
    }
    if (p.getDepthColorFormat() == ColorFormat::YUV420) {
        const unsigned neutral = (unsigned)(0.5 * (1 + mx));
        const std::size_t nc = (std::size_t)(s.width / 2) * (s.height / 2);
        for (std::size_t k = 0; k < 2 * nc; ++k) {
            if (e == 1) out.push_back((uint8_t)neutral);
            else { out.push_back((uint8_t)(neutral & 0xff)); out.push_back((uint8_t)(neutral >> 8)); }
        }
    }
    return out;
}

}  // namespace

std::shared_ptr<InputView> Application::loadInputView(int inputFrame, int inputView, Parameters const& parameters) {
    return std::make_shared<InputView>(m_config.texture_names[inputView], m_config.depth_names[inputView], inputFrame,
                                       parameters, colorScale(parameters.getColorBitDepth()),
                                       depthScale(parameters.getDepthBitDepth()), stream());
}

void Application::saveColor(std::vector<uint8_t> const& yuv, int vf, int vv, Parameters const&) {
    write_frame(m_config.outfilenames[vv], vf, yuv);
}

void Application::saveMaskedColor(std::vector<uint8_t> const& yuv, int vf, int vv, Parameters const&) {
    write_frame(m_config.outmaskedfilenames[vv], vf, yuv);
}

void Application::saveMask(std::vector<uint8_t> const& mask, int vf, int vv, Parameters const& p) {
    std::vector<uint8_t> out(mask);
    for (auto& m : out) m = m ? 255 : 0;
    out.resize(mask.size() + 2 * (std::size_t)(p.getSize().width / 2) * (p.getSize().height / 2), 128);
    write_frame(m_config.outmaskfilenames[vv], vf, out);
}

void Application::saveDepth(std::vector<float> const& depth, int vf, int vv, Parameters const& p) {
    write_frame(m_config.outdepthfilenames[vv], vf, encode_depth(depth, nullptr, p));
}

void Application::saveMaskedDepth(std::vector<float> const& depth, std::vector<uint8_t> const& mask, int vf, int vv,
                                  Parameters const& p) {
    write_frame(m_config.outmaskdepthfilenames[vv], vf, encode_depth(depth, &mask, p));
}

}  // namespace rvs
