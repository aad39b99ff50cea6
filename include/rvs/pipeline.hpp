// pipeline.hpp -- drop-in C++ call surface of the reference view synthesizer's
// per-frame path (rvs/src/Pipeline.hpp, View.hpp, SynthesizedView.hpp,
// BlendedView.hpp, SpaceTransformer.hpp, Application.hpp), running on gfx950
// through the C ABI of rvs_hip.h.
//
// Plugin points are the reference's: Pipeline's virtual factories
// createSynthesizer / createBlender / createSpaceTransformer, loadInputView and
// the want* / save* hooks.  When every factory returns the built-in classes,
// computeView runs the fused device path (rvs_synthesize: all input views of
// the frame rasterised and blended in one launch sequence); otherwise it calls
// the plugins view by view exactly like the reference.  Both produce the same
// bits.  Views hold their maps in device memory.
#ifndef RVS_PIPELINE_HPP
#define RVS_PIPELINE_HPP

#include <cstdint>
#include <future>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_vector_types.h>

#include "rvs/config.hpp"
#include "rvs_hip.h"

namespace rvs {

/** Owned device allocation (hipMalloc / hipFree). */
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept;
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* get() const { return m_ptr; }
    template <class T> T* as() const { return static_cast<T*>(m_ptr); }
    std::size_t bytes() const { return m_bytes; }
    void resize(std::size_t bytes);   // reallocates only when growing
    void upload(const void* host, std::size_t bytes, rvs_stream_t s);
    void download(void* host, std::size_t bytes, rvs_stream_t s) const;

private:
    void* m_ptr = nullptr;
    std::size_t m_bytes = 0;
};

/** Device-resident maps of one view (View.hpp:68-145). */
class View {
public:
    virtual ~View() = default;
    Size get_size() const { return m_size; }
    float* getDevColor() const { return m_color.as<float>(); }        // float3 Y,U,V
    float* getDevDepth() const { return m_depth.as<float>(); }
    float* getDevQuality() const { return m_quality.as<float>(); }
    float* getDevValidity() const { return m_validity.as<float>(); }  // triangle shape for synthesized views
    /** Host copies of the maps. */
    std::vector<float> get_color() const;
    std::vector<float> get_depth() const;
    std::vector<float> get_quality() const;
    std::vector<float> get_validity() const;

protected:
    void allocate(Size s, bool quality);
    Size m_size;
    DeviceBuffer m_color, m_depth, m_quality, m_validity;
};

/** An input view: raw planes read from the YUV files, kept resident, plus the
 * decoded colour and depth (InputView, View.hpp:156-174; image_loading.cpp). */
class InputView : public View {
public:
    InputView(std::string const& color_path, std::string const& depth_path, int frame, Parameters const& params,
              float color_scale, float depth_scale, rvs_stream_t stream);
    Parameters const& parameters() const { return m_params; }
    rvs_source source() const;       // for rvs_synthesize
    void decode(rvs_stream_t stream); // colour/depth maps (stage API path)
    bool decoded() const { return m_decoded; }

private:
    Parameters m_params;
    float m_color_scale, m_depth_scale;
    DeviceBuffer m_y, m_u, m_v, m_raw_depth, m_table;
    DeviceBuffer m_prep;              // resident culling table (rvs_source_prep)
    float m_prep_scale = 0.f;
    bool m_decoded = false;
};

/** Relative pose between an input and the virtual camera (SpaceTransformer.hpp). */
class SpaceTransformer {
public:
    virtual ~SpaceTransformer() = default;
    void set_inputPosition(Parameters const* p) { m_input = p; }
    void set_targetPosition(Parameters const* p) { m_target = p; }
    Parameters const& getInputParameters() const { return *m_input; }
    Parameters const& getVirtualParameters() const { return *m_target; }
    /** R_to^T * R_from (float, SpaceTransformer.cpp:86-92). */
    std::array<float, 9> get_rotation() const;
    /** -R_to^T * (t_to - t_from) (SpaceTransformer.cpp:77-84). */
    Vec3f get_translation() const;

private:
    Parameters const* m_input = nullptr;
    Parameters const* m_target = nullptr;
};

class GenericTransformer : public SpaceTransformer {};

/** A view synthesized from one input view (SynthesizedView.hpp). */
class SynthesizedView : public View {
public:
    void setSpaceTransformer(SpaceTransformer const* t) { m_transformer = t; }
    virtual void compute(InputView& input, rvs_stream_t stream) = 0;

protected:
    SpaceTransformer const* m_transformer = nullptr;
};

/** Triangle method (SynthetisedViewTriangle): reproject + rasterise. Maps:
 * color, depth (+inf uncovered), validity = triangle shape. */
class SynthetisedViewTriangle : public SynthesizedView {
public:
    void compute(InputView& input, rvs_stream_t stream) override;

private:
    DeviceBuffer m_uv, m_vdepth, m_ws;
};

/** Accumulator of synthesized views (BlendedView.hpp).  blend() reads the
 * view's colour, depth and validity (the triangle shape of a synthesized
 * view). */
class BlendedView : public View {
public:
    virtual void blend(View const& view, rvs_stream_t stream) = 0;
};

/** Weighted blending (BlendedViewSimple, BlendedView.cpp:180-212). */
class BlendedViewSimple : public BlendedView {
public:
    explicit BlendedViewSimple(float blending_exp) : m_exp(blending_exp) {}
    void blend(View const& view, rvs_stream_t stream) override;
    /** The same on device maps of size s (float3 colour, depth, shape). */
    void blend_maps(const float* color, const float* depth, const float* shape, Size s, rvs_stream_t stream);
    float exponent() const { return m_exp; }

private:
    float m_exp;
    bool m_empty = true;
    DeviceBuffer m_mask;
};

/** Multispectral blending (BlendedViewMultiSpec, BlendedView.cpp:62-88): every
 * view is split into a low and a high frequency band (rvs_split_frequencies;
 * ColorSpace RGB blurs all three channels, YUV only Y), each band blended with
 * its own exponent; the result is the sum of the bands' colours with the low
 * band's depth, quality and validity. */
class BlendedViewMultiSpec : public BlendedView {
public:
    BlendedViewMultiSpec(float exp_low_freq, float exp_high_freq, bool rgb = false)
        : m_low_freq(exp_low_freq), m_high_freq(exp_high_freq), m_rgb(rgb) {}
    void blend(View const& view, rvs_stream_t stream) override;
    BlendedViewSimple const& low_band() const { return m_low_freq; }
    BlendedViewSimple const& high_band() const { return m_high_freq; }

private:
    BlendedViewSimple m_low_freq, m_high_freq;
    bool m_rgb;
    bool m_empty = true;
    DeviceBuffer m_ws, m_low, m_high;
};

/** The frame loop (Pipeline.hpp / Pipeline.cpp). */
class Pipeline {
public:
    Pipeline();
    virtual ~Pipeline();
    /** Process every output frame; returns their number. */
    int execute();
    /** Force the view-by-view plugin path even with the built-in classes. */
    void setStagewise(bool on) { m_stagewise = on; }

protected:
    virtual Config const& getConfig() const = 0;
    virtual std::shared_ptr<InputView> loadInputView(int inputFrame, int inputView, Parameters const& parameters);
    virtual std::unique_ptr<SynthesizedView> createSynthesizer(int inputView, int virtualView);
    virtual std::unique_ptr<BlendedView> createBlender(int virtualView);
    virtual std::unique_ptr<SpaceTransformer> createSpaceTransformer(int virtualView);
    virtual bool wantColor() { return false; }
    virtual bool wantMaskedColor() { return false; }
    virtual bool wantMask() { return false; }
    virtual bool wantDepth() { return false; }
    virtual bool wantMaskedDepth() { return false; }
    /** The inpainted colour of the GPU path (Pipeline.hpp:134; Pipeline.cpp:284-296):
     * device float3 (Y, U, V) per pixel at `size`, the raster size (the view's
     * size times Precision).  Called on a worker thread once the frame's work
     * on stream() has completed; the buffer is left untouched until the call
     * returns, and the next call waits for this one.  The default resizes it to
     * the view's size, encodes it (write_color, image_writing.cpp:142-203) and
     * calls saveColor(yuv, ...); for the colour computeView hands over it reuses
     * the planes the frame's launch sequence already encoded. */
    virtual void saveColor(float3* color, int virtualFrame, int virtualView, Parameters const& params, Size size);
    /** Encoded YUV420 planes of the inpainted view (Y, then U, V at half size). */
    virtual void saveColor(std::vector<uint8_t> const& yuv, int virtualFrame, int virtualView,
                           Parameters const& params);
    virtual void saveMaskedColor(std::vector<uint8_t> const& yuv, int virtualFrame, int virtualView,
                                 Parameters const& params);
    virtual void saveMask(std::vector<uint8_t> const& mask, int virtualFrame, int virtualView,
                          Parameters const& params);
    virtual void saveDepth(std::vector<float> const& depth, int virtualFrame, int virtualView,
                           Parameters const& params);
    virtual void saveMaskedDepth(std::vector<float> const& depth, std::vector<uint8_t> const& mask,
                                 int virtualFrame, int virtualView, Parameters const& params);
    /** Per-view results of the view-by-view path (Pipeline.hpp:161-177): each
     * input view once synthesized, the blender after each blend, the blender
     * after the last one (before inpainting).  Called synchronously from
     * computeView; the views' device maps are complete once stream() is
     * synchronised.  The fused path (the built-in classes without
     * setStagewise(true)) forms no per-view results and calls none of them;
     * the reference declares them without calling them. */
    virtual void onIntermediateSynthesisResult(int inputFrame, int inputView, int virtualFrame, int virtualView,
                                               SynthesizedView const& synthesizedView) {}
    virtual void onIntermediateBlendingResult(int inputFrame, int inputView, int virtualFrame, int virtualView,
                                              BlendedView const& blendedView) {}
    virtual void onFinalBlendingResult(int inputFrame, int virtualFrame, int virtualView,
                                       BlendedView const& blendedView) {}
    /** fl(1/max_level) of the first view loaded by this pipeline
     * (function-static in read_color_YUV / read_depth_YUV, image_loading.cpp:134,193). */
    float colorScale(int bitdepth);
    float depthScale(int bitdepth);
    rvs_stream_t stream() const { return m_stream; }

private:
    /** Device results of one output frame; two alternate so that a frame can
     * be computed while the previous one is being saved. */
    struct FrameOut {
        DeviceBuffer color, quality, depth, validity, y, u, v, resized;
        Size raster{0, 0}, view{0, 0};
        int bitdepth = 0;
        void* done = nullptr;   // hipEvent_t recorded on stream() after the frame's last launch
    };
    void computeView(int inputFrame, int virtualFrame, int virtualView);
    FrameOut* frameOf(const float3* color);
    rvs_stream_t m_stream = nullptr;
    rvs_stream_t m_save_stream = nullptr;   // downloads / encodes of the save thread
    int m_device = 0;                       // the device current at construction
    FrameOut m_out[2];
    unsigned m_frames = 0;
    bool m_stagewise = false;
    float m_color_scale = -1.f, m_depth_scale = -1.f;
    DeviceBuffer m_ws;
    uint32_t m_entry_cap = 0;   // binned-entry list capacity of the fused path (0: default)
    std::future<void> m_save;
};

/** Configuration-file driven pipeline (Application.hpp). */
class Application : public Pipeline {
public:
    explicit Application(std::string const& config_path, std::string const& source_path = "");
    Config const& getConfig() const override { return m_config; }

protected:
    /** InputView of texture_names[k] / depth_names[k] (Application.cpp:72-81). */
    std::shared_ptr<InputView> loadInputView(int inputFrame, int inputView, Parameters const& parameters) override;
    bool wantColor() override { return !m_config.outfilenames.empty(); }
    bool wantMaskedColor() override { return !m_config.outmaskedfilenames.empty(); }
    bool wantMask() override { return !m_config.outmaskfilenames.empty(); }
    bool wantDepth() override { return !m_config.outdepthfilenames.empty(); }
    bool wantMaskedDepth() override { return !m_config.outmaskdepthfilenames.empty(); }
    using Pipeline::saveColor;
    void saveColor(std::vector<uint8_t> const& yuv, int vf, int vv, Parameters const& p) override;
    void saveMaskedColor(std::vector<uint8_t> const& yuv, int vf, int vv, Parameters const& p) override;
    void saveMask(std::vector<uint8_t> const& mask, int vf, int vv, Parameters const& p) override;
    void saveDepth(std::vector<float> const& depth, int vf, int vv, Parameters const& p) override;
    void saveMaskedDepth(std::vector<float> const& depth, std::vector<uint8_t> const& mask, int vf, int vv,
                         Parameters const& p) override;

private:
    Config m_config;
};

/** Pipeline.cpp:157-167: output frame index -> input frame index, going back
 * and forth over the input frames. */
int getExtendedIndex(int outputFrameIndex, int numberOfInputFrames);

// ---- raw YUV files (image_loading.cpp:74-213, image_writing.cpp:72-139)
/** Planes of frame `frame` of a planar YUV420 file (wraps to frame 0 past the end). */
void read_yuv420(std::string const& path, int frame, Size size, int bitdepth, std::vector<uint8_t>& y,
                 std::vector<uint8_t>& u, std::vector<uint8_t>& v);
/** Y plane of frame `frame` of a depth file (YUV420 or YUV400 layout). */
void read_depth_plane(std::string const& path, int frame, Size size, int bitdepth, ColorFormat fmt,
                      std::vector<uint8_t>& y);
/** Append (frame > 0) or create (frame == 0) a file with the given bytes. */
void write_frame(std::string const& path, int frame, std::vector<uint8_t> const& bytes);

}  // namespace rvs

#endif
