// A plugin of the drop-in Pipeline: an Application subclass that overrides the
// reference's float-colour saveColor (Pipeline.hpp:134) and the per-view hooks
// (Pipeline.hpp:161-177), records what they receive and still writes the
// configured output files through the default implementations.
//
//   plugin_check <config.json> <out_prefix> [--stagewise]
//
// Writes <out_prefix>_color<f>.bin (the device colour handed to saveColor,
// raster size, float3) and, on the view-by-view path, <out_prefix>_final<f>.bin
// (colour, quality of the blender at onFinalBlendingResult); prints one JSON
// line with the hook counts and sizes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <string>
#include <vector>

#include "rvs/pipeline.hpp"

namespace {

class Plugin : public rvs::Application {
public:
    Plugin(std::string const& cfg, std::string const& prefix) : rvs::Application(cfg), m_prefix(prefix) {}
    int synth_calls = 0, blend_calls = 0, final_calls = 0, color_calls = 0;
    int last_synth_w = 0, color_w = 0, color_h = 0;
    std::string views;   // "inputFrame:inputView" of each synthesis result

protected:
    using rvs::Application::saveColor;
    void saveColor(float3* color, int vf, int vv, rvs::Parameters const& p, rvs::Size size) override {
        std::vector<float> host(3 * (std::size_t)size.area());
        if (hipMemcpy(host.data(), color, host.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("download of the saved colour failed");
        write(m_prefix + "_color" + std::to_string(vf) + ".bin", host);
        {
            std::lock_guard<std::mutex> g(m_mu);
            ++color_calls;
            color_w = size.width;
            color_h = size.height;
        }
        rvs::Application::saveColor(color, vf, vv, p, size);   // the default: encode and write the file
    }
    void onIntermediateSynthesisResult(int inputFrame, int inputView, int, int,
                                       rvs::SynthesizedView const& v) override {
        ++synth_calls;
        last_synth_w = v.get_size().width;
        views += (views.empty() ? "" : ",") + std::to_string(inputFrame) + ":" + std::to_string(inputView);
    }
    void onIntermediateBlendingResult(int, int, int, int, rvs::BlendedView const&) override { ++blend_calls; }
    void onFinalBlendingResult(int, int vf, int, rvs::BlendedView const& b) override {
        ++final_calls;
        std::vector<float> c = b.get_color(), q = b.get_quality();
        c.insert(c.end(), q.begin(), q.end());
        write(m_prefix + "_final" + std::to_string(vf) + ".bin", c);
    }

private:
    static void write(std::string const& path, std::vector<float> const& v) {
        std::ofstream f(path, std::ios::binary);
        f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
        if (!f) throw std::runtime_error("cannot write " + path);
    }
    std::string m_prefix;
    std::mutex m_mu;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s config.json out_prefix [--stagewise]\n", argv[0]);
        return 2;
    }
    try {
        Plugin p(argv[1], argv[2]);
        if (argc > 3 && std::strcmp(argv[3], "--stagewise") == 0) p.setStagewise(true);
        const int frames = p.execute();
        std::printf("{\"frames\": %d, \"synth\": %d, \"blend\": %d, \"final\": %d, \"color\": %d, "
                    "\"color_w\": %d, \"color_h\": %d, \"synth_w\": %d, \"views\": \"%s\"}\n",
                    frames, p.synth_calls, p.blend_calls, p.final_calls, p.color_calls, p.color_w, p.color_h,
                    p.last_synth_w, p.views.c_str());
    } catch (std::exception const& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
