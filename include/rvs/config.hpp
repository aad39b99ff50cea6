// config.hpp -- camera parameters, view-synthesis configuration and pose
// traces, with the file formats and validation rules of the reference view
// synthesizer (rvs/src/Parameters.*, Config.*, PoseTraces.*).
#ifndef RVS_CONFIG_HPP
#define RVS_CONFIG_HPP

#include <array>
#include <istream>
#include <string>
#include <vector>

#include "rvs/json.hpp"

namespace rvs {

using Vec2f = std::array<float, 2>;
using Vec3f = std::array<float, 3>;
struct Size { int width = 0, height = 0; long area() const { return (long)width * height; } };
inline bool operator==(Size a, Size b) { return a.width == b.width && a.height == b.height; }
inline bool operator!=(Size a, Size b) { return !(a == b); }
struct Rect { int x = 0, y = 0, width = 0, height = 0; Size size() const { return {width, height}; } };

namespace ProjectionType {
constexpr const char* perspective = "Perspective";
constexpr const char* equirectangular = "Equirectangular";
}  // namespace ProjectionType

enum class ColorFormat { YUV420, YUV400 };
enum class DisplacementMethod { depth, polynomial };

/** One camera of a camera-parameter file (Parameters.hpp:79-160). */
class Parameters {
public:
    static Parameters readFrom(json::Node root);
    json::Node const& getRoot() const { return m_root; }
    std::string const& getProjectionType() const { return m_projectionType; }
    Vec3f getRotation() const { return m_rotation; }
    void setRotation(Vec3f r) { m_rotation = r; }
    /** Rz(yaw) Ry(pitch) Rx(roll) in float, row-major (Parameters.cpp:67-97,151-155). */
    std::array<float, 9> getRotationMatrix() const;
    Vec3f getPosition() const { return m_position; }
    void setPosition(Vec3f p) { m_position = p; }
    Vec2f getDepthRange() const { return m_depthRange; }
    Vec2f getMultiDepthRange() const { return m_multiDepthRange; }
    bool hasInvalidDepth() const { return m_hasInvalidDepth; }
    Size getPaddedSize() const { return m_resolution; }
    Size getSize() const { return m_cropRegion.size(); }
    Rect getCropRegion() const { return m_cropRegion; }
    int getColorBitDepth() const { return m_bitDepthColor; }
    int getDepthBitDepth() const { return m_bitDepthDepth; }
    ColorFormat getColorFormat() const { return m_colorFormat; }
    ColorFormat getDepthColorFormat() const { return m_depthColorFormat; }
    Vec2f getHorRange() const { return m_horRange; }
    Vec2f getVerRange() const { return m_verRange; }
    bool isFullHorRange() const { return m_isFullHorRange; }
    Vec2f getFocal() const { return m_focal; }
    void setFocal(Vec2f f) { m_focal = f; }
    /** Principal point relative to the crop region. */
    Vec2f getPrinciplePoint() const;
    void setPrinciplePoint(Vec2f p) { m_principlePoint = p; }
    DisplacementMethod getDisplacementMethod() const { return m_displacementMethod; }

private:
    explicit Parameters(json::Node root) : m_root(root) {}
    json::Node m_root;
    std::string m_projectionType;
    Vec3f m_position{}, m_rotation{};
    Vec2f m_depthRange{}, m_multiDepthRange{};
    bool m_hasInvalidDepth = true;
    Size m_resolution;
    Rect m_cropRegion;
    int m_bitDepthColor = 0, m_bitDepthDepth = 0;
    ColorFormat m_colorFormat = ColorFormat::YUV420, m_depthColorFormat = ColorFormat::YUV420;
    Vec2f m_horRange{}, m_verRange{};
    bool m_isFullHorRange = false;
    Vec2f m_focal{}, m_principlePoint{};
    DisplacementMethod m_displacementMethod = DisplacementMethod::depth;
};

struct Pose {
    Vec3f position;
    Vec3f rotation;
};

/** Pose trace CSV (PoseTraces.cpp:54-92). */
class PoseTrace : public std::vector<Pose> {
public:
    static PoseTrace loadFrom(std::istream& stream);
    static PoseTrace loadFromFile(std::string const& filename);
};

namespace BlendingMethod {
constexpr const char* simple = "Simple";
constexpr const char* multispectral = "Multispectral";
}  // namespace BlendingMethod

/** View-synthesis configuration file (Config.hpp:118-180, Config.cpp:72-400). */
class Config {
public:
    static Config loadFromFile(std::string const& filename);
    std::string version;
    std::vector<std::string> InputCameraNames, VirtualCameraNames;
    std::vector<Parameters> params_real, params_virtual;
    std::vector<std::string> texture_names, depth_names;
    std::vector<std::string> outfilenames, outmaskedfilenames, outmaskfilenames, outdepthfilenames,
        outmaskdepthfilenames;
    float validity_threshold = 5000.f;
    std::string vs_method = "Triangles";
    std::string blending_method = "Simple";
    float blending_low_freq_factor = 0.f, blending_high_freq_factor = 0.f;
    float blending_factor = 5.f;
    int start_frame = 0;
    int number_of_frames = 1;
    int number_of_output_frames = 1;
    float precision = 1.f;           /* Precision (detail::g_rescale) */
    std::string color_space = "YUV"; /* ColorSpace (detail::g_color_space) */
    PoseTrace pose_trace;

private:
    Config() = default;
};

namespace detail {
// Precision of the last configuration loaded (Config.cpp:65,393): the
// rasteriser's oversampling factor, read by SynthetisedViewTriangle.
extern float g_rescale;
}  // namespace detail

}  // namespace rvs

#endif
