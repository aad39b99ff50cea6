// config.cpp -- Parameters, Config and PoseTrace with the reference's file
// formats and validation rules (rvs/src/Parameters.cpp, Config.cpp,
// PoseTraces.cpp).  Paths inside a configuration are opened as given (relative
// to the working directory), as the reference does.
#include "rvs/config.hpp"

#include <cmath>
#include <fstream>
#include <map>
#include <regex>
#include <sstream>
#include <stdexcept>

namespace rvs {

namespace detail {
float g_rescale = 1.f;
}  // namespace detail


namespace {

template <int N>
std::array<float, N> asFloatVec(json::Node node) {
    if (node.size() != N) throw std::runtime_error("JSON parser: Expected a vector of floats");
    std::array<float, N> v{};
    for (int i = 0; i != N; ++i) v[i] = node.at(i).asFloat();
    return v;
}

template <int N>
std::array<int, N> asIntVec(json::Node node) {
    if (node.size() != N) throw std::runtime_error("JSON parser: Expected a vector of floats");
    std::array<int, N> v{};
    for (int i = 0; i != N; ++i) v[i] = node.at(i).asInt();
    return v;
}

void matmul33(const float* A, const float* B, float* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s = 0.f;
            for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * B[k * 3 + j];
            C[i * 3 + j] = s;
        }
}

}  // namespace

// ------------------------------------------------------------- Parameters

Parameters Parameters::readFrom(json::Node root) {
    Parameters P(root);
    P.m_projectionType = root.require("Projection").asString();
    P.m_position = asFloatVec<3>(root.require("Position"));
    P.m_rotation = asFloatVec<3>(root.require("Rotation"));
    P.m_depthRange = asFloatVec<2>(root.require("Depth_range"));
    if (!(P.m_depthRange[0] < P.m_depthRange[1]))
        throw std::runtime_error("Inverted depth range: [near, far] with near > far is not allowed");
    if (P.m_depthRange[1] > 1000.f)
        throw std::runtime_error("Invalid depth range: [near, far] with far > 1000 is not allowed because 1000 "
                                 "stands for infinity. Please restrict depth range or change world units.");
    P.m_hasInvalidDepth = true;
    if (auto n = root.optional("HasInvalidDepth")) P.m_hasInvalidDepth = n.asBool();
    const auto res = asIntVec<2>(root.require("Resolution"));
    P.m_resolution = {res[0], res[1]};
    P.m_bitDepthColor = root.require("BitDepthColor").asInt();
    P.m_bitDepthDepth = root.require("BitDepthDepth").asInt();
    const std::string cf = root.require("ColorSpace").asString();
    if (cf != "YUV420") throw std::runtime_error("This version of RVS only supports YUV420 color space for texture");
    P.m_colorFormat = ColorFormat::YUV420;
    const std::string dcf = root.require("DepthColorSpace").asString();
    if (dcf == "YUV420") P.m_depthColorFormat = ColorFormat::YUV420;
    else if (dcf == "YUV400") P.m_depthColorFormat = ColorFormat::YUV400;
    else throw std::runtime_error("This version of RVS only supports YUV420 and YUV400 color space for depth");
    if (P.m_projectionType == ProjectionType::equirectangular) {
        P.m_horRange = asFloatVec<2>(root.require("Hor_range"));
        P.m_isFullHorRange = false;
        try {
            const auto hi = asIntVec<2>(root.require("Hor_range"));
            P.m_isFullHorRange = hi[0] == -180 && hi[1] == 180;
        } catch (std::runtime_error&) {
        }
        P.m_verRange = asFloatVec<2>(root.require("Ver_range"));
    }
    if (auto n = root.optional("Crop_region")) {
        const auto v = asIntVec<4>(n);
        P.m_cropRegion = {v[0], v[1], v[2], v[3]};
    } else {
        P.m_cropRegion = {0, 0, P.m_resolution.width, P.m_resolution.height};
    }
    if (P.m_projectionType == ProjectionType::perspective) {
        P.m_focal = asFloatVec<2>(root.require("Focal"));
        P.m_principlePoint = asFloatVec<2>(root.require("Principle_point"));
    }
    if (auto n = root.optional("DisplacementMethod")) {
        if (n.asString() == "Depth") P.m_displacementMethod = DisplacementMethod::depth;
        if (n.asString() == "Polynomial") P.m_displacementMethod = DisplacementMethod::polynomial;
    } else {
        P.m_displacementMethod = DisplacementMethod::depth;
    }
    if (auto n = root.optional("Multi_depth_range")) P.m_multiDepthRange = asFloatVec<2>(n);
    else P.m_multiDepthRange = P.m_depthRange;
    if (!(P.m_multiDepthRange[0] < P.m_multiDepthRange[1]))
        throw std::runtime_error("Inverted depth range: [near, far] with near > far is not allowed");
    // validateUnused (Parameters.cpp:280-289)
    if (root.require("Depthmap").asInt() != 1) throw std::runtime_error("This version of RVS only supports Depthmap 1");
    root.require("Background").asInt();
    return P;
}

std::array<float, 9> Parameters::getRotationMatrix() const {
    const float radperdeg = 0.01745329252f;
    const float rz = radperdeg * m_rotation[0], ry = radperdeg * m_rotation[1], rx = radperdeg * m_rotation[2];
    const float Rz[9] = {std::cos(rz), -std::sin(rz), 0.f, std::sin(rz), std::cos(rz), 0.f, 0.f, 0.f, 1.f};
    const float Ry[9] = {std::cos(ry), 0.f, std::sin(ry), 0.f, 1.f, 0.f, -std::sin(ry), 0.f, std::cos(ry)};
    const float Rx[9] = {1.f, 0.f, 0.f, 0.f, std::cos(rx), -std::sin(rx), 0.f, std::sin(rx), std::cos(rx)};
    float T[9];
    std::array<float, 9> R{};
    matmul33(Rz, Ry, T);
    matmul33(T, Rx, R.data());
    return R;
}

Vec2f Parameters::getPrinciplePoint() const {
    return {m_principlePoint[0] - (float)m_cropRegion.x, m_principlePoint[1] - (float)m_cropRegion.y};
}

// -------------------------------------------------------------- PoseTrace

PoseTrace PoseTrace::loadFrom(std::istream& stream) {
    std::string line;
    std::getline(stream, line);
    static const std::regex header("\\s*X\\s*,\\s*Y\\s*,\\s*Z\\s*,\\s*Yaw\\s*,\\s*Pitch\\s*,\\s*Roll\\s*");
    if (!std::regex_match(line, header)) throw std::runtime_error("Format error in the pose trace header");
    static const std::regex row("([^,]+),([^,]+),([^,]+),([^,]+),([^,]+),([^,]+)");
    static const std::regex blank("\\s*");
    PoseTrace trace;
    bool trailing = false;
    while (std::getline(stream, line)) {
        std::smatch m;
        if (!trailing && std::regex_match(line, m, row)) {
            Pose p;
            for (int k = 0; k < 3; ++k) p.position[k] = std::stof(m[1 + k].str());
            for (int k = 0; k < 3; ++k) p.rotation[k] = std::stof(m[4 + k].str());
            trace.push_back(p);
        } else if (std::regex_match(line, blank)) {
            trailing = true;
        } else {
            throw std::runtime_error("Format error in a pose trace row");
        }
    }
    return trace;
}

PoseTrace PoseTrace::loadFromFile(std::string const& filename) {
    std::ifstream s(filename);
    if (!s.good()) throw std::runtime_error("Failed to load pose trace");
    return loadFrom(s);
}

// ----------------------------------------------------------------- Config

namespace {

std::vector<Parameters> loadCameras(std::string const& filepath, std::vector<std::string> const& names,
                                    json::Node overrides) {
    std::ifstream stream(filepath);
    if (!stream.good()) {
        std::ostringstream m;
        m << "Failed to open camera parameters file \"" << filepath << "\" for reading";
        throw std::runtime_error(m.str());
    }
    auto root = json::Node::readFrom(stream);
    const std::string v = root.require("Version").asString();
    if (v.substr(0, 2) != "2." && v.substr(0, 2) != "3." && v.substr(0, 2) != "4.")
        throw std::runtime_error("Version of the camera parameters file is not compatible with this version of RVS");
    std::map<std::string, Parameters> index;
    auto cams = root.require("cameras");
    for (std::size_t i = 0; i != cams.size(); ++i) {
        auto node = cams.at(i);
        const std::string name = node.require("Name").asString();
        if (index.count(name)) {
            std::ostringstream m;
            m << "Camera parameters file has duplicate camera '" << name << "'";
            throw std::runtime_error(m.str());
        }
        if (overrides) node.setOverrides(overrides);
        index.emplace(name, Parameters::readFrom(node));
    }
    std::vector<Parameters> out;
    for (auto const& n : names) {
        auto it = index.find(n);
        if (it == index.end()) {
            std::ostringstream m;
            m << "Camera parameters file does not have camera '" << n << "'";
            throw std::runtime_error(m.str());
        }
        out.push_back(it->second);
    }
    return out;
}

void inputPaths(json::Node root, const char* name, std::size_t n, std::vector<std::string>& out) {
    auto node = root.require(name);
    if (node.size() != n) {
        std::ostringstream m;
        m << "Length of " << name << " should match with InputCameraNames";
        throw std::runtime_error(m.str());
    }
    for (std::size_t i = 0; i != node.size(); ++i) out.push_back(node.at(i).asString());
}

void outputPaths(json::Node root, const char* name, std::size_t n, std::vector<std::string>& out) {
    auto node = root.optional(name);
    if (!node) return;
    if (node.size() != n) {
        std::ostringstream m;
        m << "Length of " << name << " should match with VirtualCameraNames";
        throw std::runtime_error(m.str());
    }
    for (std::size_t i = 0; i != node.size(); ++i) out.push_back(node.at(i).asString());
}

}  // namespace

Config Config::loadFromFile(std::string const& filename) {
    std::ifstream stream(filename);
    if (!stream.good()) {
        std::ostringstream m;
        m << "Failed to open configuration file \"" << filename << "\" for reading";
        throw std::runtime_error(m.str());
    }
    auto root = json::Node::readFrom(stream);
    Config c;
    c.version = root.require("Version").asString();
    if (c.version.substr(0, 2) != "2." && c.version.substr(0, 2) != "3.")
        throw std::runtime_error("Configuration file does not match the RVS version");
    auto in = root.require("InputCameraNames");
    for (std::size_t i = 0; i != in.size(); ++i) c.InputCameraNames.push_back(in.at(i).asString());
    auto vn = root.require("VirtualCameraNames");
    for (std::size_t i = 0; i != vn.size(); ++i) c.VirtualCameraNames.push_back(vn.at(i).asString());
    c.params_real = loadCameras(root.require("InputCameraParameterFile").asString(), c.InputCameraNames,
                                root.optional("InputOverrides"));
    c.params_virtual = loadCameras(root.require("VirtualCameraParameterFile").asString(), c.VirtualCameraNames,
                                   root.optional("VirtualOverrides"));
    inputPaths(root, "ViewImageNames", c.InputCameraNames.size(), c.texture_names);
    inputPaths(root, "DepthMapNames", c.InputCameraNames.size(), c.depth_names);
    const std::size_t nv = c.VirtualCameraNames.size();
    outputPaths(root, "OutputFiles", nv, c.outfilenames);
    outputPaths(root, "MaskedOutputFiles", nv, c.outmaskedfilenames);
    outputPaths(root, "OutputMasks", nv, c.outmaskfilenames);
    outputPaths(root, "DepthOutputFiles", nv, c.outdepthfilenames);
    outputPaths(root, "MaskedDepthOutputFiles", nv, c.outmaskdepthfilenames);
    if (auto n = root.optional("ValidityThreshold")) c.validity_threshold = (float)n.asDouble();
    if (auto n = root.optional("ViewSynthesisMethod")) c.vs_method = n.asString();
    if (auto n = root.optional("BlendingMethod")) c.blending_method = n.asString();
    if (auto n = root.optional("BlendingFactor")) c.blending_factor = (float)n.asDouble();
    if (c.blending_method == BlendingMethod::multispectral) {
        c.blending_low_freq_factor = (float)root.require("BlendingLowFreqFactor").asDouble();
        c.blending_high_freq_factor = (float)root.require("BlendingHighFreqFactor").asDouble();
    }
    if (auto n = root.optional("StartFrame")) c.start_frame = n.asInt();
    if (auto n = root.optional("NumberOfFrames")) c.number_of_frames = n.asInt();
    if (auto n = root.optional("NumberOfOutputFrames")) c.number_of_output_frames = n.asInt();
    else c.number_of_output_frames = c.number_of_frames;
    if (auto n = root.optional("Precision")) c.precision = (float)n.asDouble();
    detail::g_rescale = c.precision;
    if (auto n = root.optional("ColorSpace")) {
        if (n.asString() != "YUV" && n.asString() != "RGB") throw std::runtime_error("Unknown color space");
        c.color_space = n.asString();
    }
    if (auto n = root.optional("VirtualPoseTraceName")) {
        const std::string path = n.asString();
        if (!path.empty()) {
            c.pose_trace = PoseTrace::loadFromFile(path);
            if (static_cast<unsigned>(c.start_frame + c.number_of_frames) > c.pose_trace.size())
                throw std::runtime_error(
                    "Error: Number of frames to process is larger then number of entries in pose trace file");
        }
    }
    return c;
}

}  // namespace rvs
