// host_abi.cpp -- C entry points of librvs_host.so over the C++ front end
// (configuration parsing, pose traces, the configuration-driven Application),
// for callers that bind C (ctypes, cgo, JNI).
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

#include "rvs/config.hpp"
#include "rvs/json.hpp"
#include "rvs/pipeline.hpp"
#include "rvs/server.hpp"
#include "rvs_host.h"

namespace {
thread_local std::string g_err;

int put(char* out, int cap, std::string const& s) {
    if (out && cap > 0) std::snprintf(out, (size_t)cap, "%s", s.c_str());
    return (int)s.size();
}
}  // namespace

extern "C" {

const char* rvs_host_last_error(void) { return g_err.c_str(); }

int rvs_json_query(const char* text, const char* overrides_text, const char* path, const char* as, char* out,
                   int cap) {
    try {
        std::istringstream s(text);
        json::Node node = json::Node::readFrom(s);
        std::istringstream ps(path ? path : "");
        std::string seg;
        if (overrides_text && *overrides_text && node.type() == json::Node::Type::object) {
            std::istringstream os(overrides_text);
            node.setOverrides(json::Node::readFrom(os));
        }
        while (std::getline(ps, seg, '/')) {
            if (seg.empty()) continue;
            const bool numeric = seg.find_first_not_of("0123456789") == std::string::npos;
            if (numeric) node = node.at(std::stoul(seg));
            else if (seg[0] == '?') node = node.optional(seg.substr(1));
            else node = node.require(seg);
            if (overrides_text && *overrides_text && node.type() == json::Node::Type::object) {
                std::istringstream os(overrides_text);
                node.setOverrides(json::Node::readFrom(os));
            }
        }
        const std::string a(as);
        std::ostringstream o;
        o.precision(17);
        if (a == "double") o << node.asDouble();
        else if (a == "float") o << (double)node.asFloat();
        else if (a == "int") o << node.asInt();
        else if (a == "string") o << node.asString();
        else if (a == "bool") o << (node.asBool() ? "true" : "false");
        else if (a == "size") o << node.size();
        else if (a == "type") o << (int)node.type();
        else if (a == "truthy") o << (bool(node) ? "true" : "false");
        else throw std::runtime_error("unknown conversion " + a);
        put(out, cap, o.str());
        return 0;
    } catch (std::exception& e) {
        put(out, cap, e.what());
        return 1;
    }
}

int rvs_pose_trace_parse(const char* text, float* out, int max_rows) {
    try {
        std::istringstream s(text);
        rvs::PoseTrace t = rvs::PoseTrace::loadFrom(s);
        for (int i = 0; i < (int)t.size() && i < max_rows; ++i)
            for (int k = 0; k < 3; ++k) {
                out[6 * i + k] = t[i].position[k];
                out[6 * i + 3 + k] = t[i].rotation[k];
            }
        return (int)t.size();
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int rvs_config_describe(const char* path, char* out, int cap) {
    try {
        rvs::Config c = rvs::Config::loadFromFile(path);
        std::ostringstream o;
        o.precision(9);
        o << "{\"version\":\"" << c.version << "\",\"inputs\":" << c.InputCameraNames.size()
          << ",\"virtual\":" << c.VirtualCameraNames.size() << ",\"blending_factor\":" << c.blending_factor
          << ",\"blending_method\":\"" << c.blending_method << "\",\"start_frame\":" << c.start_frame
          << ",\"number_of_frames\":" << c.number_of_frames
          << ",\"number_of_output_frames\":" << c.number_of_output_frames << ",\"precision\":" << c.precision
          << ",\"validity_threshold\":" << c.validity_threshold << ",\"pose_trace\":" << c.pose_trace.size()
          << ",\"cameras\":[";
        bool first = true;
        for (auto const* list : {&c.params_real, &c.params_virtual})
            for (auto const& p : *list) {
                o << (first ? "" : ",") << "{\"proj\":\"" << p.getProjectionType() << "\",\"w\":" << p.getSize().width
                  << ",\"h\":" << p.getSize().height << ",\"bd\":" << p.getColorBitDepth()
                  << ",\"bdd\":" << p.getDepthBitDepth() << ",\"near\":" << p.getDepthRange()[0]
                  << ",\"far\":" << p.getDepthRange()[1] << ",\"full_hor\":" << (p.isFullHorRange() ? 1 : 0)
                  << ",\"invalid\":" << (p.hasInvalidDepth() ? 1 : 0) << "}";
                first = false;
            }
        o << "]}";
        put(out, cap, o.str());
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        put(out, cap, e.what());
        return 1;
    }
}

int rvs_app_run(const char* config_path, const char* source_path, int stagewise) {
    try {
        rvs::Application app(config_path, source_path ? source_path : "");
        app.setStagewise(stagewise != 0);
        return app.execute();
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// ------------------------------------------------------------------ serving

void* rvs_server_create(const rvs_source* const* scenes, const int* n_src, int n_scenes, const rvs_serve_opts* opts) {
    try {
        if (!scenes || !n_src || n_scenes < 1 || !opts) throw std::invalid_argument("rvs_server_create: bad argument");
        std::vector<std::vector<rvs_source>> sc(n_scenes);
        for (int k = 0; k < n_scenes; ++k) {
            if (!scenes[k] || n_src[k] < 1 || n_src[k] > RVS_MAX_SOURCES)
                throw std::invalid_argument("rvs_server_create: bad scene " + std::to_string(k));
            sc[k].assign(scenes[k], scenes[k] + n_src[k]);
        }
        return new rvs::Server(std::move(sc), *opts);
    } catch (std::exception const& e) {
        g_err = e.what();
        return nullptr;
    }
}

void rvs_server_destroy(void* server) { delete static_cast<rvs::Server*>(server); }

static int copy_stats(rvs::Server* s, int n, rvs_request_stats* stats) {
    if (stats)
        for (int k = 0; k < n; ++k) stats[k] = s->stats()[k];
    return 0;
}

int rvs_server_run_open(void* server, const rvs_request* reqs, int n, double rate, uint64_t seed,
                        rvs_request_stats* stats) {
    try {
        if (!server || (!reqs && n > 0)) throw std::invalid_argument("rvs_server_run_open: bad argument");
        rvs::Server* s = static_cast<rvs::Server*>(server);
        const int rc = s->run_open(reqs, n, rate, seed);
        copy_stats(s, n, stats);
        if (rc != 0) g_err = "a request failed: " + s->error();
        return rc;
    } catch (std::exception const& e) {
        g_err = e.what();
        return RVS_E_INVALID;
    }
}

int rvs_server_run_closed(void* server, const rvs_request* reqs, int n, int concurrency, rvs_request_stats* stats) {
    try {
        if (!server || (!reqs && n > 0)) throw std::invalid_argument("rvs_server_run_closed: bad argument");
        rvs::Server* s = static_cast<rvs::Server*>(server);
        const int rc = s->run_closed(reqs, n, concurrency);
        copy_stats(s, n, stats);
        if (rc != 0) g_err = "a request failed: " + s->error();
        return rc;
    } catch (std::exception const& e) {
        g_err = e.what();
        return RVS_E_INVALID;
    }
}

long long rvs_server_output(void* server, int k, void* out, size_t cap) {
    if (!server) return -1;
    rvs::Server* s = static_cast<rvs::Server*>(server);
    if (k < 0 || k >= (int)s->stats().size()) return -1;
    const std::vector<uint8_t>& o = s->output(k);
    if (o.empty()) return -1;
    if (out) std::memcpy(out, o.data(), o.size() < cap ? o.size() : cap);
    return (long long)o.size();
}

}  // extern "C"
