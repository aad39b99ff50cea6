// C driver around the reference's json::Node (rvs/src/JsonParser.hpp:57-119),
// compiled together with the reference's own JsonParser.cpp into
// oracle/_ref/libref_json.so.  Test infrastructure only: tests compare the
// host library's JSON parser with the reference parser through refjson_query.
#include "JsonParser.hpp"

#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

namespace {
int put(char* out, int cap, const std::string& s) {
    std::snprintf(out, (size_t)cap, "%s", s.c_str());
    return 0;
}
}  // namespace

// Parses `text` (optionally with `overrides_text` applied to the object reached
// by `path` when `path` ends with '!'), walks `path` ("a/0/b", numeric segments
// index arrays, empty path = root) with Node::require/at, then converts with
// `as` in {double,float,int,string,bool,size,type,truthy,optional}.
// Returns 0 and the value in `out`, or 1 and the exception text.
extern "C" int refjson_query(const char* text, const char* overrides_text, const char* path,
                             const char* as, char* out, int cap) {
    try {
        std::istringstream s(text);
        json::Node node = json::Node::readFrom(s);
        std::string p(path ? path : "");
        std::string seg;
        std::istringstream ps(p);
        if (overrides_text && *overrides_text && node.type() == json::Node::Type::object) {
            std::istringstream os(overrides_text);
            node.setOverrides(json::Node::readFrom(os));
        }
        while (std::getline(ps, seg, '/')) {
            if (seg.empty()) continue;
            bool numeric = seg.find_first_not_of("0123456789") == std::string::npos;
            if (numeric)
                node = node.at(std::stoul(seg));
            else if (seg[0] == '?')
                node = node.optional(seg.substr(1));
            else
                node = node.require(seg);
            if (overrides_text && *overrides_text && node.type() == json::Node::Type::object) {
                std::istringstream os(overrides_text);
                node.setOverrides(json::Node::readFrom(os));
            }
        }
        std::string a(as);
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
        else throw std::runtime_error("driver: unknown conversion " + a);
        return put(out, cap, o.str());
    } catch (std::exception& e) {
        put(out, cap, e.what());
        return 1;
    }
}
