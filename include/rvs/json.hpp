// json.hpp -- the configuration-file JSON reader of the reference view
// synthesizer, same interface and acceptance rules as json::Node in
// rvs/src/JsonParser.hpp:57-119 (string escapes rejected, asInt tolerance 1e-6
// on the upper side, overrides consulted first, a boolean false counts as
// absent for require()).
#ifndef RVS_JSON_HPP
#define RVS_JSON_HPP

#include <istream>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace json {

struct Value;
struct Object;

class Node {
public:
    enum class Type { number, string, array, object, boolean, null };

    /** Read JSON from an input stream (JsonParser.cpp:203-225). */
    static Node readFrom(std::istream&);
    /** Overrides consulted before the object's own keys (object nodes only). */
    void setOverrides(Node overrides);
    Type type() const;
    /** Object member, Null if absent. */
    Node optional(std::string const& key) const;
    /** Object member; throws "Parameter KEY is required but missing". */
    Node require(std::string const& key) const;
    /** Array element (std::out_of_range past the end). */
    Node at(std::size_t index) const;
    /** Number of array elements or object members. */
    std::size_t size() const;
    double asDouble() const;
    float asFloat() const;
    int asInt() const;
    std::string const& asString() const;
    bool asBool() const;
    /** false for null and for the boolean false, true otherwise. */
    explicit operator bool() const;

private:
    Node();
    Node(std::shared_ptr<Value>);
    std::shared_ptr<Value> m_value;
    std::shared_ptr<Object> m_overrides;
};

}  // namespace json

#endif
