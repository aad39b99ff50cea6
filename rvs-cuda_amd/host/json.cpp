// json.cpp -- recursive-descent reader for the configuration files
// (rvs/src/JsonParser.cpp semantics; see include/rvs/json.hpp).
#include "rvs/json.hpp"

#include <cctype>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace json {

struct Value {
    explicit Value(Node::Type t) : type(t) {}
    virtual ~Value() {}
    Node::Type type;
};

struct Object : Value {
    Object() : Value(Node::Type::object) {}
    std::map<std::string, std::shared_ptr<Value>> members;
};

namespace {

struct StringV : Value {
    StringV() : Value(Node::Type::string) {}
    std::string s;
};
struct NumberV : Value {
    NumberV() : Value(Node::Type::number) {}
    double v = 0;
};
struct ArrayV : Value {
    ArrayV() : Value(Node::Type::array) {}
    std::vector<std::shared_ptr<Value>> items;
};
struct BoolV : Value {
    BoolV() : Value(Node::Type::boolean) {}
    bool v = false;
};
struct NullV : Value {
    NullV() : Value(Node::Type::null) {}
};

class Reader {
public:
    explicit Reader(std::istream& in) : in_(in) {}

    void ws() {
        while (!in_.eof() && std::isspace(in_.peek())) in_.get();
    }

    void expect(int c) {
        const int got = in_.get();
        if (got != c) {
            std::ostringstream m;
            m << "Expected '" << char(c) << "' but found '" << char(got) << "' (0x" << std::hex << got << ")";
            throw std::runtime_error(m.str());
        }
    }

    std::shared_ptr<Value> value() {
        ws();
        const int c = in_.peek();
        switch (c) {
        case '{': return object();
        case '[': return array();
        case '"': return string();
        case 't':
        case 'f': return boolean();
        case 'n': {
            for (char ch : std::string("null")) expect(ch);
            return std::make_shared<NullV>();
        }
        default: break;
        }
        if (c == '-' || std::isdigit(c)) {
            auto n = std::make_shared<NumberV>();
            in_ >> n->v;   // the stream's own number grammar, as the reference reads numbers
            return n;
        }
        std::ostringstream m;
        m << "Invalid character " << char(c) << " (0x" << std::hex << c << ")";
        throw std::runtime_error(m.str());
    }

    std::shared_ptr<StringV> string() {
        expect('"');
        auto s = std::make_shared<StringV>();
        for (int c = in_.get(); c != '"'; c = in_.get()) {
            if (c == '\\') throw std::logic_error("JSON parser: string escaping not implemented");
            s->s.push_back(char(c));
        }
        return s;
    }

    std::shared_ptr<Value> object() {
        expect('{');
        ws();
        auto o = std::make_shared<Object>();
        bool first = true;
        while (in_.peek() != '}') {
            if (!first) {
                expect(',');
                ws();
            }
            first = false;
            auto key = string();
            ws();
            expect(':');
            ws();
            o->members[key->s] = value();
            ws();
        }
        in_.get();
        return o;
    }

    std::shared_ptr<Value> array() {
        expect('[');
        ws();
        auto a = std::make_shared<ArrayV>();
        if (in_.peek() != ']') {
            a->items.push_back(value());
            ws();
            while (in_.peek() == ',') {
                in_.get();
                a->items.push_back(value());
                ws();
            }
        }
        expect(']');
        return a;
    }

    std::shared_ptr<Value> boolean() {
        auto b = std::make_shared<BoolV>();
        b->v = in_.peek() == 't';
        for (char ch : std::string(b->v ? "true" : "false")) expect(ch);
        return b;
    }

private:
    std::istream& in_;
};

}  // namespace

Node::Node() : m_value(std::make_shared<NullV>()) {}
Node::Node(std::shared_ptr<Value> v) : m_value(std::move(v)) {}

Node Node::readFrom(std::istream& in) {
    try {
        in.exceptions(std::ios::badbit | std::ios::failbit);
        Reader r(in);
        auto v = r.value();
        r.ws();
        if (!in.eof()) {
            const int c = in.get();
            std::ostringstream m;
            m << "Stray character " << char(c) << " (0x" << std::hex << c << ")";
            throw std::runtime_error(m.str());
        }
        return Node(v);
    } catch (std::runtime_error& e) {
        throw std::runtime_error(std::string("JSON parser: ") + e.what());
    } catch (std::exception& e) {
        throw std::logic_error(std::string("JSON parser bug: ") + e.what());
    }
}

void Node::setOverrides(Node overrides) {
    if (type() != Type::object || overrides.type() != Type::object)
        throw std::logic_error("JSON overrides apply to objects only");
    m_overrides = std::static_pointer_cast<Object>(overrides.m_value);
}

Node::Type Node::type() const { return m_value->type; }

Node Node::optional(std::string const& key) const {
    if (m_overrides) {
        auto it = m_overrides->members.find(key);
        if (it != m_overrides->members.end()) return Node(it->second);
    }
    if (type() != Type::object) {
        std::ostringstream m;
        m << "JSON parser: Querying optional key '" << key << "', but node is not an object";
        throw std::runtime_error(m.str());
    }
    const auto& mem = static_cast<Object&>(*m_value).members;
    auto it = mem.find(key);
    if (it == mem.end()) return Node();
    return Node(it->second);
}

Node Node::require(std::string const& key) const {
    Node n = optional(key);
    if (n) return n;
    std::ostringstream m;
    m << "JSON parser: Parameter " << key << " is required but missing";
    throw std::runtime_error(m.str());
}

Node Node::at(std::size_t i) const {
    if (type() != Type::array) throw std::runtime_error("JSON parser: Expected an array");
    return Node(static_cast<ArrayV&>(*m_value).items.at(i));
}

std::size_t Node::size() const {
    if (type() == Type::array) return static_cast<ArrayV&>(*m_value).items.size();
    if (type() == Type::object) return static_cast<Object&>(*m_value).members.size();
    throw std::runtime_error("JSON parser: Expected an array or object");
}

double Node::asDouble() const {
    if (type() != Type::number) throw std::runtime_error("JSON parser: Expected a number");
    return static_cast<NumberV&>(*m_value).v;
}

float Node::asFloat() const { return static_cast<float>(asDouble()); }

int Node::asInt() const {
    const double v = asDouble();
    const int r = static_cast<int>(std::lround(v));
    if (v - r > 1e-6) throw std::runtime_error("JSON parser: Expected an integer value");
    return r;
}

std::string const& Node::asString() const {
    if (type() != Type::string) throw std::runtime_error("JSON parser: Expected a string");
    return static_cast<StringV&>(*m_value).s;
}

bool Node::asBool() const {
    if (type() != Type::boolean) throw std::runtime_error("JSON parser: Expected a boolean");
    return static_cast<BoolV&>(*m_value).v;
}

Node::operator bool() const {
    if (type() == Type::null) return false;
    if (type() == Type::boolean) return asBool();
    return true;
}

}  // namespace json
