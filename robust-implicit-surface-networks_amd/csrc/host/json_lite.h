// A small JSON reader for the reference's input files: function files
// (a list of typed objects), tet-mesh files ([points, tets]) and the CLI
// configuration (src/io.cpp:19-93).  Numbers are parsed with strtod, so every
// double is the correctly rounded value of its text.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rin {
namespace json {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool is_array() const { return type == Array; }
    bool is_object() const { return type == Object; }
    bool is_number() const { return type == Number; }
    size_t size() const { return type == Array ? arr.size() : (type == Object ? obj.size() : 0); }
    const Value& operator[](size_t i) const;
    const Value* find(const std::string& key) const;  // nullptr when absent
    const Value& at(const std::string& key) const;    // throws when absent
    bool contains(const std::string& key) const { return find(key) != nullptr; }
    double as_double() const;
    bool as_bool() const;
    uint64_t as_index() const;  // a non-negative integer
    const std::string& as_string() const;
};

// parse a whole document; throws std::runtime_error with the byte offset
Value parse(const std::string& text);
Value parse_file(const std::string& path);

}  // namespace json
}  // namespace rin
