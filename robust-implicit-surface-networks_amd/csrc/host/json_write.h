// JSON text in the form the reference's files have (nlohmann::json's dump()
// with default settings, src/io.cpp:154-590): no whitespace, object keys in
// ascending byte order, integers as integers, doubles in the shortest form
// Grisu2 finds (Loitsch, "Printing Floating-Point Numbers Quickly and
// Accurately with Integers", PLDI 2010) laid out as plain decimals between
// 1e-4 and 1e15 (with ".0" on integral values) and as d.ddde+XX outside.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace rin {
namespace json {

// shortest-form decimal of a double, as nlohmann::json prints it ("null" for
// NaN and infinities)
void put_double(std::string& out, double v);
void put_uint(std::string& out, uint64_t v);
// a quoted string with JSON escapes
void put_string(std::string& out, const std::string& s);

}  // namespace json
}  // namespace rin
