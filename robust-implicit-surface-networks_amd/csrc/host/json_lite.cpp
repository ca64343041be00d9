#include "json_lite.h"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace rin {
namespace json {

namespace {

struct Parser {
    const char* s;
    const char* e;
    const char* b;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("json: ") + what + " at byte " + std::to_string(s - b));
    }
    void ws() {
        while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) ++s;
    }
    bool lit(const char* w) {
        const char* p = s;
        for (; *w; ++w, ++p)
            if (p >= e || *p != *w) return false;
        s = p;
        return true;
    }
    static void put_utf8(std::string& o, uint32_t c) {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xC0 | (c >> 6));
            o += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            o += (char)(0xE0 | (c >> 12));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        } else {
            o += (char)(0xF0 | (c >> 18));
            o += (char)(0x80 | ((c >> 12) & 0x3F));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4() {
        if (e - s < 4) fail("short \\u escape");
        uint32_t c = 0;
        for (int i = 0; i < 4; ++i, ++s) {
            const char h = *s;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return c;
    }
    std::string string() {
        ++s;  // opening quote
        std::string o;
        while (true) {
            if (s >= e) fail("unterminated string");
            const char c = *s++;
            if (c == '"') break;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (s >= e) fail("unterminated escape");
            const char x = *s++;
            switch (x) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e - s >= 6 && s[0] == '\\' && s[1] == 'u') {
                        s += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(o, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return o;
    }
    void value(Value& v, int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (s >= e) fail("unexpected end");
        const char c = *s;
        if (c == '{') {
            ++s;
            v.type = Value::Object;
            ws();
            if (s < e && *s == '}') {
                ++s;
                return;
            }
            while (true) {
                ws();
                if (s >= e || *s != '"') fail("expected a key");
                std::string k = string();
                ws();
                if (s >= e || *s != ':') fail("expected ':'");
                ++s;
                v.obj.emplace_back(std::move(k), Value());
                value(v.obj.back().second, depth + 1);
                ws();
                if (s < e && *s == ',') {
                    ++s;
                    continue;
                }
                if (s < e && *s == '}') {
                    ++s;
                    return;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++s;
            v.type = Value::Array;
            ws();
            if (s < e && *s == ']') {
                ++s;
                return;
            }
            while (true) {
                v.arr.emplace_back();
                value(v.arr.back(), depth + 1);
                ws();
                if (s < e && *s == ',') {
                    ++s;
                    continue;
                }
                if (s < e && *s == ']') {
                    ++s;
                    return;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = Value::String;
            v.str = string();
            return;
        }
        if (lit("true")) {
            v.type = Value::Bool;
            v.b = true;
            return;
        }
        if (lit("false")) {
            v.type = Value::Bool;
            v.b = false;
            return;
        }
        if (lit("null")) {
            v.type = Value::Null;
            return;
        }
        if (c == '-' || (c >= '0' && c <= '9')) {
            // copy the token: strtod needs a terminated string
            const char* p = s;
            while (p < e && (*p == '-' || *p == '+' || *p == '.' || *p == 'e' || *p == 'E' || (*p >= '0' && *p <= '9')))
                ++p;
            std::string tok(s, p);
            char* end = nullptr;
            errno = 0;
            v.num = std::strtod(tok.c_str(), &end);
            if (end != tok.c_str() + tok.size() || tok.empty()) fail("bad number");
            v.type = Value::Number;
            s = p;
            return;
        }
        fail("unexpected character");
    }
};

}  // namespace

const Value& Value::operator[](size_t i) const {
    if (type != Array || i >= arr.size()) throw std::runtime_error("json: array index out of range");
    return arr[i];
}

const Value* Value::find(const std::string& key) const {
    if (type != Object) return nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const Value& Value::at(const std::string& key) const {
    const Value* v = find(key);
    if (!v) throw std::runtime_error("json: missing key \"" + key + "\"");
    return *v;
}

double Value::as_double() const {
    if (type != Number) throw std::runtime_error("json: expected a number");
    return num;
}

bool Value::as_bool() const {
    if (type != Bool) throw std::runtime_error("json: expected true or false");
    return b;
}

uint64_t Value::as_index() const {
    if (type != Number || num < 0 || num != std::floor(num) || num >= 18446744073709551616.0)
        throw std::runtime_error("json: expected a non-negative integer");
    return (uint64_t)num;
}

const std::string& Value::as_string() const {
    if (type != String) throw std::runtime_error("json: expected a string");
    return str;
}

Value parse(const std::string& text) {
    Parser P{text.data(), text.data() + text.size(), text.data()};
    Value v;
    P.value(v, 0);
    P.ws();
    if (P.s != P.e) P.fail("trailing characters");
    return v;
}

Value parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace json
}  // namespace rin
