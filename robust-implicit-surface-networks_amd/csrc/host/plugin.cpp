// The inner libraries the reference links and its callers use directly:
// simplicial_arrangement's lookup-table switches and implicit_functions'
// load_functions (include/simplicial_arrangement/lookup_table.h,
// include/implicit_functions.h).
#include <atomic>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "implicit_functions.h"
#include "json_lite.h"
#include "rin.h"
#include "simplicial_arrangement/lookup_table.h"
#include "thread_ctx.h"

namespace {
std::atomic<bool> g_lookup_on{true};

void unit(const rin::json::Value& v, double* out) {
    double a[3];
    for (int i = 0; i < 3; ++i) a[i] = v[i].as_double();
    const double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(n > 0)) throw std::runtime_error("zero axis vector");
    for (int i = 0; i < 3; ++i) out[i] = a[i] / n;
}

void vec3(const rin::json::Value& v, double* out) {
    for (int i = 0; i < 3; ++i) out[i] = v[i].as_double();
}

bool flag(const rin::json::Value& o, const char* key) {
    const rin::json::Value* v = o.find(key);
    return v && v->as_bool();
}

// one function-file entry -> descriptor (layout of rin_func_desc)
rin_func_desc describe(const rin::json::Value& o) {
    rin_func_desc d{};
    const std::string& t = o.at("type").as_string();
    double* p = d.params;
    if (t == "plane") {
        d.type = 0;
        vec3(o.at("point"), p);
        vec3(o.at("normal"), p + 3);
    } else if (t == "sphere") {
        d.type = 1;
        vec3(o.at("center"), p);
        p[3] = o.at("radius").as_double();
        p[4] = flag(o, "squared") ? 1.0 : 0.0;
    } else if (t == "cylinder") {
        d.type = 2;
        vec3(o.at("axis_point"), p);
        unit(o.at("axis_vector"), p + 3);
        p[6] = o.at("radius").as_double();
        p[7] = flag(o, "is_flipped") ? 1.0 : 0.0;
    } else if (t == "torus") {
        d.type = 3;
        vec3(o.at("center"), p);
        unit(o.at("axis_vector"), p + 3);
        p[6] = o.at("major_radius").as_double();
        p[7] = o.at("minor_radius").as_double();
        p[8] = flag(o, "is_flipped") ? 1.0 : 0.0;
    } else if (t == "zero") {
        d.type = 4;
    } else {
        throw std::runtime_error("unsupported function type \"" + t + "\"");
    }
    return d;
}
}  // namespace

namespace simplicial_arrangement {
bool load_lookup_table(Mode) {
    g_lookup_on = true;
    return true;
}
void enable_lookup_table() { g_lookup_on = true; }
void disable_lookup_table() { g_lookup_on = false; }
bool lookup_table_enabled() { return g_lookup_on; }
}  // namespace simplicial_arrangement

namespace rin {

bool load_function_values(const std::string& filename, const std::vector<std::array<double, 3>>& pts,
                          std::vector<double>& values, size_t& n_funcs) {
    std::vector<rin_func_desc> descs;
    try {
        const json::Value doc = json::parse_file(filename);
        if (!doc.is_array()) throw std::runtime_error("a function file is a JSON list");
        for (size_t i = 0; i < doc.size(); ++i) descs.push_back(describe(doc[i]));
    } catch (std::exception& e) {
        std::cout << "load_functions: " << filename << ": " << e.what() << std::endl;
        return false;
    }
    n_funcs = descs.size();
    values.assign(pts.size() * n_funcs, 0.0);
    if (descs.empty() || pts.empty()) return !descs.empty();
    try {
        rin_ctx* c = detail::thread_context();
        if (rin_eval_points(c, &pts[0][0], pts.size(), descs.data(), (uint32_t)descs.size(), values.data()) != RIN_OK) {
            std::cout << "load_functions: " << rin_last_error(c) << std::endl;
            return false;
        }
    } catch (std::exception& e) {
        std::cout << "load_functions: " << e.what() << std::endl;
        return false;
    }
    return true;
}

}  // namespace rin
