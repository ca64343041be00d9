// The reference's input/output routines (declared in include/rin/io.h), file
// formats as src/io.cpp:19-312 and :556-590 read and write them.
#include "rin/io.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

#include "json_lite.h"
#include "json_write.h"

namespace fs = std::filesystem;
using rin::json::put_double;
using rin::json::put_string;
using rin::json::put_uint;

namespace {

// src/io.cpp:12-17: relative paths are taken from the configuration's folder
std::string resolve(const fs::path& base, const std::string& p) {
    fs::path q(p);
    if (q.is_relative()) q = fs::absolute(base / q);
    return q.string();
}

std::array<double, 3> arr3(const rin::json::Value& v) {
    return {v[0].as_double(), v[1].as_double(), v[2].as_double()};
}

// Reads `[[x, y, z], ...], [[a, b, c, d], ...]` (the tet-mesh file) with one
// pass over the text: nested lists of numbers only.
struct MeshReader {
    const char* s;
    const char* e;
    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("tet mesh file: ") + what);
    }
    void ws() {
        while (s < e && (*s == ' ' || *s == '\n' || *s == '\r' || *s == '\t')) ++s;
    }
    void expect(char c) {
        ws();
        if (s >= e || *s != c) fail("unexpected character");
        ++s;
    }
    bool peek(char c) {
        ws();
        return s < e && *s == c;
    }
    double number() {
        ws();
        char buf[64];
        int n = 0;
        while (s < e && n < 63 && (*s == '-' || *s == '+' || *s == '.' || *s == 'e' || *s == 'E' || (*s >= '0' && *s <= '9')))
            buf[n++] = *s++;
        buf[n] = 0;
        char* end = nullptr;
        const double v = std::strtod(buf, &end);
        if (n == 0 || end != buf + n) fail("bad number");
        return v;
    }
    template <size_t K, class T, class Conv>
    void rows(std::vector<std::array<T, K>>& out, Conv conv) {
        expect('[');
        if (peek(']')) {
            ++s;
            return;
        }
        while (true) {
            std::array<T, K> r;
            expect('[');
            for (size_t k = 0; k < K; ++k) {
                if (k) expect(',');
                r[k] = conv(number());
            }
            expect(']');
            out.push_back(r);
            ws();
            if (peek(',')) {
                ++s;
                continue;
            }
            expect(']');
            return;
        }
    }
};

size_t to_index(double x) {
    if (!(x >= 0) || x != std::floor(x) || x >= 18446744073709551616.0)
        throw std::runtime_error("tet mesh file: a corner index is not a non-negative integer");
    return (size_t)x;
}

// streamed writing: the text goes to the file in chunks of a few megabytes
struct Out {
    std::ofstream f;
    std::string b;
    explicit Out(const std::string& name) : f(name, std::ios::binary) {}
    void flush_if_big() {
        if (b.size() > (4u << 20)) {
            f.write(b.data(), (std::streamsize)b.size());
            b.clear();
        }
    }
    bool close() {
        b += '\n';
        f.write(b.data(), (std::streamsize)b.size());
        b.clear();
        f.close();
        return true;
    }
};

// nlohmann::json's serialisation of the lists src/io.cpp builds: a list that
// received no element stays null; a list of lists of integers
template <class Rows, class Put>
void put_list(Out& o, const Rows& rows, Put put) {
    if (rows.empty()) {
        o.b += "null";
        return;
    }
    o.b += '[';
    for (size_t i = 0; i < rows.size(); ++i) {
        if (i) o.b += ',';
        put(rows[i]);
        o.flush_if_big();
    }
    o.b += ']';
}

void put_uints(std::string& b, const std::vector<size_t>& v) {
    b += '[';
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) b += ',';
        put_uint(b, v[i]);
    }
    b += ']';
}

// the members save_result and save_result_MI share, in key order
struct Common {
    const std::vector<std::array<double, 3>>& pts;
    const std::vector<PolygonFace>& faces;
    const std::vector<std::vector<size_t>>& patches;
    const std::vector<Edge>& edges;
    const std::vector<std::vector<size_t>>& chains;
    const std::vector<std::vector<size_t>>& nm;
    const std::vector<std::vector<size_t>>& shells;
    const std::vector<std::vector<size_t>>& cells;
};

template <class PutPatchLabel, class PutCellLabel>
bool write_mesh_json(const std::string& filename, const Common& c, PutPatchLabel patch_label,
                     PutCellLabel cell_label) {
    Out o(filename);
    if (!o.f) return false;
    std::string& b = o.b;
    auto lists = [&](const std::vector<std::vector<size_t>>& v) {
        put_list(o, v, [&](const std::vector<size_t>& r) { put_uints(b, r); });
    };
    // keys in ascending order, as nlohmann::json's object keeps them
    b += "{\"cells\":";
    lists(c.cells);
    b += ",\"cells_label\":";
    cell_label(o);
    b += ",\"chains\":";
    lists(c.chains);
    b += ",\"corners\":";
    std::vector<size_t> corners;
    for (size_t i = 0; i < c.nm.size(); ++i)
        if (c.nm[i].size() > 2 || c.nm[i].size() == 1) corners.push_back(i);
    if (corners.empty())
        b += "null";
    else
        put_uints(b, corners);
    b += ",\"edges\":";
    put_list(o, c.edges, [&](const Edge& e) {
        b += '[';
        put_uint(b, e.v1);
        b += ',';
        put_uint(b, e.v2);
        b += ']';
    });
    b += ",\"faces\":";
    put_list(o, c.faces, [&](const PolygonFace& f) { put_uints(b, f.vert_indices); });
    b += ",\"patches\":";
    lists(c.patches);
    b += ",\"patches_label\":[";
    patch_label(o);
    b += "],\"points\":";
    put_list(o, c.pts, [&](const std::array<double, 3>& p) {
        b += '[';
        put_double(b, p[0]);
        b += ',';
        put_double(b, p[1]);
        b += ',';
        put_double(b, p[2]);
        b += ']';
    });
    b += ",\"shells\":";
    lists(c.shells);
    b += '}';
    return o.close();
}

}  // namespace

Config parse_config_file(const std::string& filename) {
    std::ifstream probe(filename);
    if (!probe) throw std::runtime_error("Config file does not exist!");
    probe.close();
    const rin::json::Value d = rin::json::parse_file(filename);
    Config c;
    const fs::path base = fs::path(filename).parent_path();
    if (d.contains("tetMeshFile")) {
        c.tet_mesh_file = resolve(base, d.at("tetMeshFile").as_string());
        c.tet_mesh_resolution = 0;
    } else {
        c.tet_mesh_file = "";
        c.tet_mesh_resolution = (size_t)d.at("gridResolution").as_index();
        c.tet_mesh_bbox_min = arr3(d.at("gridBbox")[0]);
        c.tet_mesh_bbox_max = arr3(d.at("gridBbox")[1]);
    }
    c.func_file = resolve(base, d.at("funcFile").as_string());
    c.output_dir = resolve(base, d.at("outputDir").as_string());
    c.use_lookup = d.at("useLookup").as_bool();
    c.use_secondary_lookup = d.at("useSecondaryLookup").as_bool();
    c.use_topo_ray_shooting = d.at("useTopoRayShooting").as_bool();
    return c;
}

bool load_tet_mesh(const std::string& filename, std::vector<std::array<double, 3>>& pts,
                   std::vector<std::array<size_t, 4>>& tets) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) {
        std::cout << "tet mesh file not exist!" << std::endl;
        return false;
    }
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    MeshReader R{text.data(), text.data() + text.size()};
    pts.clear();
    tets.clear();
    R.expect('[');
    R.rows<3, double>(pts, [](double x) { return x; });
    R.expect(',');
    R.rows<4, size_t>(tets, to_index);
    R.expect(']');
    return true;
}

bool generate_tet_mesh(size_t resolution, const std::array<double, 3>& bbox_min, const std::array<double, 3>& bbox_max,
                       std::vector<std::array<double, 3>>& pts, std::vector<std::array<size_t, 4>>& tets) {
    if (resolution == 0) return false;
    const size_t N = resolution + 1, R = resolution;
    pts.resize(N * N * N);
    auto coord = [&](double t, int a) { return t * (bbox_max[a] - bbox_min[a]) + bbox_min[a]; };
    for (size_t i = 0; i < N; ++i) {
        const double x = coord(double(i) / double(N - 1), 0);
        for (size_t j = 0; j < N; ++j) {
            const double y = coord(double(j) / double(N - 1), 1);
            for (size_t k = 0; k < N; ++k) pts[(i * N + j) * N + k] = {x, y, coord(double(k) / double(N - 1), 2)};
        }
    }
    // five tets per cube, split by the parity of i + j + k
    static const int even[5][4] = {{4, 6, 1, 3}, {6, 3, 4, 7}, {1, 3, 0, 4}, {3, 1, 2, 6}, {4, 1, 6, 5}};
    static const int odd[5][4] = {{7, 0, 2, 5}, {2, 3, 0, 7}, {5, 7, 0, 4}, {7, 2, 6, 5}, {0, 1, 2, 5}};
    tets.resize(R * R * R * 5);
    for (size_t i = 0; i < R; ++i)
        for (size_t j = 0; j < R; ++j)
            for (size_t k = 0; k < R; ++k) {
                const size_t b = (i * N + j) * N + k;
                const size_t v[8] = {b, b + N * N, b + N * N + N, b + N, b + 1, b + N * N + 1, b + N * N + N + 1,
                                     b + N + 1};
                const int(*tab)[4] = ((i + j + k) % 2 == 0) ? even : odd;
                const size_t t = ((i * R + j) * R + k) * 5;
                for (int e = 0; e < 5; ++e) tets[t + e] = {v[tab[e][0]], v[tab[e][1]], v[tab[e][2]], v[tab[e][3]]};
            }
    return true;
}

bool save_result(const std::string& filename, const std::vector<std::array<double, 3>>& mesh_pts,
                 const std::vector<PolygonFace>& mesh_faces, const std::vector<std::vector<size_t>>& patches,
                 const std::vector<size_t>& patch_function_label, const std::vector<Edge>& edges,
                 const std::vector<std::vector<size_t>>& chains,
                 const std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                 const std::vector<std::vector<size_t>>& shells, const std::vector<std::vector<size_t>>& cells,
                 const std::vector<std::vector<bool>>& cell_function_label) {
    Common c{mesh_pts, mesh_faces, patches, edges, chains, non_manifold_edges_of_vert, shells, cells};
    return write_mesh_json(
        filename, c, [&](Out& o) { put_uints(o.b, patch_function_label); },
        [&](Out& o) {
            put_list(o, cell_function_label, [&](const std::vector<bool>& r) {
                o.b += '[';
                for (size_t i = 0; i < r.size(); ++i) {
                    if (i) o.b += ',';
                    o.b += r[i] ? "true" : "false";
                }
                o.b += ']';
            });
        });
}

bool save_result_MI(const std::string& filename, const std::vector<std::array<double, 3>>& mesh_pts,
                    const std::vector<PolygonFace>& mesh_faces, const std::vector<std::vector<size_t>>& patches,
                    const std::vector<std::pair<size_t, size_t>>& patch_function_label,
                    const std::vector<Edge>& edges, const std::vector<std::vector<size_t>>& chains,
                    const std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                    const std::vector<std::vector<size_t>>& shells, const std::vector<std::vector<size_t>>& cells,
                    const std::vector<size_t>& cell_function_label) {
    Common c{mesh_pts, mesh_faces, patches, edges, chains, non_manifold_edges_of_vert, shells, cells};
    return write_mesh_json(
        filename, c,
        [&](Out& o) {
            o.b += '[';
            for (size_t i = 0; i < patch_function_label.size(); ++i) {
                if (i) o.b += ',';
                o.b += '[';
                put_uint(o.b, patch_function_label[i].first);
                o.b += ',';
                put_uint(o.b, patch_function_label[i].second);
                o.b += ']';
            }
            o.b += ']';
        },
        [&](Out& o) {
            o.b += '[';
            put_uints(o.b, cell_function_label);
            o.b += ']';
        });
}

bool save_timings(const std::string& filename, const std::vector<std::string>& timing_labels,
                  const std::vector<double>& timings) {
    // an object keyed by label: ascending keys, the last value of a repeated label
    std::map<std::string, double> m;
    for (size_t i = 0; i < timings.size(); ++i) m[timing_labels[i]] = timings[i];
    Out o(filename);
    if (!o.f) return false;
    if (m.empty()) {
        o.b += "null";
    } else {
        o.b += '{';
        bool first = true;
        for (auto& kv : m) {
            if (!first) o.b += ',';
            first = false;
            put_string(o.b, kv.first);
            o.b += ':';
            put_double(o.b, kv.second);
        }
        o.b += '}';
    }
    return o.close();
}

bool save_statistics(const std::string& filename, const std::vector<std::string>& stats_labels,
                     const std::vector<size_t>& stats) {
    std::map<std::string, size_t> m;
    for (size_t i = 0; i < stats.size(); ++i) m[stats_labels[i]] = stats[i];
    Out o(filename);
    if (!o.f) return false;
    if (m.empty()) {
        o.b += "null";
    } else {
        o.b += '{';
        bool first = true;
        for (auto& kv : m) {
            if (!first) o.b += ',';
            first = false;
            put_string(o.b, kv.first);
            o.b += ':';
            put_uint(o.b, kv.second);
        }
        o.b += '}';
    }
    return o.close();
}
