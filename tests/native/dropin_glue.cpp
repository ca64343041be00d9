// Test glue: calls the drop-in implicit_arrangement() / material_interface()
// (GPU path + host graph stages) the way the reference's own tests call them
// (tests/test_implicit_networks.cpp:68), with a row-major matrix type standing
// in for Eigen at the call site, and reports the same summary as ka_glue.cpp.
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "rin/implicit_arrangement.h"

namespace {

// the call-site shape of Eigen::Matrix<double, Dynamic, Dynamic, RowMajor>
struct RowMajorMatrix {
    static constexpr bool IsRowMajor = true;
    std::vector<double> v;
    size_t r = 0, c = 0;
    const double* data() const { return v.data(); }
    size_t rows() const { return r; }
    size_t cols() const { return c; }
};

thread_local std::string g_out;

template <typename T>
void put_list(std::ostringstream& o, const std::vector<T>& v) {
    o << "[";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << v[i];
    o << "]";
}

}  // namespace

extern "C" const char* dropin_run(int mi, int robust, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T,
                                  const double* vals, uint64_t F) {
    std::vector<std::array<double, 3>> P(V);
    std::memcpy(P.data(), pts, V * 24);
    std::vector<std::array<size_t, 4>> Tt(T);
    for (uint64_t t = 0; t < T; ++t)
        for (int c = 0; c < 4; ++c) Tt[t][c] = tets[4 * t + c];
    RowMajorMatrix M;
    M.v.assign(vals, vals + V * F);
    M.r = V;
    M.c = F;
    std::vector<std::array<double, 3>> out_pts;
    std::vector<PolygonFace> faces;
    std::vector<std::vector<size_t>> patches, chains, nm, shells, cells;
    std::vector<Edge> edges;
    std::vector<std::string> tl, sl;
    std::vector<double> tv;
    std::vector<size_t> sv;
    std::vector<size_t> patch_label;
    std::vector<std::pair<size_t, size_t>> patch_label_mi;
    std::vector<std::vector<bool>> cell_label;
    std::vector<size_t> cell_label_mi;
    bool ok;
    if (mi)
        ok = material_interface(robust != 0, true, true, true, P, Tt, M, out_pts, faces, patches, patch_label_mi,
                                edges, chains, nm, shells, cells, cell_label_mi, tl, tv, sl, sv);
    else
        ok = implicit_arrangement(robust != 0, true, true, true, P, Tt, M, out_pts, faces, patches, patch_label,
                                  edges, chains, nm, shells, cells, cell_label, tl, tv, sl, sv);
    size_t corners = 0;
    for (auto& l : nm) corners += l.size() > 2;
    std::ostringstream o;
    o << "{\"ok\":" << (ok ? "true" : "false");
    o << ",\"num_verts\":" << out_pts.size() << ",\"num_faces\":" << faces.size();
    o << ",\"patches\":" << patches.size() << ",\"chains\":" << chains.size();
    o << ",\"shells\":" << shells.size() << ",\"cells\":" << cells.size() << ",\"corners\":" << corners
      << ",\"edges\":" << edges.size();
    o << ",\"stats\":{";
    for (size_t i = 0; i < sl.size(); ++i) o << (i ? "," : "") << "\"" << sl[i] << "\":" << sv[i];
    o << "},\"patch_label\":";
    if (mi) {
        o << "[";
        for (size_t i = 0; i < patch_label_mi.size(); ++i)
            o << (i ? "," : "") << "[" << patch_label_mi[i].first << "," << patch_label_mi[i].second << "]";
        o << "]";
    } else {
        put_list(o, patch_label);
    }
    o << ",\"cell_label\":";
    if (mi) {
        put_list(o, cell_label_mi);
    } else {
        o << "[";
        for (size_t i = 0; i < cell_label.size(); ++i) {
            o << (i ? "," : "") << "[";
            for (size_t j = 0; j < cell_label[i].size(); ++j) o << (j ? "," : "") << (cell_label[i][j] ? 1 : 0);
            o << "]";
        }
        o << "]";
    }
    o << "}";
    g_out = o.str();
    return g_out.c_str();
}

// the -R verdict with given per-tet counts (CPU; prints the reference's message)
extern "C" int verdict_probe(int forward_ok, int reverse_ok, const uint64_t* a, const uint64_t* b, uint64_t n) {
    std::vector<uint64_t> x(a, a + n), y(b, b + n);
    return rin::robust_verdict(forward_ok != 0, reverse_ok != 0, x, y) ? 1 : 0;
}
