// Test harness glue: runs the oracle's hot path and feeds its outputs to the
// product's host graph stages, so the reference's known-answer integration
// tests (tests/test_implicit_networks.cpp:19-732) can pin the oracle on CPU.
#include <cstring>
#include <sstream>
#include <string>

#include "../../oracle/oracle.h"
#include "cut_view.h"
#include "graph.h"

namespace {

struct OracleCuts : rin::CutSource {
    const oracle::Result* r;
    std::vector<std::pair<size_t, rin::CutView>> cache;
    std::vector<int64_t> slot;
    const rin::CutView& get(size_t tet) override {
        if (slot.empty()) slot.assign(r->nt, -1);
        if (slot[tet] < 0) {
            int64_t a = r->cut_index[tet];
            if (a < 0) throw std::runtime_error("cut requested for an empty tet");
            std::vector<uint16_t> w;
            if (r->mi)
                oracle::serialize_mi(r->mi_cuts[a], w);
            else
                oracle::serialize_ia(r->ia_cuts[a], w);
            slot[tet] = (int64_t)cache.size();
            cache.push_back({tet, rin::decode_cut(w.data())});
        }
        return cache[slot[tet]].second;
    }
};

thread_local std::string g_out;

template <typename T>
void put_list(std::ostringstream& o, const std::vector<T>& v) {
    o << "[";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? "," : "") << v[i];
    o << "]";
}

}  // namespace

extern "C" const char* ka_run(int mi, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T,
                              const double* vals, uint64_t F) {
    std::ostringstream o;
    try {
        std::vector<std::array<double, 3>> P(V);
        std::memcpy(P.data(), pts, V * 24);
        std::vector<std::array<uint64_t, 4>> Tt(T);
        std::memcpy(Tt.data(), tets, T * 32);
        std::vector<double> vv(vals, vals + V * F);
        oracle::Result r;
        if (mi)
            oracle::run_mi(P, Tt, vv, F, oracle::F_KEEP_CUTS, r);
        else
            oracle::run_ia(P, Tt, vv, F, oracle::F_KEEP_CUTS, r);

        std::vector<std::array<double, 3>> pts3(P.begin(), P.end());
        std::vector<std::array<size_t, 4>> tets4(T);
        for (uint64_t t = 0; t < T; ++t)
            for (int c = 0; c < 4; ++c) tets4[t][c] = Tt[t][c];
        std::vector<size_t> fit(r.func_in_tet.begin(), r.func_in_tet.end());
        std::vector<size_t> sidx(r.start_index.begin(), r.start_index.end());
        std::vector<PolygonFace> faces(r.faces.size());
        for (size_t i = 0; i < r.faces.size(); ++i) {
            faces[i].vert_indices.assign(r.faces[i].vs.begin(), r.faces[i].vs.end());
            for (auto& tf : r.faces[i].tet_faces) faces[i].tet_face_indices.push_back(tf);
            faces[i].func_index = {r.faces[i].f0, r.faces[i].f1};
        }
        std::vector<IsoVert> iv;
        std::vector<MI_Vert> mv;
        for (const auto& v : r.verts) {
            if (mi) {
                MI_Vert m{};
                m.tet_index = v.tet;
                m.tet_vert_index = v.tet_vert;
                m.simplex_size = v.simplex_size;
                for (int q = 0; q < 4; ++q) {
                    m.simplex_vert_indices[q] = v.sv[q];
                    m.material_indices[q] = v.fn[q];
                }
                mv.push_back(m);
            } else {
                IsoVert m{};
                m.tet_index = v.tet;
                m.tet_vert_index = v.tet_vert;
                m.simplex_size = v.simplex_size;
                for (int q = 0; q < 4; ++q) m.simplex_vert_indices[q] = v.sv[q];
                for (int q = 0; q < 3; ++q) m.func_indices[q] = v.fn[q];
                iv.push_back(m);
            }
        }
        std::vector<uint8_t> degen(V, 0);
        bool found = false;
        if (mi) {
            for (uint64_t v = 0; v < V; ++v) degen[v] = r.degenerate_vertex[v];
        } else {
            for (uint64_t v = 0; v < V; ++v)
                for (uint64_t f = 0; f < F; ++f)
                    if (vv[v * F + f] == 0) degen[v] = 1;
        }
        for (auto d : degen) found |= d != 0;

        OracleCuts cuts;
        cuts.r = &r;
        rin::GraphIn in;
        in.mi = mi;
        in.pts = &pts3;
        in.tets = &tets4;
        in.iso_verts = &iv;
        in.mi_verts = &mv;
        in.faces = &faces;
        in.func_in_tet = &fit;
        in.start_index = &sidx;
        in.degenerate_vertex = &degen;
        in.found_degenerate = found;
        in.n_func = F;
        for (uint64_t f = 0; f < F; ++f) {
            in.sample_sign.push_back(vv[f] > 0);
            in.sample_value.push_back(vv[f]);
        }
        in.cuts = &cuts;
        rin::GraphOut g;
        bool ok = rin::run_graph(in, g);
        size_t corners = 0;
        for (auto& l : g.nm_edges_of_vert) corners += l.size() > 2;
        o << "{\"ok\":" << (ok ? "true" : "false") << ",\"error\":\"" << g.error << "\"";
        o << ",\"num_verts\":" << r.verts.size() << ",\"num_faces\":" << r.faces.size();
        o << ",\"num_intersecting\":" << r.num_intersecting << ",\"num_degenerate\":" << r.num_degenerate;
        o << ",\"buckets\":[" << r.nb[0] << "," << r.nb[1] << "," << r.nb[2] << "]";
        o << ",\"patches\":" << g.patches.size() << ",\"chains\":" << g.chains.size();
        o << ",\"shells\":" << g.shells.size() << ",\"cells\":" << g.cells.size();
        o << ",\"components\":" << g.num_components << ",\"corners\":" << corners << ",\"edges\":" << g.edges.size();
        o << ",\"patch_label\":";
        if (mi) {
            o << "[";
            for (size_t i = 0; i < g.patch_label_mi.size(); ++i)
                o << (i ? "," : "") << "[" << g.patch_label_mi[i].first << "," << g.patch_label_mi[i].second << "]";
            o << "]";
        } else {
            put_list(o, g.patch_label);
        }
        o << ",\"cell_label\":";
        if (mi) {
            put_list(o, g.cell_label_mi);
        } else {
            o << "[";
            for (size_t i = 0; i < g.cell_label.size(); ++i) {
                o << (i ? "," : "") << "[";
                for (size_t j = 0; j < g.cell_label[i].size(); ++j) o << (j ? "," : "") << (g.cell_label[i][j] ? 1 : 0);
                o << "]";
            }
            o << "]";
        }
        o << "}";
    } catch (std::exception& e) {
        o.str("");
        o << "{\"ok\":false,\"error\":\"exception: " << e.what() << "\"}";
    }
    g_out = o.str();
    return g_out.c_str();
}
