// Test harness glue: runs the oracle's hot path and feeds its outputs to the
// product's host graph stages, so the reference's known-answer integration
// tests (tests/test_implicit_networks.cpp:19-732) can pin the oracle on CPU.
#include <cstring>
#include <sstream>
#include <string>

#include "../../oracle/oracle.h"
#include "cut_view.h"
#include "graph.h"
#include "rin/io.h"
#ifdef RIN_HAVE_NLOHMANN
#include <fstream>

#include RIN_HAVE_NLOHMANN
#endif

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

// the oracle's hot path and the product's graph stages on one input
struct KA {
    oracle::Result r;
    std::vector<PolygonFace> faces;
    rin::GraphOut g;
    bool ok = false;
};

void compute(int mi, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, const double* vals, uint64_t F,
             KA& out) {
    std::vector<std::array<double, 3>> P(V);
    std::memcpy(P.data(), pts, V * 24);
    std::vector<std::array<uint64_t, 4>> Tt(T);
    std::memcpy(Tt.data(), tets, T * 32);
    std::vector<double> vv(vals, vals + V * F);
    oracle::Result& r = out.r;
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
    std::vector<PolygonFace>& faces = out.faces;
    faces.resize(r.faces.size());
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
    out.ok = rin::run_graph(in, out.g);
}

}  // namespace

extern "C" const char* ka_run(int mi, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T,
                              const double* vals, uint64_t F) {
    std::ostringstream o;
    try {
        KA ka;
        compute(mi, pts, V, tets, T, vals, F, ka);
        const oracle::Result& r = ka.r;
        const rin::GraphOut& g = ka.g;
        const bool ok = ka.ok;
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

// The same computation written to out_dir by this build's writers
// (include/rin/io.h: mesh.json, timings.json, stats.json) and, when the image
// has nlohmann::json, by a checker that builds the documents with nlohmann's
// own types the way src/io.cpp:154-312 and :556-590 describe them
// (mesh_nlohmann.json, timings_nlohmann.json, stats_nlohmann.json).
// Returns 0, or -1 on failure; 1 when the nlohmann files were written too.
extern "C" int ka_write(int mi, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, const double* vals,
                        uint64_t F, const char* out_dir) {
    try {
        KA ka;
        compute(mi, pts, V, tets, T, vals, F, ka);
        std::vector<std::array<double, 3>> xyz(ka.r.pts.begin(), ka.r.pts.end());
        const rin::GraphOut& g = ka.g;
        const std::string dir(out_dir);
        if (mi)
            save_result_MI(dir + "/mesh.json", xyz, ka.faces, g.patches, g.patch_label_mi, g.edges, g.chains,
                           g.nm_edges_of_vert, g.shells, g.cells, g.cell_label_mi);
        else
            save_result(dir + "/mesh.json", xyz, ka.faces, g.patches, g.patch_label, g.edges, g.chains,
                        g.nm_edges_of_vert, g.shells, g.cells, g.cell_label);
        std::vector<std::string> tl, sl;
        std::vector<double> tv;
        std::vector<size_t> sv;
        for (auto& t : g.timings) {
            tl.push_back(t.first);
            tv.push_back(t.second);
        }
        for (auto& t : ka.r.timings) {
            tl.push_back(t.first);
            tv.push_back(t.second);
        }
        for (auto& st : g.stats) {
            sl.push_back(st.first);
            sv.push_back(st.second);
        }
        sl.push_back("num_pts");
        sv.push_back(V);
        save_timings(dir + "/timings.json", tl, tv);
        save_statistics(dir + "/stats.json", sl, sv);
#ifdef RIN_HAVE_NLOHMANN
        using json = nlohmann::json;
        auto lists = [](const std::vector<std::vector<size_t>>& v) {
            json j;
            for (const auto& r : v) j.push_back(json(r));
            return j;
        };
        json jp, jf, je, jc, jl;
        for (const auto& p : xyz) jp.push_back(json(p));
        for (const auto& f : ka.faces) jf.push_back(json(f.vert_indices));
        for (const auto& e : g.edges) je.push_back({e.v1, e.v2});
        for (size_t i = 0; i < g.nm_edges_of_vert.size(); ++i)
            if (g.nm_edges_of_vert[i].size() > 2 || g.nm_edges_of_vert[i].size() == 1) jc.push_back(i);
        json jpl;
        if (mi) {
            jpl.push_back(json(g.patch_label_mi));
            jl.push_back(json(g.cell_label_mi));
        } else {
            jpl.push_back(json(g.patch_label));
            for (const auto& c : g.cell_label) jl.push_back(json(c));
        }
        json out;
        out["points"] = jp;
        out["faces"] = jf;
        out["patches"] = lists(g.patches);
        out["patches_label"] = jpl;
        out["edges"] = je;
        out["chains"] = lists(g.chains);
        out["corners"] = jc;
        out["shells"] = lists(g.shells);
        out["cells"] = lists(g.cells);
        out["cells_label"] = jl;
        std::ofstream(dir + "/mesh_nlohmann.json") << out << std::endl;
        json jt, js;
        for (size_t i = 0; i < tv.size(); ++i) jt[tl[i]] = tv[i];
        for (size_t i = 0; i < sv.size(); ++i) js[sl[i]] = sv[i];
        std::ofstream(dir + "/timings_nlohmann.json") << jt << std::endl;
        std::ofstream(dir + "/stats_nlohmann.json") << js << std::endl;
        return 1;
#else
        return 0;
#endif
    } catch (std::exception&) {
        return -1;
    }
}

// io.h entry points for the CPU tests
extern "C" int io_generate(uint64_t res, const double* lo, const double* hi, double* pts, uint64_t* tets) {
    std::vector<std::array<double, 3>> P;
    std::vector<std::array<size_t, 4>> Tt;
    if (!generate_tet_mesh(res, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}, P, Tt)) return -1;
    std::memcpy(pts, P.data(), P.size() * 24);
    for (size_t t = 0; t < Tt.size(); ++t)
        for (int c = 0; c < 4; ++c) tets[4 * t + c] = Tt[t][c];
    return 0;
}

thread_local std::vector<std::array<double, 3>> g_mesh_pts;
thread_local std::vector<std::array<size_t, 4>> g_mesh_tets;

// loads into thread-local storage; returns V and T, then io_mesh_copy fetches
extern "C" int io_load(const char* path, uint64_t* V, uint64_t* T) {
    if (!load_tet_mesh(path, g_mesh_pts, g_mesh_tets)) return -1;
    *V = g_mesh_pts.size();
    *T = g_mesh_tets.size();
    return 0;
}

extern "C" void io_mesh_copy(double* pts, uint64_t* tets) {
    std::memcpy(pts, g_mesh_pts.data(), g_mesh_pts.size() * 24);
    for (size_t t = 0; t < g_mesh_tets.size(); ++t)
        for (int c = 0; c < 4; ++c) tets[4 * t + c] = g_mesh_tets[t][c];
}

extern "C" const char* io_config(const char* path) {
    std::ostringstream o;
    try {
        Config c = parse_config_file(path);
        o << "{\"tet_mesh_file\":\"" << c.tet_mesh_file << "\",\"func_file\":\"" << c.func_file
          << "\",\"output_dir\":\"" << c.output_dir << "\",\"use_lookup\":" << c.use_lookup
          << ",\"use_secondary_lookup\":" << c.use_secondary_lookup
          << ",\"use_topo_ray_shooting\":" << c.use_topo_ray_shooting << ",\"res\":" << c.tet_mesh_resolution
          << ",\"bbox\":[" << c.tet_mesh_bbox_min[0] << "," << c.tet_mesh_bbox_min[1] << "," << c.tet_mesh_bbox_min[2]
          << "," << c.tet_mesh_bbox_max[0] << "," << c.tet_mesh_bbox_max[1] << "," << c.tet_mesh_bbox_max[2] << "]}";
    } catch (std::exception& e) {
        o.str("");
        o << "{\"error\":\"" << e.what() << "\"}";
    }
    g_out = o.str();
    return g_out.c_str();
}
