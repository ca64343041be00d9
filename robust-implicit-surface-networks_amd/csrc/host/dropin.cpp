// Drop-in implicit_arrangement() / material_interface() (declared in
// include/rin/implicit_arrangement.h): the per-tet path on the GPU through the
// C ABI, the graph stages on the host (graph.cpp).
//
// Driver order and reported labels follow src/implicit_arrangement.cpp:13-647
// and src/material_interface.cpp:13-694.
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "cut_view.h"
#include "graph.h"
#include "rin.h"
#include "rin/implicit_arrangement.h"
#include "simplicial_arrangement/lookup_table.h"
#include "thread_ctx.h"

namespace {

thread_local int t_device = 0;

// the grid of the last call, kept on the device: callers that pass the same
// mesh again (the same generate_tet_mesh grid, or the same explicit mesh) reuse
// its device tables
struct GridKey {
    bool regular = false;
    int res = 0;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    uint64_t V = 0, T = 0, hash = 0;
    bool operator==(const GridKey& o) const {
        return regular == o.regular && res == o.res && V == o.V && T == o.T && hash == o.hash &&
               std::memcmp(lo, o.lo, sizeof(lo)) == 0 && std::memcmp(hi, o.hi, sizeof(hi)) == 0;
    }
};

// one context per host thread (contexts are independent)
struct Ctx {
    rin_ctx* c = nullptr;
    int device = -1;
    rin_grid* grid = nullptr;
    GridKey key;
    ~Ctx() { reset(); }
    void reset() {
        if (grid) rin_grid_release(c, grid);
        grid = nullptr;
        if (c) rin_destroy(c);
        c = nullptr;
    }
    rin_ctx* get() {
        if (!c || device != t_device) {
            reset();
            if (rin_create(&c, t_device) != RIN_OK) throw std::runtime_error("rin_create failed");
            device = t_device;
        }
        return c;
    }
};
thread_local Ctx t_ctx;

// Is (pts, tets) exactly what generate_tet_mesh (src/io.cpp:95-152) makes for
// some resolution and box?  Then the device builds the grid itself (implicit
// tets, closed-form edge and face ids, star tables for the dedup).
bool regular_grid_of(const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<size_t, 4>>& tets,
                     GridKey& k) {
    const size_t T = tets.size();
    if (T == 0 || T % 5) return false;
    size_t r = (size_t)std::llround(std::cbrt((double)(T / 5)));
    while (r * r * r * 5 > T) --r;
    while ((r + 1) * (r + 1) * (r + 1) * 5 <= T) ++r;
    if (r == 0 || r * r * r * 5 != T || r > 1600) return false;
    const size_t N = r + 1;
    if (pts.size() != N * N * N) return false;
    const double* lo = pts.front().data();
    const double* hi = pts.back().data();
    auto coord = [&](size_t i, int a) { return (double(i) / double(N - 1)) * (hi[a] - lo[a]) + lo[a]; };
    for (size_t i = 0; i < N; ++i) {
        const double x = coord(i, 0);
        for (size_t j = 0; j < N; ++j) {
            const double y = coord(j, 1);
            for (size_t q = 0; q < N; ++q) {
                const auto& p = pts[(i * N + j) * N + q];
                if (p[0] != x || p[1] != y || p[2] != coord(q, 2)) return false;
            }
        }
    }
    static const int even[5][4] = {{4, 6, 1, 3}, {6, 3, 4, 7}, {1, 3, 0, 4}, {3, 1, 2, 6}, {4, 1, 6, 5}};
    static const int odd[5][4] = {{7, 0, 2, 5}, {2, 3, 0, 7}, {5, 7, 0, 4}, {7, 2, 6, 5}, {0, 1, 2, 5}};
    for (size_t i = 0; i < r; ++i)
        for (size_t j = 0; j < r; ++j)
            for (size_t q = 0; q < r; ++q) {
                const size_t b = (i * N + j) * N + q;
                const size_t v[8] = {b, b + N * N, b + N * N + N, b + N, b + 1, b + N * N + 1, b + N * N + N + 1,
                                     b + N + 1};
                const int(*tab)[4] = ((i + j + q) % 2 == 0) ? even : odd;
                const size_t t = ((i * r + j) * r + q) * 5;
                for (int e = 0; e < 5; ++e)
                    for (int c = 0; c < 4; ++c)
                        if (tets[t + e][c] != v[tab[e][c]]) return false;
            }
    k.regular = true;
    k.res = (int)r;
    for (int a = 0; a < 3; ++a) {
        k.lo[a] = lo[a];
        k.hi[a] = hi[a];
    }
    k.V = pts.size();
    k.T = T;
    return true;
}

uint64_t mix_words(const void* p, size_t bytes, uint64_t h) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 29;
    }
    for (; i < bytes; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
    return h;
}

struct Failure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(rin_ctx* c, int st, const char* what) {
    if (st != RIN_OK) throw Failure(std::string(what) + ": " + rin_last_error(c));
}

// RAII owners of the C ABI handles (the grid stays with the thread's context)
struct Handles {
    rin_ctx* c;
    rin_grid* g = nullptr;
    rin_values* v = nullptr;
    std::vector<rin_result*> res;
    explicit Handles(rin_ctx* ctx) : c(ctx) {}
    ~Handles() {
        for (auto* r : res) rin_result_release(c, r);
        if (v) rin_values_release(c, v);
    }
    rin_result* run(uint32_t flags) {
        rin_result* r = nullptr;
        check(c, rin_run(c, g, v, flags, &r), "rin_run");
        res.push_back(r);
        return r;
    }
};

// per-tet complexes fetched from the device result on demand, in batches
struct DeviceCuts : rin::CutSource {
    rin_ctx* c;
    const rin_result* r;
    std::unordered_map<size_t, rin::CutView> cache;
    DeviceCuts(rin_ctx* ctx, const rin_result* res) : c(ctx), r(res) {}
    void fetch(const std::vector<size_t>& tets) {
        std::vector<uint64_t> need;
        for (size_t t : tets)
            if (!cache.count(t)) need.push_back(t);
        if (need.empty()) return;
        std::vector<uint64_t> start(need.size() + 1);
        // records are at most a few thousand words; size generously, retry once
        std::vector<uint16_t> words(need.size() * 1024 + 64);
        int st = rin_result_records(c, r, need.data(), need.size(), start.data(), words.data(), words.size());
        if (st == RIN_ERR_ARG && std::string(rin_last_error(c)).find("too small") != std::string::npos) {
            words.assign(need.size() * 8192 + 64, 0);
            st = rin_result_records(c, r, need.data(), need.size(), start.data(), words.data(), words.size());
        }
        check(c, st, "rin_result_records");
        for (size_t i = 0; i < need.size(); ++i)
            if (start[i + 1] > start[i]) cache.emplace(need[i], rin::decode_cut(words.data() + start[i]));
    }
    const rin::CutView& get(size_t tet) override {
        auto it = cache.find(tet);
        if (it == cache.end()) {
            fetch({tet});
            it = cache.find(tet);
            if (it == cache.end()) throw std::runtime_error("cut requested for an inactive tet");
        }
        return it->second;
    }
    void prefetch(const std::vector<size_t>& tets) override { fetch(tets); }
};

bool is_row_major_ok(const rin::RowMajorValues& f, size_t n_pts) {
    return f.data != nullptr && f.rows == n_pts && f.cols > 0;
}

void append_stage_times(rin_ctx* c, const rin_result* r, std::vector<std::string>& labels,
                        std::vector<double>& secs) {
    const int n = rin_result_stage_times(c, r, nullptr, 0);
    std::vector<double> ms(n > 0 ? n : 0);
    if (n > 0) rin_result_stage_times(c, r, ms.data(), n);
    for (int i = 0; i < n; ++i) {
        labels.emplace_back(rin_result_stage_label(c, r, i));
        secs.push_back(ms[i] * 1e-3);
    }
}

// The device grid for (pts, tets) — the thread's cached one when the mesh is
// the last call's — and the function values uploaded onto it.
void setup(Handles& H, const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<size_t, 4>>& tets,
           const rin::RowMajorValues& funcVals) {
    rin_ctx* c = H.c;
    GridKey k;
    if (!regular_grid_of(pts, tets, k)) {
        k = GridKey{};
        k.V = pts.size();
        k.T = tets.size();
        k.hash = mix_words(pts.data(), pts.size() * sizeof(pts[0]), 0x12345);
        k.hash = mix_words(tets.data(), tets.size() * sizeof(tets[0]), k.hash);
    }
    if (!(t_ctx.grid && t_ctx.key == k)) {
        if (t_ctx.grid) rin_grid_release(c, t_ctx.grid);
        t_ctx.grid = nullptr;
        rin_grid* g = nullptr;
        if (k.regular) {
            check(c, rin_grid_regular(c, k.res, k.lo, k.hi, &g), "rin_grid_regular");
        } else {
            std::vector<uint64_t> t64(4 * tets.size());
            for (size_t t = 0; t < tets.size(); ++t)
                for (int q = 0; q < 4; ++q) t64[4 * t + q] = tets[t][q];
            check(c, rin_grid_explicit(c, &pts[0][0], pts.size(), t64.data(), tets.size(), &g), "rin_grid_explicit");
        }
        t_ctx.grid = g;
        t_ctx.key = k;
    }
    H.g = t_ctx.grid;
    check(c, rin_values_upload(c, H.g, funcVals.data, (uint32_t)funcVals.cols, &H.v), "rin_values_upload");
}

// -R check: per-tet (vertex, face, cell) counts forward vs reversed
// (src/implicit_arrangement.cpp:136-243, src/material_interface.cpp:174-287)
bool robust_check(Handles& H, uint32_t mode, std::vector<std::string>& tl, std::vector<double>& tv) {
    rin_result* fwd = nullptr;
    if (rin_run(H.c, H.g, H.v, mode | RIN_FLAG_CUT_ONLY, &fwd) != RIN_OK) return rin::robust_verdict(false, false, {}, {});
    H.res.push_back(fwd);
    append_stage_times(H.c, fwd, tl, tv);
    rin_result* rev = nullptr;
    if (rin_run(H.c, H.g, H.v, mode | RIN_FLAG_CUT_ONLY | RIN_FLAG_REVERSE, &rev) != RIN_OK)
        return rin::robust_verdict(true, false, {}, {});
    H.res.push_back(rev);
    rin_sizes s{};
    check(H.c, rin_result_sizes(H.c, fwd, &s), "rin_result_sizes");
    std::vector<uint64_t> a(3 * s.num_intersecting), b(3 * s.num_intersecting);
    check(H.c, rin_result_cut_counts(H.c, fwd, a.data()), "rin_result_cut_counts");
    check(H.c, rin_result_cut_counts(H.c, rev, b.data()), "rin_result_cut_counts");
    return rin::robust_verdict(true, true, a, b);
}

struct DeviceMesh {
    rin_sizes s{};
    std::vector<uint64_t> fit, start, rec11, vstart, verts, tfstart, tf, func;
    std::vector<double> xyz;
};

void fetch_mesh(rin_ctx* c, const rin_result* r, DeviceMesh& m) {
    check(c, rin_result_sizes(c, r, &m.s), "rin_result_sizes");
    m.fit.resize(m.s.n_func_in_tet + 1);
    m.start.resize(m.s.n_tets + 1);
    check(c, rin_result_csr(c, r, m.fit.data(), m.start.data()), "rin_result_csr");
    m.rec11.resize(11 * m.s.n_verts + 1);
    m.xyz.resize(3 * m.s.n_verts + 1);
    check(c, rin_result_verts(c, r, m.rec11.data(), m.xyz.data()), "rin_result_verts");
    m.vstart.resize(m.s.n_faces + 1);
    m.verts.resize(m.s.n_face_verts + 1);
    m.tfstart.resize(m.s.n_faces + 1);
    m.tf.resize(2 * m.s.n_tet_faces + 1);
    m.func.resize(2 * m.s.n_faces + 1);
    check(c, rin_result_faces(c, r, m.vstart.data(), m.verts.data(), m.tfstart.data(), m.tf.data(), m.func.data()),
          "rin_result_faces");
}

void to_faces(const DeviceMesh& m, std::vector<PolygonFace>& faces) {
    const size_t base = faces.size();
    faces.resize(base + m.s.n_faces);
    for (uint64_t f = 0; f < m.s.n_faces; ++f) {
        PolygonFace& P = faces[base + f];
        P.vert_indices.assign(m.verts.begin() + m.vstart[f], m.verts.begin() + m.vstart[f + 1]);
        for (uint64_t q = m.tfstart[f]; q < m.tfstart[f + 1]; ++q) P.tet_face_indices.push_back({m.tf[2 * q], m.tf[2 * q + 1]});
        P.func_index = {m.func[2 * f], m.func[2 * f + 1]};
    }
}

// iso-edges computed on the device (rin_result_edges), in compute_mesh_edges'
// form and order (src/mesh_connectivity.cpp:10-56)
void device_edges(rin_ctx* c, const rin_result* r, const DeviceMesh& m, std::vector<std::vector<size_t>>& eof,
                  std::vector<Edge>& edges) {
    uint64_t ne = 0;
    check(c, rin_result_edge_count(c, r, &ne), "rin_result_edge_count");
    const uint64_t no = m.s.n_face_verts;
    std::vector<uint64_t> v12(2 * ne + 1), st(ne + 1), occ(2 * no + 1), eo(no + 1);
    check(c, rin_result_edges(c, r, v12.data(), st.data(), occ.data(), eo.data()), "rin_result_edges");
    edges.resize(ne);
    for (uint64_t e = 0; e < ne; ++e) {
        Edge& E = edges[e];
        E.v1 = v12[2 * e];
        E.v2 = v12[2 * e + 1];
        E.face_edge_indices.resize(st[e + 1] - st[e]);
        for (uint64_t q = st[e], k = 0; q < st[e + 1]; ++q, ++k)
            E.face_edge_indices[k] = {occ[2 * q], occ[2 * q + 1]};
    }
    eof.resize(m.s.n_faces);
    for (uint64_t f = 0; f < m.s.n_faces; ++f) eof[f].assign(eo.begin() + m.vstart[f], eo.begin() + m.vstart[f + 1]);
}

// patches computed on the device (rin_result_patches), in compute_patches'
// form and order (src/mesh_connectivity.cpp:58-93)
void device_patches(rin_ctx* c, const rin_result* r, const DeviceMesh& m, std::vector<std::vector<size_t>>& patches) {
    uint64_t np = 0;
    check(c, rin_result_patch_count(c, r, &np), "rin_result_patch_count");
    std::vector<uint64_t> st(np + 1), fc(m.s.n_faces + 1);
    check(c, rin_result_patches(c, r, st.data(), fc.data()), "rin_result_patches");
    patches.resize(np);
    for (uint64_t p = 0; p < np; ++p) patches[p].assign(fc.begin() + st[p], fc.begin() + st[p + 1]);
}

void device_chains(rin_ctx* c, const rin_result* r, std::vector<std::vector<size_t>>& chains) {
    uint64_t nc = 0, ne = 0;
    check(c, rin_result_chain_count(c, r, &nc), "rin_result_chain_count");
    check(c, rin_result_edge_count(c, r, &ne), "rin_result_edge_count");
    std::vector<uint64_t> st(nc + 1), ed(ne + 1);
    check(c, rin_result_chains(c, r, st.data(), ed.data()), "rin_result_chains");
    chains.resize(nc);
    for (uint64_t i = 0; i < nc; ++i) chains[i].assign(ed.begin() + st[i], ed.begin() + st[i + 1]);
}

template <class T>
void push_stat(std::vector<std::string>& sl, std::vector<size_t>& sv, const char* label, T v) {
    sl.emplace_back(label);
    sv.push_back((size_t)v);
}

}  // namespace

namespace rin {
void set_device(int device) { t_device = device; }

bool robust_verdict(bool forward_ok, bool reverse_ok, const std::vector<uint64_t>& forward_counts,
                    const std::vector<uint64_t>& reverse_counts) {
    if (!forward_ok) {
        std::cout << "type 2 failure (crash in the normal order)." << std::endl;
        return false;
    }
    if (!reverse_ok) {
        std::cout << "type 3 failure (crash in the reverse order)." << std::endl;
        return false;
    }
    if (forward_counts != reverse_counts) {
        std::cout << "type 1 failure (inconsistency)." << std::endl;
        return false;
    }
    std::cout << "success." << std::endl;
    return true;
}
namespace detail {
rin_ctx* thread_context() { return t_ctx.get(); }
}  // namespace detail
}  // namespace rin

bool implicit_arrangement(bool robust_test, bool use_lookup, bool use_secondary_lookup, bool use_topo_ray_shooting,
                          const std::vector<std::array<double, 3>>& pts,
                          const std::vector<std::array<size_t, 4>>& tets, const rin::RowMajorValues& funcVals,
                          std::vector<std::array<double, 3>>& iso_pts, std::vector<PolygonFace>& iso_faces,
                          std::vector<std::vector<size_t>>& patches, std::vector<size_t>& patch_function_label,
                          std::vector<Edge>& iso_edges, std::vector<std::vector<size_t>>& chains,
                          std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                          std::vector<std::vector<size_t>>& shells,
                          std::vector<std::vector<size_t>>& arrangement_cells,
                          std::vector<std::vector<bool>>& cell_function_label, std::vector<std::string>& timing_labels,
                          std::vector<double>& timings, std::vector<std::string>& stats_labels,
                          std::vector<size_t>& stats) {
    (void)use_secondary_lookup;   // the 2-function table changes speed, not results
    (void)use_topo_ray_shooting;  // cells always come from topological ray shooting
    const size_t n_pts = pts.size(), n_tets = tets.size();
    std::cout << "tet mesh: " << n_pts << " verts, " << n_tets << " tets." << std::endl;
    push_stat(stats_labels, stats, "num_pts", n_pts);
    push_stat(stats_labels, stats, "num_tets", n_tets);
    if (!is_row_major_ok(funcVals, n_pts)) {
        std::cout << "funcVals must be (num_pts x num_funcs)" << std::endl;
        return false;
    }
    const size_t F = funcVals.cols;
    try {
        rin_ctx* c = t_ctx.get();
        Handles H(c);
        setup(H, pts, tets, funcVals);
        const uint32_t lookup =
            (use_lookup && simplicial_arrangement::lookup_table_enabled()) ? 0u : RIN_FLAG_NO_LOOKUP;
        if (robust_test) {
            const bool ok = robust_check(H, lookup, timing_labels, timings);
            rin_sizes s{};
            check(c, rin_result_sizes(c, H.res[0], &s), "rin_result_sizes");
            push_stat(stats_labels, stats, "num_degenerate_vertex", s.num_degenerate);
            push_stat(stats_labels, stats, "num_intersecting_tet", s.num_intersecting);
            push_stat(stats_labels, stats, "num_1_func", s.bucket[0]);
            push_stat(stats_labels, stats, "num_2_func", s.bucket[1]);
            push_stat(stats_labels, stats, "num_more_func", s.bucket[2]);
            return ok;
        }
        rin_result* r = H.run(RIN_FLAG_BATCHED | RIN_FLAG_KEEP_RECORDS | lookup);
        append_stage_times(c, r, timing_labels, timings);
        DeviceMesh m;
        fetch_mesh(c, r, m);
        std::cout << "num_degenerate_vertex = " << m.s.num_degenerate << std::endl;
        std::cout << "num_intersecting_tet = " << m.s.num_intersecting << std::endl;
        push_stat(stats_labels, stats, "num_degenerate_vertex", m.s.num_degenerate);
        push_stat(stats_labels, stats, "num_intersecting_tet", m.s.num_intersecting);
        push_stat(stats_labels, stats, "num_1_func", m.s.bucket[0]);
        push_stat(stats_labels, stats, "num_2_func", m.s.bucket[1]);
        push_stat(stats_labels, stats, "num_more_func", m.s.bucket[2]);
        std::cout << "num iso-vertices = " << m.s.n_verts << std::endl;
        std::cout << "num iso-faces = " << m.s.n_faces << std::endl;
        push_stat(stats_labels, stats, "num_iso_verts", m.s.n_verts);
        push_stat(stats_labels, stats, "num_iso_faces", m.s.n_faces);

        std::vector<IsoVert> iv(m.s.n_verts);
        for (uint64_t i = 0; i < m.s.n_verts; ++i) {
            const uint64_t* w = &m.rec11[11 * i];
            IsoVert& v = iv[i];
            v.tet_index = w[0];
            v.tet_vert_index = w[1];
            v.simplex_size = w[2];
            for (int q = 0; q < 4; ++q) v.simplex_vert_indices[q] = w[3 + q];
            for (int q = 0; q < 3; ++q) v.func_indices[q] = w[7 + q];
        }
        const size_t p0 = iso_pts.size();
        iso_pts.resize(p0 + m.s.n_verts);
        for (uint64_t i = 0; i < m.s.n_verts; ++i)
            for (int d = 0; d < 3; ++d) iso_pts[p0 + i][d] = m.xyz[3 * i + d];
        std::vector<PolygonFace> faces;
        to_faces(m, faces);

        // host graph stages (src/implicit_arrangement.cpp:404-645)
        std::vector<size_t> fit(m.fit.begin(), m.fit.begin() + m.s.n_func_in_tet);
        std::vector<size_t> sidx(m.start.begin(), m.start.end());
        std::vector<uint8_t> degen(n_pts, 0);
        bool found = false;
        for (size_t v = 0; v < n_pts; ++v)
            for (size_t f = 0; f < F; ++f)
                if (funcVals.data[v * F + f] == 0) {
                    degen[v] = 1;
                    found = true;
                    break;
                }
        DeviceCuts cuts(c, r);
        rin::GraphIn in;
        in.mi = false;
        in.pts = &pts;
        in.tets = &tets;
        in.iso_verts = &iv;
        in.faces = &faces;
        in.func_in_tet = &fit;
        in.start_index = &sidx;
        in.degenerate_vertex = &degen;
        in.found_degenerate = found;
        in.n_func = F;
        for (size_t f = 0; f < F; ++f) in.sample_sign.push_back(funcVals.data[f] > 0);
        in.cuts = &cuts;
        in.edges_from = [&](std::vector<std::vector<size_t>>& eof, std::vector<Edge>& e) { device_edges(c, r, m, eof, e); };
        in.patches_from = [&](std::vector<std::vector<size_t>>& p) { device_patches(c, r, m, p); };
        in.chains_from = [&](std::vector<std::vector<size_t>>& ch) { device_chains(c, r, ch); };
        rin::GraphOut g;
        const bool ok = rin::run_graph(in, g);
        for (auto& t : g.timings) {
            timing_labels.push_back(t.first);
            timings.push_back(t.second);
        }
        for (auto& s : g.stats) push_stat(stats_labels, stats, s.first.c_str(), s.second);
        iso_faces.insert(iso_faces.end(), faces.begin(), faces.end());
        iso_edges.insert(iso_edges.end(), g.edges.begin(), g.edges.end());
        patches.insert(patches.end(), g.patches.begin(), g.patches.end());
        patch_function_label.insert(patch_function_label.end(), g.patch_label.begin(), g.patch_label.end());
        chains.insert(chains.end(), g.chains.begin(), g.chains.end());
        non_manifold_edges_of_vert.insert(non_manifold_edges_of_vert.end(), g.nm_edges_of_vert.begin(),
                                          g.nm_edges_of_vert.end());
        shells.insert(shells.end(), g.shells.begin(), g.shells.end());
        arrangement_cells.insert(arrangement_cells.end(), g.cells.begin(), g.cells.end());
        cell_function_label.insert(cell_function_label.end(), g.cell_label.begin(), g.cell_label.end());
        if (!ok) std::cout << g.error << std::endl;
        return ok;
    } catch (std::exception& e) {
        std::cout << e.what() << std::endl;
        return false;
    }
}

bool material_interface(bool robust_test, bool use_lookup, bool use_secondary_lookup, bool use_topo_ray_shooting,
                        const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<size_t, 4>>& tets,
                        const rin::RowMajorValues& funcVals, std::vector<std::array<double, 3>>& MI_pts,
                        std::vector<PolygonFace>& MI_faces, std::vector<std::vector<size_t>>& patches,
                        std::vector<std::pair<size_t, size_t>>& patch_function_label, std::vector<Edge>& MI_edges,
                        std::vector<std::vector<size_t>>& chains,
                        std::vector<std::vector<size_t>>& non_manifold_edges_of_vert,
                        std::vector<std::vector<size_t>>& shells, std::vector<std::vector<size_t>>& material_cells,
                        std::vector<size_t>& cell_function_label, std::vector<std::string>& timing_labels,
                        std::vector<double>& timings, std::vector<std::string>& stats_labels,
                        std::vector<size_t>& stats) {
    (void)use_secondary_lookup;
    (void)use_topo_ray_shooting;
    const size_t n_pts = pts.size(), n_tets = tets.size();
    std::cout << "tet mesh: " << n_pts << " verts, " << n_tets << " tets." << std::endl;
    push_stat(stats_labels, stats, "num_pts", n_pts);
    push_stat(stats_labels, stats, "num_tets", n_tets);
    if (!is_row_major_ok(funcVals, n_pts)) {
        std::cout << "funcVals must be (num_pts x num_funcs)" << std::endl;
        return false;
    }
    const size_t F = funcVals.cols;
    try {
        rin_ctx* c = t_ctx.get();
        Handles H(c);
        setup(H, pts, tets, funcVals);
        const uint32_t mode =
            RIN_FLAG_MI | ((use_lookup && simplicial_arrangement::lookup_table_enabled()) ? 0u : RIN_FLAG_NO_LOOKUP);
        if (robust_test) {
            const bool ok = robust_check(H, mode, timing_labels, timings);
            rin_sizes s{};
            check(c, rin_result_sizes(c, H.res[0], &s), "rin_result_sizes");
            push_stat(stats_labels, stats, "num_intersecting_tet", s.num_intersecting);
            push_stat(stats_labels, stats, "num_2_func", s.bucket[0]);
            push_stat(stats_labels, stats, "num_3_func", s.bucket[1]);
            push_stat(stats_labels, stats, "num_more_func", s.bucket[2]);
            return ok;
        }
        rin_result* r = H.run(RIN_FLAG_BATCHED | RIN_FLAG_KEEP_RECORDS | mode);
        append_stage_times(c, r, timing_labels, timings);
        DeviceMesh m;
        fetch_mesh(c, r, m);
        push_stat(stats_labels, stats, "num_intersecting_tet", m.s.num_intersecting);
        push_stat(stats_labels, stats, "num_2_func", m.s.bucket[0]);
        push_stat(stats_labels, stats, "num_3_func", m.s.bucket[1]);
        push_stat(stats_labels, stats, "num_more_func", m.s.bucket[2]);
        std::cout << "num MI-vertices = " << m.s.n_verts << std::endl;
        std::cout << "num MI-faces = " << m.s.n_faces << std::endl;
        push_stat(stats_labels, stats, "num_MI_verts", m.s.n_verts);
        push_stat(stats_labels, stats, "num_MI_faces", m.s.n_faces);

        std::vector<MI_Vert> mv(m.s.n_verts);
        for (uint64_t i = 0; i < m.s.n_verts; ++i) {
            const uint64_t* w = &m.rec11[11 * i];
            MI_Vert& v = mv[i];
            v.tet_index = w[0];
            v.tet_vert_index = w[1];
            v.simplex_size = w[2];
            for (int q = 0; q < 4; ++q) {
                v.simplex_vert_indices[q] = w[3 + q];
                v.material_indices[q] = w[7 + q];
            }
        }
        const size_t p0 = MI_pts.size();
        MI_pts.resize(p0 + m.s.n_verts);
        for (uint64_t i = 0; i < m.s.n_verts; ++i)
            for (int d = 0; d < 3; ++d) MI_pts[p0 + i][d] = m.xyz[3 * i + d];
        std::vector<PolygonFace> faces;
        to_faces(m, faces);

        // host graph stages (src/material_interface.cpp:449-694)
        std::vector<size_t> fit(m.fit.begin(), m.fit.begin() + m.s.n_func_in_tet);
        std::vector<size_t> sidx(m.start.begin(), m.start.end());
        // degenerate vertex: tied maximum (src/material_interface.cpp:53-92)
        std::vector<uint8_t> degen(n_pts, 0);
        bool found = false;
        for (size_t v = 0; v < n_pts; ++v) {
            const double* row = funcVals.data + v * F;
            size_t best = 0;
            for (size_t f = 1; f < F; ++f)
                if (row[f] > row[best]) best = f;
            for (size_t f = 0; f < F; ++f)
                if (f != best && row[f] == row[best]) {
                    degen[v] = 1;
                    found = true;
                    break;
                }
        }
        DeviceCuts cuts(c, r);
        rin::GraphIn in;
        in.mi = true;
        in.pts = &pts;
        in.tets = &tets;
        in.mi_verts = &mv;
        in.faces = &faces;
        in.func_in_tet = &fit;
        in.start_index = &sidx;
        in.degenerate_vertex = &degen;
        in.found_degenerate = found;
        in.n_func = F;
        for (size_t f = 0; f < F; ++f) in.sample_value.push_back(funcVals.data[f]);
        in.cuts = &cuts;
        in.edges_from = [&](std::vector<std::vector<size_t>>& eof, std::vector<Edge>& e) { device_edges(c, r, m, eof, e); };
        in.patches_from = [&](std::vector<std::vector<size_t>>& p) { device_patches(c, r, m, p); };
        in.chains_from = [&](std::vector<std::vector<size_t>>& ch) { device_chains(c, r, ch); };
        rin::GraphOut g;
        const bool ok = rin::run_graph(in, g);
        for (auto& t : g.timings) {
            timing_labels.push_back(t.first);
            timings.push_back(t.second);
        }
        for (auto& s : g.stats) push_stat(stats_labels, stats, s.first.c_str(), s.second);
        MI_faces.insert(MI_faces.end(), faces.begin(), faces.end());
        MI_edges.insert(MI_edges.end(), g.edges.begin(), g.edges.end());
        patches.insert(patches.end(), g.patches.begin(), g.patches.end());
        patch_function_label.insert(patch_function_label.end(), g.patch_label_mi.begin(), g.patch_label_mi.end());
        chains.insert(chains.end(), g.chains.begin(), g.chains.end());
        non_manifold_edges_of_vert.insert(non_manifold_edges_of_vert.end(), g.nm_edges_of_vert.begin(),
                                          g.nm_edges_of_vert.end());
        shells.insert(shells.end(), g.shells.begin(), g.shells.end());
        material_cells.insert(material_cells.end(), g.cells.begin(), g.cells.end());
        cell_function_label.insert(cell_function_label.end(), g.cell_label_mi.begin(), g.cell_label_mi.end());
        if (!ok) std::cout << g.error << std::endl;
        return ok;
    } catch (std::exception& e) {
        std::cout << e.what() << std::endl;
        return false;
    }
}
