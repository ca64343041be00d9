// Host graph stages; each function names the reference lines whose behaviour
// (traversal order, tie-breaking, error cases) it keeps.
#include "graph.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <functional>
#include <map>
#include <set>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace rin {

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct PairHash {
    size_t operator()(const std::pair<size_t, size_t>& p) const {
        return std::hash<size_t>()(p.first * 0x9E3779B97F4A7C15ull ^ (p.second + 0x632BE59BD9B4E019ull));
    }
};
template <size_t N>
struct ArrHash {
    size_t operator()(const std::array<size_t, N>& a) const {
        size_t h = 0xcbf29ce484222325ull;
        for (size_t x : a) h = (h ^ x) * 0x100000001b3ull;
        return h;
    }
};
using PairMap = std::unordered_map<std::pair<size_t, size_t>, size_t, PairHash>;

// (min, second-min, max) of a polygon's vertex ids (src/extract_mesh.h:68-92)
template <typename T>
std::array<T, 3> face_key(const std::vector<T>& vs) {
    T mn = vs[0], mx = vs[0];
    size_t mp = 0;
    for (size_t i = 1; i < vs.size(); ++i) {
        if (vs[i] < mn) {
            mn = vs[i];
            mp = i;
        } else if (vs[i] > mx) {
            mx = vs[i];
        }
    }
    T sm = mx + 1;
    for (size_t i = 0; i < vs.size(); ++i)
        if (i != mp && vs[i] < sm) sm = vs[i];
    return {mn, sm, mx};
}

// src/topo_ray_shooting.h:127-138
inline bool xyz_less(const std::array<double, 3>& p, const std::array<double, 3>& q) {
    if (p[0] != q[0]) return p[0] < q[0];
    if (p[1] != q[1]) return p[1] < q[1];
    return p[2] < q[2];
}

}  // namespace

// ------------------------------------------------------------------ basic stages

void compute_mesh_edges(const std::vector<PolygonFace>& faces, std::vector<std::vector<size_t>>& edges_of_face,
                        std::vector<Edge>& edges) {
    PairMap id;
    size_t total = 0;
    for (const auto& f : faces) total += f.vert_indices.size();
    id.reserve(total);
    edges.reserve(total / 2);
    edges_of_face.reserve(faces.size());
    for (size_t i = 0; i < faces.size(); ++i) {
        const auto& vs = faces[i].vert_indices;
        const size_t n = vs.size();
        edges_of_face.emplace_back(n);
        for (size_t j = 0; j < n; ++j) {
            size_t a = vs[j], b = vs[(j + 1) % n];
            if (a > b) std::swap(a, b);
            auto it = id.try_emplace({a, b}, edges.size());
            if (it.second) {
                edges.push_back(Edge{a, b, {}});
            }
            edges[it.first->second].face_edge_indices.emplace_back(i, j);
            edges_of_face.back()[j] = it.first->second;
        }
    }
}

void compute_patches(const std::vector<std::vector<size_t>>& edges_of_face, const std::vector<Edge>& edges,
                     const std::vector<PolygonFace>& faces, std::vector<std::vector<size_t>>& patches,
                     std::vector<std::pair<size_t, size_t>>& labels) {
    std::vector<char> seen(edges_of_face.size(), 0);
    for (size_t i = 0; i < edges_of_face.size(); ++i) {
        if (seen[i]) continue;
        patches.emplace_back();
        auto& patch = patches.back();
        labels.push_back(faces[i].func_index);
        std::deque<size_t> Q{i};
        patch.push_back(i);
        seen[i] = 1;
        while (!Q.empty()) {
            size_t f = Q.front();
            Q.pop_front();
            for (size_t e : edges_of_face[f]) {
                const auto& fe = edges[e].face_edge_indices;
                if (fe.size() != 2) continue;  // only manifold edges join faces
                size_t g = (fe[0].first == f) ? fe[1].first : fe[0].first;
                if (!seen[g]) {
                    seen[g] = 1;
                    patch.push_back(g);
                    Q.push_back(g);
                }
            }
        }
    }
}

void compute_chains(const std::vector<Edge>& edges, const std::vector<std::vector<size_t>>& nm,
                    std::vector<std::vector<size_t>>& chains) {
    std::vector<char> seen(edges.size(), 0);
    for (size_t i = 0; i < edges.size(); ++i) {
        if (seen[i] || edges[i].face_edge_indices.size() <= 2) continue;
        chains.emplace_back();
        auto& chain = chains.back();
        std::deque<size_t> Q{i};
        chain.push_back(i);
        seen[i] = 1;
        while (!Q.empty()) {
            size_t e = Q.front();
            Q.pop_front();
            for (size_t v : {edges[e].v1, edges[e].v2}) {
                if (nm[v].size() != 2) continue;  // chains stop at corners
                size_t o = (nm[v][0] == e) ? nm[v][1] : nm[v][0];
                if (!seen[o]) {
                    seen[o] = 1;
                    chain.push_back(o);
                    Q.push_back(o);
                }
            }
        }
    }
}

void compute_shells_and_components(size_t np, const std::vector<std::vector<HalfPair>>& hpp,
                                   std::vector<std::vector<size_t>>& shells, std::vector<size_t>& shell_of_hp,
                                   std::vector<std::vector<size_t>>& comps, std::vector<size_t>& comp_of_patch) {
    // half-patch (p, +1) is 2p, (p, -1) is 2p+1
    std::vector<std::vector<size_t>> adj(2 * np);
    for (const auto& list : hpp)
        for (const auto& pr : list) {
            size_t a = 2 * pr.first.first + (pr.first.second == 1 ? 0 : 1);
            size_t b = 2 * pr.second.first + (pr.second.second == 1 ? 0 : 1);
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
    shells.clear();
    shell_of_hp.assign(2 * np, 0);
    std::vector<char> seen(2 * np, 0);
    for (size_t i = 0; i < 2 * np; ++i) {
        if (seen[i]) continue;
        size_t sid = shells.size();
        shells.emplace_back();
        std::deque<size_t> Q{i};
        shells.back().push_back(i);
        shell_of_hp[i] = sid;
        seen[i] = 1;
        while (!Q.empty()) {
            size_t h = Q.front();
            Q.pop_front();
            for (size_t o : adj[h])
                if (!seen[o]) {
                    seen[o] = 1;
                    shells.back().push_back(o);
                    shell_of_hp[o] = sid;
                    Q.push_back(o);
                }
        }
    }
    comps.clear();
    comp_of_patch.assign(np, 0);
    std::vector<char> pseen(np, 0);
    for (size_t i = 0; i < np; ++i) {
        if (pseen[i]) continue;
        size_t cid = comps.size();
        comps.emplace_back();
        std::deque<size_t> Q{i};
        comps.back().push_back(i);
        comp_of_patch[i] = cid;
        pseen[i] = 1;
        while (!Q.empty()) {
            size_t p = Q.front();
            Q.pop_front();
            for (size_t side = 0; side < 2; ++side)
                for (size_t o : adj[2 * p + side]) {
                    size_t q = o / 2;
                    if (!pseen[q]) {
                        pseen[q] = 1;
                        comps.back().push_back(q);
                        comp_of_patch[q] = cid;
                        Q.push_back(q);
                    }
                }
        }
    }
}

void compute_arrangement_cells(size_t ns, const std::vector<std::pair<size_t, size_t>>& links,
                               std::vector<std::vector<size_t>>& cells) {
    // shell `ns` stands for the sink (the unbounded side)
    std::vector<std::vector<size_t>> adj(ns + 1);
    for (const auto& l : links) {
        size_t a = (l.first == Mesh_None) ? ns : l.first;
        size_t b = (l.second == Mesh_None) ? ns : l.second;
        adj[a].push_back(b);
        adj[b].push_back(a);
    }
    std::vector<char> seen(ns + 1, 0);
    for (size_t i = 0; i <= ns; ++i) {
        if (seen[i]) continue;
        std::vector<size_t> comp{i};
        std::deque<size_t> Q{i};
        seen[i] = 1;
        while (!Q.empty()) {
            size_t s = Q.front();
            Q.pop_front();
            for (size_t o : adj[s])
                if (!seen[o]) {
                    seen[o] = 1;
                    comp.push_back(o);
                    Q.push_back(o);
                }
        }
        std::vector<size_t> cell;
        for (size_t s : comp)
            if (s < ns) cell.push_back(s);
        cells.push_back(cell);
    }
}

std::vector<std::vector<bool>> sign_propagation(const std::vector<std::vector<size_t>>& cells,
                                                const std::vector<size_t>& shell_of_hp,
                                                const std::vector<std::vector<size_t>>& shells,
                                                const std::vector<size_t>& plabel, size_t n_func,
                                                const std::vector<bool>& sample) {
    std::vector<size_t> cell_of_shell(shells.size(), 0);
    for (size_t i = 0; i < cells.size(); ++i)
        for (size_t s : cells[i]) cell_of_shell[s] = i;
    std::vector<std::vector<bool>> label(cells.size(), std::vector<bool>(n_func, false));
    std::vector<char> visited(cells.size(), 0);
    std::vector<char> fseen(n_func, 0);
    std::vector<std::vector<size_t>> pending(n_func);
    std::deque<size_t> Q;
    if (!cells.empty()) Q.push_back(0);
    while (!Q.empty()) {
        size_t ci = Q.front();
        Q.pop_front();
        if (visited[ci]) continue;
        visited[ci] = 1;
        std::vector<bool>& cur = label[ci];
        // neighbouring cells in first-seen order, last write of (function, side) wins
        std::vector<std::pair<size_t, std::pair<size_t, bool>>> nbr;
        for (size_t s : cells[ci])
            for (size_t hp : shells[s]) {
                size_t fn = plabel[hp / 2];
                bool sgn = (hp % 2 == 0);
                size_t opp = cell_of_shell[shell_of_hp[sgn ? hp + 1 : hp - 1]];
                bool found = false;
                for (auto& nb : nbr)
                    if (nb.first == opp) {
                        nb.second = {fn, sgn};
                        found = true;
                    }
                if (!found) nbr.push_back({opp, {fn, sgn}});
                if (fseen[fn] && cur[fn] != sgn) throw std::runtime_error("ERROR: Inconsistent Cell Function Labels.");
                cur[fn] = sgn;
                fseen[fn] = 1;
                for (size_t c : pending[fn]) label[c][fn] = sgn;
                pending[fn].clear();
            }
        for (size_t fn = 0; fn < n_func; ++fn) {
            if (fseen[fn]) continue;
            pending[fn].push_back(ci);
            for (auto& nb : nbr) pending[fn].push_back(nb.first);
        }
        for (auto& nb : nbr) {
            if (!visited[nb.first]) Q.push_back(nb.first);
            label[nb.first] = cur;
            label[nb.first][nb.second.first] = !nb.second.second;
        }
    }
    for (size_t fn = 0; fn < n_func; ++fn)
        if (!fseen[fn])
            for (auto& l : label) l[fn] = sample[fn];
    return label;
}

std::vector<size_t> sign_propagation_MI(const std::vector<std::vector<size_t>>& cells,
                                        const std::vector<size_t>& shell_of_hp,
                                        const std::vector<std::vector<size_t>>& shells,
                                        const std::vector<std::pair<size_t, size_t>>& plabel, size_t n_func,
                                        const std::vector<double>& sample) {
    (void)shell_of_hp;
    (void)n_func;
    std::vector<size_t> label(cells.size(), Mesh_None);
    for (size_t ci = 0; ci < cells.size(); ++ci) {
        for (size_t s : cells[ci]) {
            if (s == Mesh_None) {
                // the whole box is one material: the maximum at the sample vertex
                size_t id = 0;
                double mx = sample[0];
                for (size_t i = 1; i < sample.size(); ++i)
                    if (sample[i] > mx) {
                        mx = sample[i];
                        id = i;
                    }
                label[ci] = id;
                break;
            }
            for (size_t hp : shells[s]) {
                size_t fn = (hp % 2 == 0) ? plabel[hp / 2].first : plabel[hp / 2].second;
                if (label[ci] == Mesh_None)
                    label[ci] = fn;
                else if (label[ci] != fn)
                    throw std::runtime_error("ERROR: Inconsistent Cell Function Labels.");
            }
        }
    }
    return label;
}

// ------------------------------------------------------------------ face order

namespace {

// Walk around an iso-edge inside one tet (src/pair_faces.cpp:138-352): start at
// the first incident face, go to its positive side, then to its negative side.
void pair_in_one_tet(const CutView& C, const std::vector<PolygonFace>& faces, const Edge& e,
                     std::vector<HalfPair>& out) {
    const size_t nf = C.faces.size();
    std::vector<char> inc(nf, 0);
    std::vector<size_t> iso_of(nf, Mesh_None);
    for (const auto& fe : e.face_edge_indices) {
        size_t lf = faces[fe.first].tet_face_indices[0].second;
        inc[lf] = 1;
        iso_of[lf] = fe.first;
    }
    // cell lookup for either kind of complex
    std::vector<size_t> cell_of_label;
    if (C.mi) {
        size_t mx = 0;
        for (const auto& c : C.cells) mx = std::max(mx, c.material_label);
        cell_of_label.assign(mx + 1, Mesh_None);
        for (size_t i = 0; i < C.cells.size(); ++i) cell_of_label[C.cells[i].material_label] = i;
    }
    auto pos_cell = [&](size_t f) -> size_t {
        if (!C.mi) return C.faces[f].positive_cell;
        size_t l = C.faces[f].positive_material_label;
        return l < cell_of_label.size() ? cell_of_label[l] : Mesh_None;
    };
    auto neg_cell = [&](size_t f) -> size_t {
        if (!C.mi) return C.faces[f].negative_cell;
        size_t l = C.faces[f].negative_material_label;
        return l < cell_of_label.size() ? cell_of_label[l] : Mesh_None;
    };
    std::vector<char> visited(C.cells.size(), 0);
    for (int dir = 0; dir < 2; ++dir) {
        size_t iso1 = e.face_edge_indices[0].first;
        size_t f1 = faces[iso1].tet_face_indices[0].second;
        int s1 = dir == 0 ? 1 : -1;
        size_t cell = dir == 0 ? pos_cell(f1) : neg_cell(f1);
        while (cell != Mesh_None && !visited[cell]) {
            visited[cell] = 1;
            size_t f2 = Mesh_None;
            for (size_t g : C.cells[cell].faces)
                if (inc[g] && g != f1) f2 = g;
            if (f2 == Mesh_None) break;
            int s2;
            if (pos_cell(f2) == cell) {
                cell = neg_cell(f2);
                s2 = 1;
            } else {
                cell = pos_cell(f2);
                s2 = -1;
            }
            out.push_back({{iso1, s1}, {iso_of[f2], s2}});
            f1 = f2;
            iso1 = iso_of[f2];
            s1 = -s2;
        }
    }
}

using TF = std::pair<size_t, size_t>;  // (tet, local face)

// Walk around an iso-edge on a tet edge or tet face shared by several tets
// (src/pair_faces.cpp:354-718).
void pair_in_tets(const GraphIn& in, const Edge& e, const std::vector<size_t>& simplex,
                  const std::vector<size_t>& tet_ids, std::vector<HalfPair>& out) {
    const auto& tets = *in.tets;
    const auto& faces = *in.faces;
    const auto& fit = *in.func_in_tet;
    const auto& sidx = *in.start_index;
    std::unordered_map<TF, size_t, PairHash> iso_of;
    for (const auto& fe : e.face_edge_indices)
        for (const auto& tf : faces[fe.first].tet_face_indices) iso_of[tf] = fe.first;
    // tet planes that are the same triangle, seen from two tets
    std::unordered_map<TF, TF, PairHash> same_plane;
    if (simplex.size() == 3) {
        TF a, b;
        for (int q = 0; q < 2; ++q) {
            size_t t = tet_ids[q];
            size_t p = 0;
            for (size_t i = 0; i < 4; ++i) {
                size_t v = tets[t][i];
                if (v != simplex[0] && v != simplex[1] && v != simplex[2]) {
                    p = i;
                    break;
                }
            }
            (q == 0 ? a : b) = {t, p};
        }
        same_plane[a] = b;
        same_plane[b] = a;
    } else {
        std::unordered_map<std::array<size_t, 3>, TF, ArrHash<3>> plane_of_tri;
        for (size_t t : tet_ids)
            for (size_t i = 0; i < 4; ++i) {
                size_t v = tets[t][i];
                if (v == simplex[0] || v == simplex[1]) continue;
                std::array<size_t, 3> tri{tets[t][(i + 1) % 4], tets[t][(i + 2) % 4], tets[t][(i + 3) % 4]};
                std::sort(tri.begin(), tri.end());
                auto it = plane_of_tri.try_emplace(tri, TF{t, i});
                if (!it.second) {
                    same_plane[{t, i}] = it.first->second;
                    same_plane[it.first->second] = {t, i};
                }
            }
    }
    // pair the faces lying on those shared planes
    std::unordered_map<TF, TF, PairHash> opposite;
    size_t nb = 0;
    std::unordered_map<size_t, size_t> on_vert;
    std::unordered_map<std::array<size_t, 3>, size_t, ArrHash<3>> on_edge;
    std::unordered_map<std::array<size_t, 5>, size_t, ArrHash<5>> on_face;
    std::unordered_map<std::array<size_t, 3>, TF, ArrHash<3>> on_tri;
    for (size_t t : tet_ids) {
        const CutView* Cp = nullptr;
        if (in.cuts && sidx[t + 1] > sidx[t]) Cp = &in.cuts->get(t);
        if (!Cp) {
            // an empty tet: its four faces are whole triangles.  The reference
            // increments the wrong counter in this loop
            // (src/pair_faces.cpp:455); an edge whose two ends are zeros of a
            // function makes that function active in every incident tet, so the
            // loop is not reached.
            throw std::runtime_error("empty tet around a degenerate iso-edge");
        }
        const CutView& C = *Cp;
        const size_t s0 = sidx[t];
        std::vector<char> bv(C.vertices.size(), 0), bf(C.faces.size(), 0);
        for (size_t f = 0; f < C.faces.size(); ++f) {
            size_t sp = C.faces[f].supporting_plane;
            if (sp < 4 && same_plane.count({t, sp})) {
                bf[f] = 1;
                for (size_t v : C.faces[f].vertices) bv[v] = 1;
            }
        }
        std::vector<size_t> bid(C.vertices.size(), Mesh_None);
        for (size_t j = 0; j < C.vertices.size(); ++j) {
            if (!bv[j]) continue;
            size_t impl[3], bnd[3];
            int ni = 0, nbd = 0;
            for (int q = 0; q < 3; ++q) {
                size_t p = C.vertices[j][q];
                if (p > 3)
                    impl[ni++] = fit[p - 4 + s0];
                else
                    bnd[nbd++] = p;
            }
            if (nbd == 2) {
                size_t ev[2];
                int ne = 0;
                for (size_t c = 0; c < 4; ++c)
                    if (c != bnd[0] && c != bnd[1]) ev[ne++] = tets[t][c];
                std::array<size_t, 3> key{std::min(ev[0], ev[1]), std::max(ev[0], ev[1]), impl[0]};
                auto it = on_edge.try_emplace(key, nb);
                if (it.second) ++nb;
                bid[j] = it.first->second;
            } else if (nbd == 1) {
                size_t fv[3];
                int ne = 0;
                for (size_t c = 0; c < 4; ++c)
                    if (c != bnd[0]) fv[ne++] = tets[t][c];
                std::sort(fv, fv + 3);
                std::array<size_t, 5> key{fv[0], fv[1], fv[2], impl[0], impl[1]};
                auto it = on_face.try_emplace(key, nb);
                if (it.second) ++nb;
                bid[j] = it.first->second;
            } else if (nbd == 3) {
                size_t c = 0;
                while (c == bnd[0] || c == bnd[1] || c == bnd[2]) ++c;
                auto it = on_vert.try_emplace(tets[t][c], nb);
                if (it.second) ++nb;
                bid[j] = it.first->second;
            }
        }
        for (size_t f = 0; f < C.faces.size(); ++f) {
            if (!bf[f]) continue;
            std::vector<size_t> fvs;
            for (size_t v : C.faces[f].vertices) fvs.push_back(bid[v]);
            auto it = on_tri.try_emplace(face_key(fvs), TF{t, f});
            if (!it.second) {
                opposite[it.first->second] = {t, f};
                opposite[{t, f}] = it.first->second;
            }
        }
    }
    auto cut_of = [&](size_t t) -> const CutView& { return in.cuts->get(t); };
    // orientation of a local face relative to the smallest implicit plane of its group
    auto half_iso = [&](TF tf, int orient, size_t& iso, int& iso_orient) {
        iso = iso_of.at(tf);
        const CutView& C = cut_of(tf.first);
        size_t sp = C.faces[tf.second].supporting_plane;
        if (sp > 3) {
            iso_orient = orient;
            return;
        }
        size_t uid = C.unique_plane_indices.at(sp);
        size_t mp = Mesh_None;
        for (size_t p : C.unique_planes[uid])
            if (p > 3 && p < mp) mp = p;
        iso_orient = (C.unique_plane_orientations.at(mp) != C.unique_plane_orientations.at(sp)) ? -orient : orient;
    };
    auto opp = [&](TF tf) -> TF {
        auto it = opposite.find(tf);
        if (it == opposite.end()) throw std::runtime_error("unpaired tet-boundary face around iso-edge");
        return it->second;
    };
    // next face around the edge from (face, side)
    std::function<void(TF, int, TF&, int&)> next = [&](TF f, int orient, TF& fn, int& on) {
        const CutView& C = cut_of(f.first);
        size_t cell = orient == 1 ? C.faces[f.second].positive_cell : C.faces[f.second].negative_cell;
        if (cell == Mesh_None) {
            next(opp(f), 1, fn, on);
            return;
        }
        for (size_t g : C.cells[cell].faces) {
            if (g == f.second) continue;
            if (iso_of.count({f.first, g}) || opposite.count({f.first, g})) {
                fn = {f.first, g};
                break;
            }
        }
        on = (C.faces[fn.second].positive_cell == cell) ? 1 : -1;
    };
    size_t iso0 = e.face_edge_indices[0].first;
    TF cur = faces[iso0].tet_face_indices[0];
    int ocur = 1;
    size_t iso_next = Mesh_None;
    size_t guard = 0;
    while (iso_next != iso0) {
        if (++guard > 4 * e.face_edge_indices.size() + 16) throw std::runtime_error("face order walk did not close");
        TF nf;
        int on;
        next(cur, ocur, nf, on);
        size_t g2 = 0;
        while (!iso_of.count(nf)) {
            if (++g2 > 1024) throw std::runtime_error("face order walk lost");
            TF f2 = nf;
            next(f2, -on, nf, on);
        }
        size_t ic, in2;
        int oc, on2;
        half_iso(cur, ocur, ic, oc);
        half_iso(nf, on, in2, on2);
        out.push_back({{ic, oc}, {in2, on2}});
        iso_next = in2;
        cur = nf;
        ocur = -on;
    }
}

template <typename VT>
bool face_order(const GraphIn& in, const std::vector<VT>& verts, const Edge& e,
                const std::unordered_map<size_t, std::vector<size_t>>& inc_tets, std::vector<HalfPair>& out) {
    const auto& faces = *in.faces;
    std::set<size_t> tets_of;
    for (const auto& fe : e.face_edge_indices)
        for (const auto& tf : faces[fe.first].tet_face_indices) tets_of.insert(tf.first);
    if (tets_of.size() == 1) {
        pair_in_one_tet(in.cuts->get(*tets_of.begin()), faces, e, out);
        return true;
    }
    if (in.mi) {
        // src/pair_faces.cpp:131-134
        throw std::runtime_error("More than one tets containing the iso-faces incident to the iso-edge.");
    }
    const auto& a = verts[e.v1];
    const auto& b = verts[e.v2];
    bool on_edge = false;
    size_t x = 0, y = 0, z = 0;
    if (a.simplex_size == 1 && b.simplex_size == 1) {
        on_edge = true;
        x = a.simplex_vert_indices[0];
        y = b.simplex_vert_indices[0];
    } else if (a.simplex_size == 2 && b.simplex_size == 1) {
        x = a.simplex_vert_indices[0];
        y = a.simplex_vert_indices[1];
        z = b.simplex_vert_indices[0];
        on_edge = (z == x || z == y);
    } else if (a.simplex_size == 1 && b.simplex_size == 2) {
        x = b.simplex_vert_indices[0];
        y = b.simplex_vert_indices[1];
        z = a.simplex_vert_indices[0];
        on_edge = (z == x || z == y);
    } else if (a.simplex_size == 2 && b.simplex_size == 2) {
        x = a.simplex_vert_indices[0];
        y = a.simplex_vert_indices[1];
        on_edge = (x == b.simplex_vert_indices[0] && y == b.simplex_vert_indices[1]);
    }
    if (on_edge) {
        const auto& tx = inc_tets.at(x);
        std::unordered_set<size_t> s1(tx.begin(), tx.end());
        std::vector<size_t> common;
        for (size_t t : inc_tets.at(y))
            if (s1.count(t)) common.push_back(t);
        in.cuts->prefetch(common);
        pair_in_tets(in, e, {x, y}, common, out);
    } else {
        size_t t1 = *tets_of.begin(), t2 = *tets_of.rbegin();
        const auto& T1 = (*in.tets)[t1];
        std::vector<size_t> common;
        for (size_t v : (*in.tets)[t2])
            if (std::find(T1.begin(), T1.end(), v) != T1.end()) common.push_back(v);
        pair_in_tets(in, e, common, {t1, t2}, out);
    }
    return true;
}

// ------------------------------------------------------------------ ray shooting

// vertices of the tet edge local v -> local u in order along the edge
// (src/topo_ray_shooting.cpp:529-887)
void edge_order(const CutView& C, size_t v, size_t u, std::vector<size_t>& out) {
    const size_t nlab = C.mi ? 4 : 3;
    bool flag[4] = {true, true, true, true};
    flag[v] = false;
    flag[u] = false;
    size_t vid = Mesh_None, uid = Mesh_None, inner = 0;
    std::vector<char> on(C.vertices.size(), 0);
    for (size_t i = 0; i < C.vertices.size(); ++i) {
        size_t cnt = 0, other = Mesh_None;
        for (size_t q = 0; q < nlab; ++q) {
            size_t p = C.vertices[i][q];
            if (p < 4) {
                if (flag[p])
                    ++cnt;
                else
                    other = p;
            }
        }
        if (cnt == 2) {
            on[i] = 1;
            if (other == u)
                vid = i;
            else if (other == v)
                uid = i;
            else
                ++inner;
        }
    }
    if (inner == 0) {
        out.push_back(vid);
        out.push_back(uid);
        return;
    }
    size_t pid = 0;
    for (size_t i = 0; i < 4; ++i)
        if (flag[i]) {
            pid = i;
            break;
        }
    std::vector<size_t> tri;
    for (size_t f = 0; f < C.faces.size(); ++f) {
        const auto& F = C.faces[f];
        bool on_plane = C.mi ? (F.positive_material_label == pid)
                             : (F.negative_cell == Mesh_None && F.supporting_plane == pid);
        if (on_plane) tri.push_back(f);
    }
    std::map<std::pair<size_t, size_t>, std::pair<size_t, size_t>> fof;
    for (size_t f : tri) {
        const auto& vs = C.faces[f].vertices;
        const size_t n = vs.size();
        for (size_t i = 0; i < n; ++i) {
            size_t a = vs[i], b = vs[(i + 1) % n];
            for (auto key : {std::make_pair(a, b), std::make_pair(b, a)}) {
                auto it = fof.try_emplace(key, std::make_pair(f, Mesh_None));
                if (!it.second) it.first->second.second = f;
            }
        }
    }
    size_t fstart = Mesh_None;
    for (size_t f : tri) {
        bool has_v = false;
        size_t cnt = 0;
        for (size_t x : C.faces[f].vertices) {
            has_v |= (x == vid);
            cnt += on[x];
        }
        if (has_v && cnt == 2) {
            fstart = f;
            break;
        }
    }
    out.push_back(vid);
    std::vector<char> vf(C.faces.size(), 0);
    size_t vcur = vid, fcur = fstart;
    std::pair<size_t, size_t> enext{Mesh_None, Mesh_None};
    size_t guard = 0;
    while (vcur != uid) {
        if (fcur == Mesh_None || ++guard > 4 * C.faces.size() + 8) throw std::runtime_error("edge walk failed");
        std::pair<size_t, size_t> eon{Mesh_None, Mesh_None};
        const auto& vs = C.faces[fcur].vertices;
        const size_t n = vs.size();
        for (size_t i = 0; i < n; ++i) {
            size_t a = vs[i], b = vs[(i + 1) % n];
            if (on[a] && !on[b]) {
                auto& two = fof[{a, b}];
                size_t of = (two.first == fcur) ? two.second : two.first;
                if (a == vid || (of != Mesh_None && vf[of]))
                    ;  // the edge we came through
                else
                    enext = {a, b};
            } else if (on[b] && !on[a]) {
                auto& two = fof[{a, b}];
                size_t of = (two.first == fcur) ? two.second : two.first;
                if (b == vid || (of != Mesh_None && vf[of]))
                    ;
                else
                    enext = {a, b};
            } else if (on[a] && on[b]) {
                eon = {a, b};
            }
        }
        if (eon.first != Mesh_None) {
            vcur = (eon.first == vcur) ? eon.second : eon.first;
            out.push_back(vcur);
        }
        vf[fcur] = 1;
        auto& two = fof[enext];
        fcur = (two.first == fcur) ? two.second : two.first;
    }
}

// the cell on the positive side (IA) / of the negative label (MI) of the face
// holding edge (a, b), then its face through a but not b (and through b but not a)
void passing_faces(const CutView& C, size_t a, size_t b, std::pair<size_t, int>* fa, std::pair<size_t, int>* fb) {
    size_t inc = Mesh_None;
    for (size_t f = 0; f < C.faces.size() && inc == Mesh_None; ++f) {
        const auto& vs = C.faces[f].vertices;
        for (size_t j = 0; j < vs.size(); ++j) {
            size_t x = vs[j], y = vs[(j + 1) % vs.size()];
            if ((x == a && y == b) || (x == b && y == a)) {
                inc = f;
                break;
            }
        }
    }
    if (inc == Mesh_None) throw std::runtime_error("no face through the ray edge");
    size_t cell;
    std::vector<size_t> cell_of_label;
    if (C.mi) {
        size_t mx = 0;
        for (const auto& c : C.cells) mx = std::max(mx, c.material_label);
        cell_of_label.assign(mx + 1, Mesh_None);
        for (size_t i = 0; i < C.cells.size(); ++i) cell_of_label[C.cells[i].material_label] = i;
        cell = cell_of_label.at(C.faces[inc].negative_material_label);
    } else {
        cell = C.faces[inc].positive_cell;
    }
    auto side = [&](size_t f) -> int {
        if (C.mi) {
            size_t l = C.faces[f].positive_material_label;
            return (l < cell_of_label.size() && cell_of_label[l] == cell) ? 1 : -1;
        }
        return C.faces[f].positive_cell == cell ? 1 : -1;
    };
    for (size_t f : C.cells.at(cell).faces) {
        bool ha = false, hb = false;
        for (size_t x : C.faces[f].vertices) {
            if (x == a)
                ha = true;
            else if (x == b)
                hb = true;
        }
        if (ha && !hb && fa) {
            *fa = {f, side(f)};
            if (!fb) return;
        } else if (!ha && hb && fb) {
            *fb = {f, side(f)};
        }
    }
}

template <typename VT>
void ray_shooting(const GraphIn& in, const std::vector<VT>& verts, const std::vector<std::vector<size_t>>& patches,
                  const std::vector<size_t>& patch_of_face, const std::vector<std::vector<size_t>>& shells,
                  const std::vector<size_t>& shell_of_hp, const std::vector<std::vector<size_t>>& comps,
                  const std::vector<size_t>& comp_of_patch, std::vector<std::vector<size_t>>& cells) {
    const auto& pts = *in.pts;
    const auto& tets = *in.tets;
    const auto& faces = *in.faces;
    // src/topo_ray_shooting.cpp:332-353 — each vertex points to the smallest
    // (x,y,z) corner of the last tet containing it
    std::vector<size_t> nxt(pts.size(), Mesh_None);
    for (const auto& t : tets) {
        size_t mi = 0;
        for (size_t i = 1; i < 4; ++i)
            if (xyz_less(pts[t[i]], pts[t[mi]])) mi = i;
        for (size_t i = 0; i < 4; ++i)
            if (i != mi) nxt[t[i]] = t[mi];
    }
    // src/topo_ray_shooting.cpp:355-438 — extremal tree edge per component
    std::vector<size_t> ext(2 * comps.size(), Mesh_None);
    std::vector<size_t> vert_on(pts.size(), Mesh_None);
    std::unordered_map<TF, size_t, PairHash> iso_of_tf;
    std::unordered_map<TF, std::pair<size_t, size_t>, PairHash> comp_of_tv;
    std::vector<char> vseen(verts.size(), 0);
    for (size_t i = 0; i < patches.size(); ++i) {
        size_t cid = comp_of_patch[i];
        size_t& u1 = ext[2 * cid];
        size_t& u2 = ext[2 * cid + 1];
        for (size_t f : patches[i]) {
            for (const auto& tf : faces[f].tet_face_indices) iso_of_tf.try_emplace(tf, f);
            for (size_t vi : faces[f].vert_indices) {
                if (vseen[vi]) continue;
                vseen[vi] = 1;
                const auto& V = verts[vi];
                if (V.simplex_size != 2) continue;
                size_t a = V.simplex_vert_indices[0], b = V.simplex_vert_indices[1];
                if (nxt[b] == a) std::swap(a, b);  // orient along the tree edge
                if (nxt[a] != b) continue;
                if (u1 == Mesh_None) {
                    u1 = a;
                    u2 = b;
                } else if (b == u2) {
                    if (xyz_less(pts[a], pts[u1])) u1 = a;
                } else if (xyz_less(pts[b], pts[u2])) {
                    u1 = a;
                    u2 = b;
                }
                vert_on[a] = vi;
                comp_of_tv.try_emplace(TF{V.tet_index, V.tet_vert_index}, std::make_pair(vi, cid));
            }
        }
    }
    auto shell_of = [&](size_t tet, std::pair<size_t, int> fo) {
        size_t f = iso_of_tf.at({tet, fo.first});
        return fo.second == 1 ? shell_of_hp[2 * patch_of_face[f]] : shell_of_hp[2 * patch_of_face[f] + 1];
    };
    auto local_of = [&](size_t tet, size_t g1, size_t g2, size_t& l1, size_t& l2) {
        for (size_t j = 0; j < 4; ++j) {
            if (tets[tet][j] == g1)
                l1 = j;
            else if (tets[tet][j] == g2)
                l2 = j;
        }
    };
    std::vector<size_t> need;
    for (size_t i = 0; i < comps.size(); ++i) {
        if (ext[2 * i] == Mesh_None) throw std::runtime_error("component without a tree-edge vertex");
        need.push_back(verts[vert_on[ext[2 * i]]].tet_index);
    }
    in.cuts->prefetch(need);
    std::vector<std::pair<size_t, size_t>> links;
    for (size_t i = 0; i < comps.size(); ++i) {
        size_t e1 = ext[2 * i], e2 = ext[2 * i + 1];
        size_t tid = verts[vert_on[e1]].tet_index;
        const CutView& C = in.cuts->get(tid);
        size_t l1 = 0, l2 = 0;
        local_of(tid, e1, e2, l1, l2);
        std::vector<size_t> ord;
        edge_order(C, l1, l2, ord);
        size_t js = Mesh_None;
        for (size_t j = 0; j + 1 < ord.size(); ++j) {
            auto it = comp_of_tv.find({tid, ord[j]});
            if (it != comp_of_tv.end() && it->second.second == i) js = j;
        }
        if (js == Mesh_None) throw std::runtime_error("ray start vertex not found");
        if (js + 2 < ord.size()) {
            std::pair<size_t, int> fa, fb;
            passing_faces(C, ord[js], ord[js + 1], &fa, &fb);
            links.emplace_back(shell_of(tid, fa), shell_of(tid, fb));
        } else {
            std::pair<size_t, int> fa;
            passing_faces(C, ord[js], ord.back(), &fa, nullptr);
            size_t s_start = shell_of(tid, fa);
            size_t vc = e2;
            while (nxt[vc] != Mesh_None && vert_on[vc] == Mesh_None) vc = nxt[vc];
            if (vert_on[vc] != Mesh_None) {
                size_t et = verts[vert_on[vc]].tet_index;
                const CutView& E = in.cuts->get(et);
                size_t a1 = 0, a2 = 0;
                local_of(et, vc, nxt[vc], a1, a2);
                std::vector<size_t> o2;
                edge_order(E, a1, a2, o2);
                std::pair<size_t, int> fe;
                passing_faces(E, o2[1], o2.front(), &fe, nullptr);
                links.emplace_back(s_start, shell_of(et, fe));
            } else {
                links.emplace_back(s_start, Mesh_None);
            }
        }
    }
    compute_arrangement_cells(shells.size(), links, cells);
}

}  // namespace

// ------------------------------------------------------------------ driver

bool run_graph(const GraphIn& in, GraphOut& out) {
    const auto& faces = *in.faces;
    const size_t nverts = in.mi ? in.mi_verts->size() : in.iso_verts->size();
    double t = now();
    auto lap = [&](const char* label) {
        double t2 = now();
        out.timings.emplace_back(label, t2 - t);
        t = t2;
    };
    std::vector<std::vector<size_t>> eof;
    if (in.edges_from)
        in.edges_from(eof, out.edges);
    else
        compute_mesh_edges(faces, eof, out.edges);
    lap(in.mi ? "edge-face connectivity" : "isoEdge-face connectivity");
    out.stats.emplace_back(in.mi ? "num_MI_edges" : "num_iso_edges", out.edges.size());

    std::vector<std::pair<size_t, size_t>> plabel;
    if (in.patches_from) {
        in.patches_from(out.patches);
        for (const auto& p : out.patches) plabel.push_back(faces[p[0]].func_index);
    } else {
        compute_patches(eof, out.edges, faces, out.patches, plabel);
    }
    if (in.mi)
        out.patch_label_mi = plabel;
    else
        for (auto& p : plabel) out.patch_label.push_back(p.first);
    lap("patches");
    out.stats.emplace_back("num_patches", out.patches.size());

    std::vector<size_t> patch_of_face(faces.size());
    for (size_t i = 0; i < out.patches.size(); ++i)
        for (size_t f : out.patches[i]) patch_of_face[f] = i;
    lap("face-patch map");

    out.nm_edges_of_vert.assign(nverts, {});
    for (size_t i = 0; i < out.edges.size(); ++i)
        if (out.edges[i].face_edge_indices.size() > 2) {
            out.nm_edges_of_vert[out.edges[i].v1].push_back(i);
            out.nm_edges_of_vert[out.edges[i].v2].push_back(i);
        }
    if (in.chains_from)
        in.chains_from(out.chains);
    else
        compute_chains(out.edges, out.nm_edges_of_vert, out.chains);
    lap("chains");
    out.stats.emplace_back("num_chains", out.chains.size());

    // tets around vertices that are zeros of some function / material ties
    std::unordered_map<size_t, std::vector<size_t>> inc_tets;
    if (in.found_degenerate && in.degenerate_vertex) {
        const auto& tets = *in.tets;
        for (size_t i = 0; i < tets.size(); ++i)
            for (size_t j = 0; j < 4; ++j)
                if ((*in.degenerate_vertex)[tets[i][j]]) inc_tets[tets[i][j]].push_back(i);
    }
    lap("vert-tet connectivity");

    std::vector<std::vector<HalfPair>> hfp(out.chains.size()), hpp(out.chains.size());
    {
        std::vector<size_t> need;
        for (const auto& ch : out.chains)
            for (const auto& fe : out.edges[ch[0]].face_edge_indices)
                for (const auto& tf : faces[fe.first].tet_face_indices) need.push_back(tf.first);
        in.cuts->prefetch(need);
    }
    try {
        for (size_t i = 0; i < out.chains.size(); ++i) {
            const Edge& e = out.edges[out.chains[i][0]];
            if (in.mi)
                face_order(in, *in.mi_verts, e, inc_tets, hfp[i]);
            else
                face_order(in, *in.iso_verts, e, inc_tets, hfp[i]);
        }
    } catch (std::exception& ex) {
        out.error = std::string("order patches failed: ") + ex.what();
        return false;
    }
    for (size_t i = 0; i < hfp.size(); ++i)
        for (const auto& pr : hfp[i])
            hpp[i].push_back({{patch_of_face[pr.first.first], pr.first.second},
                              {patch_of_face[pr.second.first], pr.second.second}});
    lap("order patches around chains");

    std::vector<size_t> shell_of_hp, comp_of_patch;
    std::vector<std::vector<size_t>> comps;
    compute_shells_and_components(out.patches.size(), hpp, out.shells, shell_of_hp, comps, comp_of_patch);
    out.num_components = comps.size();
    lap("shells and components");
    out.stats.emplace_back("num_shells", out.shells.size());
    out.stats.emplace_back("num_components", comps.size());

    if (comps.size() < 2) {
        for (size_t i = 0; i < out.shells.size(); ++i) out.cells.push_back({i});
        if (in.mi && out.cells.empty()) out.cells.push_back({Mesh_None});
    } else {
        if (in.mi)
            ray_shooting(in, *in.mi_verts, out.patches, patch_of_face, out.shells, shell_of_hp, comps, comp_of_patch,
                         out.cells);
        else
            ray_shooting(in, *in.iso_verts, out.patches, patch_of_face, out.shells, shell_of_hp, comps,
                         comp_of_patch, out.cells);
    }
    lap(in.mi ? "material cells" : "arrangement cells");
    out.stats.emplace_back("num_cells", out.cells.size());

    if (in.mi)
        out.cell_label_mi = sign_propagation_MI(out.cells, shell_of_hp, out.shells, out.patch_label_mi, in.n_func,
                                                in.sample_value);
    else
        out.cell_label = sign_propagation(out.cells, shell_of_hp, out.shells, out.patch_label, in.n_func,
                                          in.sample_sign);
    return true;
}

}  // namespace rin
