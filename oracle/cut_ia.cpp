// ORACLE — test infrastructure only.
//
// Per-tet implicit-arrangement cut: the contract of the external
// simplicial_arrangement::compute_arrangement called at
// src/implicit_arrangement.cpp:283 (field use in src/extract_mesh.cpp:54-260,
// src/pair_faces.cpp:138-238,474-651, src/topo_ray_shooting.cpp:529-707).
//
// Restated as incremental plane insertion into the tet's convex cell complex:
//   * planes 0..3 are the tet faces (plane i = barycentric coordinate i, the
//     face opposite corner i, positive inside); implicit planes get ids 4.. in
//     insertion order;
//   * a vertex is a sorted triple of plane ids; corner i is the triple of the
//     three faces through it;
//   * inserting plane p classifies every vertex exactly, creates one vertex per
//     edge with a strict sign change, splits faces (the positive part keeps the
//     face id, the negative part is appended), splits cells (the positive part
//     keeps the cell id, the negative part is appended) and appends one face
//     on p per split cell;
//   * a plane containing an existing face is a duplicate: it joins that face's
//     plane group (unique_planes) and changes nothing else.
#include <map>
#include <stdexcept>

#include "exact.h"
#include "oracle.h"

namespace oracle {

namespace {

// row of plane `id` as corner values
void plane_row(const std::vector<std::array<double, 4>>& P, int id, double* row) {
    for (int c = 0; c < 4; ++c) row[c] = 0.0;
    if (id < 4)
        row[id] = 1.0;
    else
        for (int c = 0; c < 4; ++c) row[c] = P[id - 4][c];
}

// sign of plane p at the point where planes t[0],t[1],t[2] meet:
//   sign(det[t0;t1;t2;p]) * sign(det[t0;t1;t2;1])
int orient(const std::vector<std::array<double, 4>>& P, const std::array<int, 3>& t, int p) {
    double M[16], N[16];
    for (int r = 0; r < 3; ++r) {
        plane_row(P, t[r], M + 4 * r);
        plane_row(P, t[r], N + 4 * r);
    }
    plane_row(P, p, M + 12);
    for (int c = 0; c < 4; ++c) N[12 + c] = 1.0;
    int a = det_sign(4, M);
    int b = det_sign(4, N);
    if (b == 0) throw std::runtime_error("oracle ia_cut: degenerate vertex");
    return a * b;
}

std::array<int, 3> sorted3(int a, int b, int c) {
    if (a > b) std::swap(a, b);
    if (b > c) std::swap(b, c);
    if (a > b) std::swap(a, b);
    return {a, b, c};
}

}  // namespace

IACut ia_cut(const std::vector<std::array<double, 4>>& planes) {
    IACut C;
    const int k = (int)planes.size();
    C.nplanes = 4 + k;
    C.verts = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    // inward-oriented tet faces (positive volume tet): normal toward the inside
    const int cyc[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
    for (int i = 0; i < 4; ++i) {
        IAFace f;
        f.plane = i;
        f.pos = 0;
        f.neg = NONE;
        for (int j = 0; j < 3; ++j) {
            f.vs.push_back(cyc[i][j]);
            f.es.push_back(cyc[i][(j + 2) % 3]);  // edge a->b lies on the plane of the third corner
        }
        C.faces.push_back(f);
    }
    C.ncells = 1;
    C.grp.assign(C.nplanes, NONE);
    C.ori.assign(C.nplanes, 1);
    for (int i = 0; i < 4; ++i) C.grp[i] = i;

    for (int j = 0; j < k; ++j) {
        const int P = 4 + j;
        const int nv0 = (int)C.verts.size();
        std::vector<int> s(nv0);
        bool any_pos = false, any_neg = false;
        for (int v = 0; v < nv0; ++v) {
            s[v] = orient(planes, C.verts[v], P);
            any_pos |= s[v] > 0;
            any_neg |= s[v] < 0;
        }
        // duplicate plane: some face lies entirely on p
        int coplanar = NONE;
        for (int f = 0; f < (int)C.faces.size() && coplanar == NONE; ++f) {
            bool allz = true;
            for (int v : C.faces[f].vs) allz &= (s[v] == 0);
            if (allz) coplanar = f;
        }
        if (coplanar != NONE) {
            const IAFace& F = C.faces[coplanar];
            int sv = 0;
            for (int g = 0; g < (int)C.faces.size() && sv == 0; ++g) {
                if (C.faces[g].pos != F.pos && C.faces[g].neg != F.pos) continue;
                for (int v : C.faces[g].vs)
                    if (s[v] != 0) {
                        sv = s[v];
                        break;
                    }
            }
            C.grp[P] = C.grp[F.plane];
            C.ori[P] = (sv > 0) ? C.ori[F.plane] : 1 - C.ori[F.plane];
            C.dup = true;
            continue;
        }
        C.grp[P] = P;
        C.ori[P] = 1;
        if (!any_pos || !any_neg) continue;

        // one new vertex per edge with a strict sign change
        std::map<std::pair<int, int>, int> cross;
        for (const IAFace& f : C.faces) {
            const int n = (int)f.vs.size();
            for (int i = 0; i < n; ++i) {
                int u = f.vs[i], w = f.vs[(i + 1) % n];
                if (s[u] * s[w] < 0) {
                    auto key = std::make_pair(std::min(u, w), std::max(u, w));
                    if (!cross.count(key)) {
                        cross[key] = (int)C.verts.size();
                        C.verts.push_back(sorted3(f.plane, f.es[i], P));
                        s.push_back(0);
                    }
                }
            }
        }

        // split faces
        const int nf0 = (int)C.faces.size();
        for (int fi = 0; fi < nf0; ++fi) {
            IAFace& f = C.faces[fi];
            bool fp = false, fn = false;
            for (int v : f.vs) {
                fp |= s[v] > 0;
                fn |= s[v] < 0;
            }
            if (!(fp && fn)) continue;
            // augmented cycle: original vertices with crossing points inserted
            std::vector<int> av, ae;
            const int n = (int)f.vs.size();
            for (int i = 0; i < n; ++i) {
                int u = f.vs[i], w = f.vs[(i + 1) % n];
                av.push_back(u);
                ae.push_back(f.es[i]);
                if (s[u] * s[w] < 0) {
                    av.push_back(cross[{std::min(u, w), std::max(u, w)}]);
                    ae.push_back(f.es[i]);
                }
            }
            auto part = [&](int sgn, std::vector<int>& ov, std::vector<int>& oe) {
                const int m = (int)av.size();
                std::vector<int> idx;
                for (int i = 0; i < m; ++i)
                    if (s[av[i]] * sgn >= 0) idx.push_back(i);
                for (size_t q = 0; q < idx.size(); ++q) {
                    int ia = idx[q], ib = idx[(q + 1) % idx.size()];
                    ov.push_back(av[ia]);
                    oe.push_back(((ia + 1) % m == ib) ? ae[ia] : P);
                }
            };
            IAFace neg;
            neg.plane = f.plane;
            neg.pos = f.pos;
            neg.neg = f.neg;
            part(-1, neg.vs, neg.es);
            std::vector<int> pv, pe;
            part(+1, pv, pe);
            f.vs = pv;
            f.es = pe;
            C.faces.push_back(neg);
        }

        // split cells
        const int nc0 = C.ncells;
        for (int c = 0; c < nc0; ++c) {
            bool cp = false, cn = false;
            for (const IAFace& g : C.faces) {
                if (g.pos != c && g.neg != c) continue;
                for (int v : g.vs) {
                    cp |= s[v] > 0;
                    cn |= s[v] < 0;
                }
            }
            if (!(cp && cn)) continue;
            const int c2 = C.ncells++;
            for (IAFace& g : C.faces) {
                if (g.pos != c && g.neg != c) continue;
                bool gn = false;
                for (int v : g.vs) gn |= s[v] < 0;
                if (!gn) continue;
                if (g.pos == c) g.pos = c2;
                if (g.neg == c) g.neg = c2;
            }
            // boundary of the new face: edges on p of the positive part's faces
            std::vector<std::array<int, 3>> edges;  // from, to, plane
            for (const IAFace& g : C.faces) {
                if (g.pos != c && g.neg != c) continue;
                const int n = (int)g.vs.size();
                for (int i = 0; i < n; ++i) {
                    int a = g.vs[i], b = g.vs[(i + 1) % n];
                    if (s[a] == 0 && s[b] == 0) {
                        if (g.pos == c)
                            edges.push_back({b, a, g.plane});
                        else
                            edges.push_back({a, b, g.plane});
                    }
                }
            }
            IAFace nf;
            nf.plane = P;
            nf.pos = c;
            nf.neg = c2;
            if (edges.empty()) throw std::runtime_error("oracle ia_cut: empty cut polygon");
            int cur = 0;
            for (size_t it = 0; it < edges.size(); ++it) {
                nf.vs.push_back(edges[cur][0]);
                nf.es.push_back(edges[cur][2]);
                int to = edges[cur][1];
                int nxt = -1;
                for (size_t q = 0; q < edges.size(); ++q)
                    if (edges[q][0] == to) {
                        nxt = (int)q;
                        break;
                    }
                if (nxt < 0) throw std::runtime_error("oracle ia_cut: open cut polygon");
                cur = nxt;
                if (cur == 0) {
                    if (it + 1 != edges.size()) throw std::runtime_error("oracle ia_cut: split cut polygon");
                    break;
                }
            }
            C.faces.push_back(nf);
        }
    }
    return C;
}

}  // namespace oracle
