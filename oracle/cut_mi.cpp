// ORACLE — test infrastructure only.
//
// Per-tet material-interface cut: the contract of the external
// simplicial_arrangement::compute_material_interface called at
// src/material_interface.cpp:323-327 (fields used in
// src/extract_mesh.cpp:628-986, src/pair_faces.cpp:240-352,
// src/topo_ray_shooting.cpp:709-887,945-1129).
//
// Labels 0..3 are the tet faces (the "boundary materials"), materials get
// labels 4.. in insertion order.  The complex starts as the whole tet filled
// by material 4.  Inserting material m:
//   * every vertex v gets s(v) = sign(f_m(v) - max_c f_c(v)), exactly;
//   * if m never wins strictly it changes nothing; if it ties a whole cell it
//     is a duplicate of that cell's material (unique_materials);
//   * otherwise one vertex is created per edge with a strict sign change,
//     faces are clipped to their s<=0 part (the remainder keeps the face id),
//     tet-boundary face parts where m wins are merged into one face per tet
//     face, every cell crossed by f_m = f_X gets a new face (pos m, neg X),
//     fully dominated cells disappear, unreferenced vertices disappear and
//     tet corners where m wins are relabelled.
// Face labels: boundary faces have pos = tet face id and neg = cell material;
// interior faces have pos = the later-inserted material.  Cycles are oriented
// with the normal toward the pos side.
#include <algorithm>
#include <map>
#include <stdexcept>

#include "exact.h"
#include "oracle.h"

namespace oracle {

namespace {

void label_row(const std::vector<std::array<double, 4>>& M, int id, double* row) {
    for (int c = 0; c < 5; ++c) row[c] = 0.0;
    if (id < 4) {
        row[id] = 1.0;
    } else {
        for (int c = 0; c < 4; ++c) row[c] = M[id - 4][c];
        row[4] = -1.0;
    }
}

// sign of f_m - (common value of the vertex's materials) at the vertex:
// unknowns (l0..l3, t), rows of the vertex's labels, then the query row
// (f_m, -1) or the normalisation row (1,1,1,1,0).
int orient_mi(const std::vector<std::array<double, 4>>& Mv, const std::array<int, 4>& t, int m) {
    double A[25], B[25];
    for (int r = 0; r < 4; ++r) {
        label_row(Mv, t[r], A + 5 * r);
        label_row(Mv, t[r], B + 5 * r);
    }
    label_row(Mv, m, A + 20);
    for (int c = 0; c < 4; ++c) B[20 + c] = 1.0;
    B[24] = 0.0;
    int a = det_sign(5, A);
    int b = det_sign(5, B);
    if (b == 0) throw std::runtime_error("oracle mi_cut: degenerate vertex");
    return a * b;
}

std::array<int, 4> sorted4(int a, int b, int c, int d) {
    std::array<int, 4> r{a, b, c, d};
    std::sort(r.begin(), r.end());
    return r;
}

struct Edge3 {
    int a, b, l;
};

// chain directed edges into one cycle starting at edges[0]
bool chain(const std::vector<Edge3>& E, std::vector<int>& vs, std::vector<int>& es) {
    vs.clear();
    es.clear();
    if (E.empty()) return false;
    size_t cur = 0;
    for (size_t it = 0; it < E.size(); ++it) {
        vs.push_back(E[cur].a);
        es.push_back(E[cur].l);
        int to = E[cur].b;
        size_t nxt = E.size();
        for (size_t q = 0; q < E.size(); ++q)
            if (E[q].a == to) {
                nxt = q;
                break;
            }
        if (nxt == E.size()) return false;
        cur = nxt;
        if (cur == 0) return it + 1 == E.size();
    }
    return false;
}

}  // namespace

MICut mi_cut(const std::vector<std::array<double, 4>>& mats) {
    MICut C;
    const int k = (int)mats.size();
    C.nlabels = 4 + k;
    C.grp.assign(C.nlabels, NONE);
    for (int i = 0; i < 4; ++i) C.grp[i] = i;
    if (k == 0) return C;
    for (int i = 0; i < 4; ++i) {
        std::array<int, 4> t{};
        int n = 0;
        for (int j = 0; j < 4; ++j)
            if (j != i) t[n++] = j;
        t[3] = 4;
        C.verts.push_back(sorted4(t[0], t[1], t[2], t[3]));
    }
    // outward-oriented tet faces: normal toward the boundary label (outside)
    const int cyc[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    for (int i = 0; i < 4; ++i) {
        MIFace f;
        f.pos = i;
        f.neg = 4;
        for (int j = 0; j < 3; ++j) {
            f.vs.push_back(cyc[i][j]);
            f.es.push_back(cyc[i][(j + 2) % 3]);
        }
        C.faces.push_back(f);
    }
    C.cells = {4};
    C.grp[4] = 4;

    for (int j = 1; j < k; ++j) {
        const int m = 4 + j;
        const int nv0 = (int)C.verts.size();
        std::vector<int> s(nv0);
        bool any_pos = false;
        for (int v = 0; v < nv0; ++v) {
            s[v] = orient_mi(mats, C.verts[v], m);
            any_pos |= s[v] > 0;
        }
        // per-cell sign summary
        std::map<int, int> cpos, cneg;
        for (int X : C.cells) cpos[X] = cneg[X] = 0;
        for (const MIFace& f : C.faces) {
            for (int lab : {f.pos, f.neg}) {
                if (lab < 4) continue;
                for (int v : f.vs) {
                    cpos[lab] |= s[v] > 0;
                    cneg[lab] |= s[v] < 0;
                }
            }
        }
        if (!any_pos) {
            C.grp[m] = m;
            for (int X : C.cells) {
                if (!cpos[X] && !cneg[X]) {
                    C.grp[m] = C.grp[X];
                    C.dup = true;
                    break;
                }
            }
            continue;
        }
        C.grp[m] = m;
        std::map<int, bool> consumed, split;
        for (int X : C.cells) {
            consumed[X] = cpos[X] && !cneg[X];
            split[X] = cpos[X] && cneg[X];
        }

        std::map<std::pair<int, int>, int> cross;
        for (const MIFace& f : C.faces) {
            const int n = (int)f.vs.size();
            for (int i = 0; i < n; ++i) {
                int u = f.vs[i], w = f.vs[(i + 1) % n];
                if (s[u] * s[w] < 0) {
                    auto key = std::make_pair(std::min(u, w), std::max(u, w));
                    if (!cross.count(key)) {
                        cross[key] = (int)C.verts.size();
                        C.verts.push_back(sorted4(f.pos, f.neg, f.es[i], m));
                        s.push_back(0);
                    }
                }
            }
        }

        std::vector<MIFace> nf;
        std::vector<Edge3> plus[4];
        for (const MIFace& f : C.faces) {
            bool fp = false, fn = false;
            for (int v : f.vs) {
                fp |= s[v] > 0;
                fn |= s[v] < 0;
            }
            const int n = (int)f.vs.size();
            if (fn && !fp) {
                nf.push_back(f);
            } else if (fn && fp) {
                std::vector<int> av, ae;
                for (int i = 0; i < n; ++i) {
                    int u = f.vs[i], w = f.vs[(i + 1) % n];
                    av.push_back(u);
                    ae.push_back(f.es[i]);
                    if (s[u] * s[w] < 0) {
                        av.push_back(cross[{std::min(u, w), std::max(u, w)}]);
                        ae.push_back(f.es[i]);
                    }
                }
                const int L = (int)av.size();
                auto part = [&](int sgn, int cutlab, std::vector<int>& ov, std::vector<int>& oe) {
                    std::vector<int> idx;
                    for (int i = 0; i < L; ++i)
                        if (s[av[i]] * sgn >= 0) idx.push_back(i);
                    for (size_t q = 0; q < idx.size(); ++q) {
                        int ia = idx[q], ib = idx[(q + 1) % idx.size()];
                        ov.push_back(av[ia]);
                        oe.push_back(((ia + 1) % L == ib) ? ae[ia] : cutlab);
                    }
                };
                MIFace g;
                g.pos = f.pos;
                g.neg = f.neg;
                part(-1, m, g.vs, g.es);
                nf.push_back(g);
                if (f.pos < 4) {
                    std::vector<int> pv, pe;
                    part(+1, f.neg, pv, pe);
                    for (size_t q = 0; q < pv.size(); ++q)
                        plus[f.pos].push_back({pv[q], pv[(q + 1) % pv.size()], pe[q]});
                }
            } else if (fp && !fn) {
                if (f.pos < 4)
                    for (int i = 0; i < n; ++i) plus[f.pos].push_back({f.vs[i], f.vs[(i + 1) % n], f.es[i]});
            } else {
                // face on the zero set of f_m - max: relabel the dominated sides
                bool pc = f.pos >= 4 && consumed[f.pos];
                bool nc = consumed[f.neg];
                if (f.pos < 4) {
                    if (nc)
                        for (int i = 0; i < n; ++i) plus[f.pos].push_back({f.vs[i], f.vs[(i + 1) % n], f.es[i]});
                    else
                        nf.push_back(f);
                } else if (pc && nc) {
                    // interior to the new material's cell
                } else if (pc) {
                    MIFace g = f;
                    g.pos = m;
                    nf.push_back(g);
                } else if (nc) {
                    // the new material becomes the pos side: reverse the cycle
                    MIFace g;
                    g.pos = m;
                    g.neg = f.pos;
                    for (int i = 0; i < n; ++i) {
                        g.vs.push_back(f.vs[(n - i) % n]);
                        g.es.push_back(f.es[(2 * n - 1 - i) % n]);
                    }
                    nf.push_back(g);
                } else {
                    nf.push_back(f);
                }
            }
        }
        // new faces between m and each crossed cell (collected before the
        // merged boundary faces are appended)
        std::vector<MIFace> cutf;
        for (int X : C.cells) {
            if (!split[X]) continue;
            std::vector<Edge3> E;
            for (const MIFace& g : nf) {
                if (g.pos != X && g.neg != X) continue;
                const int n = (int)g.vs.size();
                int other = (g.pos == X) ? g.neg : g.pos;
                for (int i = 0; i < n; ++i) {
                    int a = g.vs[i], b = g.vs[(i + 1) % n];
                    if (s[a] == 0 && s[b] == 0) {
                        if (g.neg == X)
                            E.push_back({b, a, other});
                        else
                            E.push_back({a, b, other});
                    }
                }
            }
            MIFace F;
            F.pos = m;
            F.neg = X;
            if (!chain(E, F.vs, F.es)) throw std::runtime_error("oracle mi_cut: bad cut polygon");
            cutf.push_back(F);
        }
        // merge the parts of each tet face now owned by m
        for (int i = 0; i < 4; ++i) {
            if (plus[i].empty()) continue;
            std::vector<Edge3> E;
            for (size_t q = 0; q < plus[i].size(); ++q) {
                bool shared = false;
                for (size_t r = 0; r < plus[i].size(); ++r)
                    if (plus[i][r].a == plus[i][q].b && plus[i][r].b == plus[i][q].a) {
                        shared = true;
                        break;
                    }
                if (!shared) E.push_back(plus[i][q]);
            }
            MIFace F;
            F.pos = i;
            F.neg = m;
            std::vector<int> vs, es;
            if (!chain(E, vs, es)) throw std::runtime_error("oracle mi_cut: bad merged face");
            // drop vertices between two collinear edges (same third label)
            const int n = (int)vs.size();
            for (int q = 0; q < n; ++q) {
                int prev = (q + n - 1) % n;
                if (vs[q] >= 4 && es[prev] == es[q]) continue;
                F.vs.push_back(vs[q]);
                F.es.push_back(es[q]);
            }
            nf.push_back(F);
        }
        for (auto& F : cutf) nf.push_back(F);

        // cells
        std::vector<int> cells;
        for (int X : C.cells)
            if (!consumed[X]) cells.push_back(X);
        cells.push_back(m);
        C.cells = cells;

        // relabel corners won by m, drop unreferenced vertices
        for (int c = 0; c < 4; ++c)
            if (s[c] > 0) {
                std::array<int, 4> t{};
                int n = 0;
                for (int q = 0; q < 4; ++q)
                    if (q != c) t[n++] = q;
                t[3] = m;
                C.verts[c] = sorted4(t[0], t[1], t[2], t[3]);
            }
        std::vector<char> used(C.verts.size(), 0);
        for (int c = 0; c < 4; ++c) used[c] = 1;
        for (auto& f : nf)
            for (int v : f.vs) used[v] = 1;
        std::vector<int> remap(C.verts.size(), NONE);
        std::vector<std::array<int, 4>> nv;
        for (size_t v = 0; v < C.verts.size(); ++v)
            if (used[v]) {
                remap[v] = (int)nv.size();
                nv.push_back(C.verts[v]);
            }
        for (auto& f : nf)
            for (int& v : f.vs) v = remap[v];
        C.verts = nv;
        C.faces = nf;
    }
    return C;
}

}  // namespace oracle
