// ORACLE — test infrastructure only.
// Material-interface half of the restated hot path.
#include <algorithm>
#include <chrono>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <unordered_map>

#include "oracle.h"

namespace oracle {

namespace {
double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
constexpr uint64_t U64NONE = UINT64_MAX;
}  // namespace

void run_mi(const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<uint64_t, 4>>& tets,
            const std::vector<double>& vals, uint64_t F, int flags, Result& r) {
    r = Result();
    r.mi = true;
    r.nv = pts.size();
    r.nt = tets.size();
    r.nf = F;
    const uint64_t V = r.nv, T = r.nt;
    auto fv = [&](uint64_t v, uint64_t f) { return vals[v * F + f]; };

    // src/material_interface.cpp:53-92 — first maximum wins; ties list every maximum
    double t0 = now();
    r.highest.resize(V);
    r.degenerate_vertex.assign(V, 0);
    std::unordered_map<uint64_t, std::vector<uint64_t>> ties;
    for (uint64_t v = 0; v < V; ++v) {
        double mx = fv(v, 0);
        uint64_t id = 0, cnt = 1;
        for (uint64_t j = 1; j < F; ++j) {
            if (fv(v, j) > mx) {
                mx = fv(v, j);
                id = j;
                cnt = 1;
            } else if (fv(v, j) == mx) {
                ++cnt;
            }
        }
        r.highest[v] = id;
        if (cnt > 1) {
            // vertices with tied maxima (the reference's is_degenerate_vertex,
            // src/material_interface.cpp:79-81)
            ++r.num_degenerate;
            r.degenerate_vertex[v] = 1;
            auto& L = ties[v];
            for (uint64_t j = 0; j < F; ++j)
                if (fv(v, j) == mx) L.push_back(j);
        }
    }
    double t1 = now();
    r.timings.push_back({"highest func", t1 - t0});

    // :94-152 — high materials of the corners, then every material exceeding
    // the corner-wise minimum of the high materials at two or more corners
    r.start_index.push_back(0);
    for (uint64_t t = 0; t < T; ++t) {
        std::set<uint64_t> mats;
        for (int c = 0; c < 4; ++c) {
            uint64_t v = tets[t][c];
            if (r.degenerate_vertex[v])
                mats.insert(ties[v].begin(), ties[v].end());
            else
                mats.insert(r.highest[v]);
        }
        if (mats.size() < 2) {
            r.start_index.push_back(r.func_in_tet.size());
            continue;
        }
        double mn[4];
        for (int c = 0; c < 4; ++c) mn[c] = std::numeric_limits<double>::max();
        for (uint64_t m : mats)
            for (int c = 0; c < 4; ++c)
                if (fv(tets[t][c], m) < mn[c]) mn[c] = fv(tets[t][c], m);
        for (uint64_t j = 0; j < F; ++j) {
            int gc = 0;
            for (int c = 0; c < 4; ++c) gc += fv(tets[t][c], j) > mn[c];
            if (gc > 1) mats.insert(j);
        }
        ++r.num_intersecting;
        r.func_in_tet.insert(r.func_in_tet.end(), mats.begin(), mats.end());
        r.start_index.push_back(r.func_in_tet.size());
    }
    double t2 = now();
    r.timings.push_back({"filter", t2 - t1});

    // :288-351
    r.cut_index.assign(T, -1);
    std::vector<uint64_t> act;
    for (uint64_t t = 0; t < T; ++t)
        if (r.start_index[t + 1] > r.start_index[t]) {
            r.cut_index[t] = (int64_t)act.size();
            act.push_back(t);
        }
    r.mi_cuts.resize(act.size());
    if (flags & F_LOOKUP) mi_table2(0);
    double bt[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bt[:3])
    for (long long a = 0; a < (long long)act.size(); ++a) {
        uint64_t t = act[a];
        double c0 = now();
        uint64_t s0 = r.start_index[t], k = r.start_index[t + 1] - s0;
        if ((flags & F_LOOKUP) && k == 2 && !(flags & F_REVERSE)) {
            const uint64_t m0 = r.func_in_tet[s0], m1 = r.func_in_tet[s0 + 1];
            int pat = 0;
            for (int c = 3; c >= 0; --c) {
                const double a0 = fv(tets[t][c], m0), a1 = fv(tets[t][c], m1);
                pat = pat * 3 + (a1 > a0 ? 2 : (a1 < a0 ? 0 : 1));
            }
            r.mi_cuts[a] = mi_table2(pat);
            bt[0] += now() - c0;
            continue;
        }
        std::vector<std::array<double, 4>> mats(k);
        for (uint64_t q = 0; q < k; ++q) {
            uint64_t f = r.func_in_tet[s0 + ((flags & F_REVERSE) ? k - 1 - q : q)];
            for (int c = 0; c < 4; ++c) mats[q][c] = fv(tets[t][c], f);
        }
        r.mi_cuts[a] = mi_cut(mats);
        bt[k == 2 ? 0 : (k == 3 ? 1 : 2)] += now() - c0;
    }
    for (uint64_t t : act) {
        uint64_t k = r.start_index[t + 1] - r.start_index[t];
        ++r.nb[k <= 2 ? 0 : (k == 3 ? 1 : 2)];
    }
    double t3 = now();
    r.timings.push_back({"MI(2 func)", bt[0]});
    r.timings.push_back({"MI(3 func)", bt[1]});
    r.timings.push_back({"MI(>=4 func)", bt[2]});
    r.timings.push_back({"MI(wall)", t3 - t2});

    // src/extract_mesh.cpp:569-986
    std::unordered_map<uint64_t, uint64_t> on_vert;
    std::map<std::array<uint64_t, 4>, uint64_t> on_edge;
    std::map<std::array<uint64_t, 6>, uint64_t> on_face;
    std::map<std::array<long long, 3>, uint64_t> mat_of_bface;
    std::map<std::array<long long, 3>, std::vector<uint64_t>> mats_of_bface;
    for (uint64_t t = 0; t < T; ++t) {
        if (r.cut_index[t] < 0) continue;
        const MICut& C = r.mi_cuts[r.cut_index[t]];
        const uint64_t s0 = r.start_index[t];
        const auto& tet = tets[t];
        auto gmat = [&](int lab) { return r.func_in_tet[s0 + lab - 4]; };
        std::vector<char> mi_v(C.verts.size(), 0), mi_f(C.faces.size(), 0);
        for (size_t fi = 0; fi < C.faces.size(); ++fi)
            if (C.faces[fi].pos > 3) {
                mi_f[fi] = 1;
                for (int v : C.faces[fi].vs) mi_v[v] = 1;
            }
        std::vector<long long> gid(C.verts.size(), 0);
        for (size_t j = 0; j < C.verts.size(); ++j) {
            int bnd[4], nb = 0;
            uint64_t fm[4];
            int nm = 0;
            for (int q = 0; q < 4; ++q) {
                int l = C.verts[j][q];
                if (l > 3)
                    fm[nm++] = gmat(l);
                else
                    bnd[nb++] = l;
            }
            if (!mi_v[j]) {
                // a vertex on no interface face is a tet corner; corner v is
                // encoded as -v-1 until an interface face needs it (:670-687)
                int c = 0;
                while (c < 4 && std::find(bnd, bnd + nb, c) != bnd + nb) ++c;
                if (c >= 4 || nb != 3) throw std::runtime_error("oracle: non-corner vertex off the interface");
                gid[j] = -(long long)tet[c] - 1;
                continue;
            }
            if (nb == 2) {
                uint64_t e[2];
                int ne = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != bnd[0] && c != bnd[1]) e[ne++] = tet[c];
                uint64_t a = std::min(e[0], e[1]), b = std::max(e[0], e[1]);
                uint64_t m1 = std::min(fm[0], fm[1]), m2 = std::max(fm[0], fm[1]);
                auto it = on_edge.try_emplace({a, b, m1, m2}, r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 2;
                    iv.sv[0] = a;
                    iv.sv[1] = b;
                    iv.fn[0] = m1;
                    iv.fn[1] = m2;
                    iv.fn[2] = iv.fn[3] = U64NONE;
                    r.verts.push_back(iv);
                }
                gid[j] = (long long)it.first->second;
            } else if (nb == 1) {
                uint64_t e[3];
                int ne = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != bnd[0]) e[ne++] = tet[c];
                std::sort(e, e + 3);
                std::sort(fm, fm + 3);
                auto it = on_face.try_emplace({e[0], e[1], e[2], fm[0], fm[1], fm[2]}, r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 3;
                    for (int q = 0; q < 3; ++q) {
                        iv.sv[q] = e[q];
                        iv.fn[q] = fm[q];
                    }
                    iv.fn[3] = U64NONE;
                    r.verts.push_back(iv);
                }
                gid[j] = (long long)it.first->second;
            } else if (nb == 0) {
                IsoVert iv{};
                iv.tet = t;
                iv.tet_vert = j;
                iv.simplex_size = 4;
                for (int q = 0; q < 4; ++q) {
                    iv.sv[q] = tet[q];
                    iv.fn[q] = fm[q];
                }
                gid[j] = (long long)r.verts.size();
                r.verts.push_back(iv);
            } else {
                int c = 0;
                while (std::find(bnd, bnd + 3, c) != bnd + 3) ++c;
                auto it = on_vert.try_emplace(tet[c], r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 1;
                    iv.sv[0] = tet[c];
                    for (int q = 0; q < 4; ++q) iv.fn[q] = U64NONE;
                    r.verts.push_back(iv);
                }
                gid[j] = (long long)it.first->second;
            }
        }
        for (size_t fi = 0; fi < C.faces.size(); ++fi) {
            const MIFace& f = C.faces[fi];
            if (mi_f[fi]) {
                IsoFace F;
                for (int v : f.vs) F.vs.push_back((uint64_t)gid[v]);
                F.tet_faces.push_back({t, fi});
                F.f0 = gmat(f.pos);
                F.f1 = gmat(f.neg);
                r.faces.push_back(F);
                continue;
            }
            // tet-boundary face: an interface only when the materials on the
            // two sides differ (:833-951, same map/erase semantics)
            std::vector<long long> bv;
            for (int v : f.vs) {
                long long g = gid[v];
                if (g >= 0 && r.verts[g].simplex_size == 1)
                    bv.push_back(-(long long)r.verts[g].sv[0] - 1);
                else
                    bv.push_back(g);
            }
            long long mn = bv[0], mx = bv[0];
            size_t mp = 0;
            for (size_t q = 1; q < bv.size(); ++q) {
                if (bv[q] < mn) {
                    mn = bv[q];
                    mp = q;
                } else if (bv[q] > mx)
                    mx = bv[q];
            }
            long long sm = mx + 1;
            for (size_t q = 0; q < bv.size(); ++q)
                if (q != mp && bv[q] < sm) sm = bv[q];
            std::array<long long, 3> key{mn, sm, mx};
            uint64_t m = U64NONE;
            std::vector<uint64_t> ms;
            if (!C.dup) {
                m = gmat(f.neg);
            } else {
                int rep = C.grp[f.neg];
                int cnt = 0;
                for (int l = 4; l < C.nlabels; ++l) cnt += C.grp[l] == rep;
                if (cnt == 1)
                    m = gmat(f.neg);
                else
                    for (int l = 4; l < C.nlabels; ++l)
                        if (C.grp[l] == rep) ms.push_back(gmat(l));
            }
            bool iface = false;
            uint64_t other = U64NONE;
            if (m != U64NONE) {
                auto it = mat_of_bface.try_emplace(key, m);
                if (!it.second) {
                    if (it.first->second == m)
                        mat_of_bface.erase(it.first);
                    else {
                        iface = true;
                        other = it.first->second;
                    }
                } else {
                    auto it2 = mats_of_bface.find(key);
                    if (it2 != mats_of_bface.end()) {
                        bool has = std::find(it2->second.begin(), it2->second.end(), m) != it2->second.end();
                        if (has) {
                            mat_of_bface.erase(it.first);
                            mats_of_bface.erase(it2);
                        } else {
                            iface = true;
                            other = it2->second.front();
                        }
                    }
                }
            } else {
                auto it = mats_of_bface.try_emplace(key, ms);
                if (!it.second) {
                    bool common = false;
                    for (uint64_t a : it.first->second)
                        for (uint64_t b : ms) common |= a == b;
                    if (common)
                        mats_of_bface.erase(it.first);
                    else {
                        iface = true;
                        other = it.first->second.front();
                    }
                } else {
                    auto it2 = mat_of_bface.find(key);
                    if (it2 != mat_of_bface.end()) {
                        uint64_t mo = it2->second;
                        bool has = std::find(ms.begin(), ms.end(), mo) != ms.end();
                        if (has) {
                            mat_of_bface.erase(it2);
                            mats_of_bface.erase(it.first);
                        } else {
                            iface = true;
                            other = mo;
                        }
                    }
                }
            }
            if (!iface) continue;
            IsoFace F;
            for (size_t q = 0; q < f.vs.size(); ++q) {
                if (bv[q] < 0) {
                    uint64_t key1 = (uint64_t)(-bv[q] - 1);
                    auto it = on_vert.try_emplace(key1, r.verts.size());
                    if (it.second) {
                        IsoVert iv{};
                        iv.tet = t;
                        iv.tet_vert = f.vs[q];
                        iv.simplex_size = 1;
                        iv.sv[0] = key1;
                        for (int z = 0; z < 4; ++z) iv.fn[z] = U64NONE;
                        r.verts.push_back(iv);
                    }
                    F.vs.push_back(it.first->second);
                } else {
                    F.vs.push_back((uint64_t)gid[f.vs[q]]);
                }
            }
            F.tet_faces.push_back({t, fi});
            // the reference reads the boundary label through the CSR here
            // (:979); the material across the tet face is used instead
            F.f0 = other;
            F.f1 = gmat(f.neg);
            r.faces.push_back(F);
        }
    }
    double t4 = now();
    r.timings.push_back({"extract mesh", t4 - t3});

    // src/extract_mesh.cpp:1541-1637
    r.pts.resize(r.verts.size());
    for (size_t i = 0; i < r.verts.size(); ++i) {
        const IsoVert& iv = r.verts[i];
        auto P = [&](int q) { return pts[iv.sv[q]]; };
        if (iv.simplex_size == 1) {
            r.pts[i] = P(0);
        } else if (iv.simplex_size == 2) {
            double f1 = fv(iv.sv[0], iv.fn[0]) - fv(iv.sv[0], iv.fn[1]);
            double f2 = fv(iv.sv[1], iv.fn[0]) - fv(iv.sv[1], iv.fn[1]);
            double b0 = f2 / (f2 - f1);
            double b1 = 1 - b0;
            for (int a = 0; a < 3; ++a) r.pts[i][a] = b0 * P(0)[a] + b1 * P(1)[a];
        } else if (iv.simplex_size == 3) {
            double p1[3], p2[3];
            for (int q = 0; q < 3; ++q) {
                p1[q] = fv(iv.sv[q], iv.fn[0]) - fv(iv.sv[q], iv.fn[1]);
                p2[q] = fv(iv.sv[q], iv.fn[1]) - fv(iv.sv[q], iv.fn[2]);
            }
            double n1 = p1[2] * p2[1] - p1[1] * p2[2];
            double n2 = p1[0] * p2[2] - p1[2] * p2[0];
            double n3 = p1[1] * p2[0] - p1[0] * p2[1];
            double d = n1 + n2 + n3;
            double b0 = n1 / d, b1 = n2 / d, b2 = n3 / d;
            for (int a = 0; a < 3; ++a) r.pts[i][a] = b0 * P(0)[a] + b1 * P(1)[a] + b2 * P(2)[a];
        } else {
            double a1[4], a2[4], a3[4];
            for (int q = 0; q < 4; ++q) {
                a1[q] = fv(iv.sv[q], iv.fn[0]) - fv(iv.sv[q], iv.fn[1]);
                a2[q] = fv(iv.sv[q], iv.fn[1]) - fv(iv.sv[q], iv.fn[2]);
                a3[q] = fv(iv.sv[q], iv.fn[2]) - fv(iv.sv[q], iv.fn[3]);
            }
            double n1 = a1[3] * (a2[2] * a3[1] - a2[1] * a3[2]) + a1[2] * (a2[1] * a3[3] - a2[3] * a3[1]) +
                        a1[1] * (a2[3] * a3[2] - a2[2] * a3[3]);
            double n2 = a1[3] * (a2[0] * a3[2] - a2[2] * a3[0]) + a1[2] * (a2[3] * a3[0] - a2[0] * a3[3]) +
                        a1[0] * (a2[2] * a3[3] - a2[3] * a3[2]);
            double n3 = a1[3] * (a2[1] * a3[0] - a2[0] * a3[1]) + a1[1] * (a2[0] * a3[3] - a2[3] * a3[0]) +
                        a1[0] * (a2[3] * a3[1] - a2[1] * a3[3]);
            double n4 = a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]) + a1[1] * (a2[2] * a3[0] - a2[0] * a3[2]) +
                        a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]);
            double d = n1 + n2 + n3 + n4;
            double b0 = n1 / d, b1 = n2 / d, b2 = n3 / d, b3 = n4 / d;
            for (int a = 0; a < 3; ++a)
                r.pts[i][a] = b0 * P(0)[a] + b1 * P(1)[a] + b2 * P(2)[a] + b3 * P(3)[a];
        }
    }
    double t5 = now();
    r.timings.push_back({"compute xyz", t5 - t4});
    if (!(flags & F_KEEP_CUTS)) r.mi_cuts.clear(), r.mi_cuts.shrink_to_fit();
}

}  // namespace oracle
