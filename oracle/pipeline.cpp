// ORACLE — test infrastructure only.
// Stage-for-stage restatement of the reference's hot path; every function
// cites the reference lines it follows.
#include <algorithm>
#include <chrono>
#include <cmath>
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
struct Key3Hash {
    size_t operator()(const std::array<uint64_t, 3>& k) const {
        return std::hash<uint64_t>()(k[0] * 0x9E3779B97F4A7C15ull ^ (k[1] + 0x7F4A7C15ull) * 31 ^ k[2]);
    }
};
}  // namespace

// src/io.cpp:95-152 — regular grid, x-major vertex order, 5 tets per cube with
// the corner order of the reference, alternating by the parity of i+j+k.
void grid(int res, const double bmin[3], const double bmax[3], std::vector<std::array<double, 3>>& pts,
          std::vector<std::array<uint64_t, 4>>& tets) {
    const uint64_t N = (uint64_t)res + 1;
    pts.resize(N * N * N);
    auto coord = [&](double t, int a) { return t * (bmax[a] - bmin[a]) + bmin[a]; };
    for (uint64_t i = 0; i < N; ++i) {
        double x = coord(double(i) / double(N - 1), 0);
        for (uint64_t j = 0; j < N; ++j) {
            double y = coord(double(j) / double(N - 1), 1);
            for (uint64_t k = 0; k < N; ++k) {
                double z = coord(double(k) / double(N - 1), 2);
                pts[i * N * N + j * N + k] = {x, y, z};
            }
        }
    }
    const uint64_t R = (uint64_t)res;
    tets.resize(R * R * R * 5);
    for (uint64_t i = 0; i < R; ++i)
        for (uint64_t j = 0; j < R; ++j)
            for (uint64_t k = 0; k < R; ++k) {
                uint64_t t = (i * R * R + j * R + k) * 5;
                uint64_t v[8] = {i * N * N + j * N + k,
                                 (i + 1) * N * N + j * N + k,
                                 (i + 1) * N * N + (j + 1) * N + k,
                                 i * N * N + (j + 1) * N + k,
                                 i * N * N + j * N + k + 1,
                                 (i + 1) * N * N + j * N + k + 1,
                                 (i + 1) * N * N + (j + 1) * N + k + 1,
                                 i * N * N + (j + 1) * N + k + 1};
                static const int even[5][4] = {{4, 6, 1, 3}, {6, 3, 4, 7}, {1, 3, 0, 4}, {3, 1, 2, 6}, {4, 1, 6, 5}};
                static const int odd[5][4] = {{7, 0, 2, 5}, {2, 3, 0, 7}, {5, 7, 0, 4}, {7, 2, 6, 5}, {0, 1, 2, 5}};
                const int(*tab)[4] = ((i + j + k) % 2 == 0) ? even : odd;
                for (int q = 0; q < 5; ++q)
                    tets[t + q] = {v[tab[q][0]], v[tab[q][1]], v[tab[q][2]], v[tab[q][3]]};
            }
}

// Function values (app/implicit_arrangement.cpp:59, external load_functions).
// The formulas are this project's: spheres and cylinders are positive inside
// (tests/test_implicit_networks.cpp:433 pins the sphere sign), planes are
// (x - p).n.  Evaluated strictly left to right without contraction.
static double eval_one(const Func& f, const std::array<double, 3>& x) {
    const double* p = f.p;
    switch (f.type) {
        case 0:
            return (x[0] - p[0]) * p[3] + (x[1] - p[1]) * p[4] + (x[2] - p[2]) * p[5];
        case 1: {
            double d0 = x[0] - p[0], d1 = x[1] - p[1], d2 = x[2] - p[2];
            double q = d0 * d0 + d1 * d1 + d2 * d2;
            return p[4] != 0.0 ? p[3] * p[3] - q : p[3] - std::sqrt(q);
        }
        case 2: {
            double d0 = x[0] - p[0], d1 = x[1] - p[1], d2 = x[2] - p[2];
            double h = d0 * p[3] + d1 * p[4] + d2 * p[5];
            double e0 = d0 - h * p[3], e1 = d1 - h * p[4], e2 = d2 - h * p[5];
            double v = p[6] - std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            return p[7] != 0.0 ? -v : v;
        }
        case 3: {
            double d0 = x[0] - p[0], d1 = x[1] - p[1], d2 = x[2] - p[2];
            double h = d0 * p[3] + d1 * p[4] + d2 * p[5];
            double e0 = d0 - h * p[3], e1 = d1 - h * p[4], e2 = d2 - h * p[5];
            double u = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2) - p[6];
            double v = p[7] - std::sqrt(u * u + h * h);
            return p[8] != 0.0 ? -v : v;
        }
        case 4:
            return 0.0;
    }
    throw std::runtime_error("oracle: unknown function type");
}

void eval(const std::vector<Func>& funcs, const std::vector<std::array<double, 3>>& pts, std::vector<double>& vals) {
    const size_t F = funcs.size();
    vals.resize(pts.size() * F);
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < (long long)pts.size(); ++v)
        for (size_t j = 0; j < F; ++j) vals[v * F + j] = eval_one(funcs[j], pts[v]);
}

// ------------------------------------------------------------------ tables
namespace {
template <class Cut, class Fn>
std::vector<Cut> build_table(Fn cut, bool mi) {
    std::vector<Cut> tab(81);
    for (int pat = 0; pat < 81; ++pat) {
        std::array<double, 4> row;
        int p = pat;
        for (int c = 0; c < 4; ++c, p /= 3) row[c] = double(p % 3 - 1);
        if (mi)
            tab[pat] = cut(std::vector<std::array<double, 4>>{{0, 0, 0, 0}, row});
        else
            tab[pat] = cut(std::vector<std::array<double, 4>>{row});
    }
    return tab;
}
}  // namespace

const IACut& ia_table1(int pattern) {
    static const std::vector<IACut> tab = build_table<IACut>(ia_cut, false);
    return tab[pattern];
}
const MICut& mi_table2(int pattern) {
    static const std::vector<MICut> tab = build_table<MICut>(mi_cut, true);
    return tab[pattern];
}

// ------------------------------------------------------------------ IA
void run_ia(const std::vector<std::array<double, 3>>& pts, const std::vector<std::array<uint64_t, 4>>& tets,
            const std::vector<double>& vals, uint64_t F, int flags, Result& r) {
    r = Result();
    r.mi = false;
    r.nv = pts.size();
    r.nt = tets.size();
    r.nf = F;
    const uint64_t V = r.nv, T = r.nt;
    auto fv = [&](uint64_t v, uint64_t f) { return vals[v * F + f]; };

    // src/implicit_arrangement.cpp:53-77 — signs; the degenerate count counts
    // zero entries, not vertices (:72)
    double t0 = now();
    std::vector<int8_t> sg(V * F);
    for (uint64_t v = 0; v < V; ++v)
        for (uint64_t f = 0; f < F; ++f) {
            double x = fv(v, f);
            int s = x > 0 ? 1 : (x < 0 ? -1 : 0);
            sg[v * F + f] = (int8_t)s;
            if (s == 0) ++r.num_degenerate;
        }
    double t1 = now();
    r.timings.push_back({"func signs", t1 - t0});

    // :82-116 — a function is active unless all four corners share one strict sign
    r.start_index.reserve(T + 1);
    r.start_index.push_back(0);
    for (uint64_t t = 0; t < T; ++t) {
        for (uint64_t f = 0; f < F; ++f) {
            int pc = 0, nc = 0;
            for (int c = 0; c < 4; ++c) {
                int s = sg[tets[t][c] * F + f];
                pc += s == 1;
                nc += s == -1;
            }
            if (pc < 4 && nc < 4) r.func_in_tet.push_back(f);
        }
        if (r.func_in_tet.size() > r.start_index.back()) ++r.num_intersecting;
        r.start_index.push_back(r.func_in_tet.size());
    }
    double t2 = now();
    r.timings.push_back({"filter", t2 - t1});

    // :244-306 — one cut per active tet, planes in CSR order
    r.cut_index.assign(T, -1);
    std::vector<uint64_t> act;
    for (uint64_t t = 0; t < T; ++t)
        if (r.start_index[t + 1] > r.start_index[t]) {
            r.cut_index[t] = (int64_t)act.size();
            act.push_back(t);
        }
    r.ia_cuts.resize(act.size());
    if (flags & F_LOOKUP) ia_table1(0);
    double bt[3] = {0, 0, 0};
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bt[:3])
    for (long long a = 0; a < (long long)act.size(); ++a) {
        uint64_t t = act[a];
        double c0 = now();
        uint64_t s0 = r.start_index[t], k = r.start_index[t + 1] - s0;
        if ((flags & F_LOOKUP) && k == 1) {
            const uint64_t f = r.func_in_tet[s0];
            int pat = 0;
            for (int c = 3; c >= 0; --c) pat = pat * 3 + (sg[tets[t][c] * F + f] + 1);
            r.ia_cuts[a] = ia_table1(pat);
            bt[0] += now() - c0;
            continue;
        }
        std::vector<std::array<double, 4>> planes(k);
        for (uint64_t q = 0; q < k; ++q) {
            uint64_t f = r.func_in_tet[s0 + ((flags & F_REVERSE) ? k - 1 - q : q)];
            for (int c = 0; c < 4; ++c) planes[q][c] = fv(tets[t][c], f);
        }
        r.ia_cuts[a] = ia_cut(planes);
        bt[k == 1 ? 0 : (k == 2 ? 1 : 2)] += now() - c0;
    }
    for (uint64_t t : act) {
        uint64_t k = r.start_index[t + 1] - r.start_index[t];
        ++r.nb[k == 1 ? 0 : (k == 2 ? 1 : 2)];
    }
    double t3 = now();
    r.timings.push_back({"simp_arr(1 func)", bt[0]});
    r.timings.push_back({"simp_arr(2 func)", bt[1]});
    r.timings.push_back({"simp_arr(>=3 func)", bt[2]});
    r.timings.push_back({"simp_arr(wall)", t3 - t2});

    // src/extract_mesh.cpp:10-265 — first-occurrence numbering over (tet, local)
    std::unordered_map<uint64_t, uint64_t> on_vert;
    std::map<std::array<uint64_t, 3>, uint64_t> on_edge;
    std::map<std::array<uint64_t, 5>, uint64_t> on_face;
    std::unordered_map<std::array<uint64_t, 3>, uint64_t, Key3Hash> face_on_bndry;
    for (uint64_t t = 0; t < T; ++t) {
        if (r.cut_index[t] < 0) continue;
        const IACut& C = r.ia_cuts[r.cut_index[t]];
        const uint64_t s0 = r.start_index[t];
        const auto& tet = tets[t];
        std::vector<char> iso_v(C.verts.size(), 0), iso_f(C.faces.size(), 0);
        for (size_t fi = 0; fi < C.faces.size(); ++fi) {
            const IAFace& f = C.faces[fi];
            bool iso = f.plane > 3;
            if (C.dup && !iso) {
                int rep = C.grp[f.plane];
                for (int q = 0; q < C.nplanes; ++q) iso |= (q > 3 && C.grp[q] == rep);
            }
            if (iso) {
                iso_f[fi] = 1;
                for (int v : f.vs) iso_v[v] = 1;
            }
        }
        std::vector<uint64_t> gid(C.verts.size(), UINT64_MAX);
        for (size_t j = 0; j < C.verts.size(); ++j) {
            if (!iso_v[j]) continue;
            uint64_t impl[3];
            int bnd[3], ni = 0, nb = 0;
            for (int q = 0; q < 3; ++q) {
                int p = C.verts[j][q];
                if (p > 3)
                    impl[ni++] = r.func_in_tet[s0 + p - 4];
                else
                    bnd[nb++] = p;
            }
            if (nb == 2) {
                uint64_t e[2];
                int ne = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != bnd[0] && c != bnd[1]) e[ne++] = tet[c];
                uint64_t a = std::min(e[0], e[1]), b = std::max(e[0], e[1]);
                auto it = on_edge.try_emplace({a, b, impl[0]}, r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 2;
                    iv.sv[0] = a;
                    iv.sv[1] = b;
                    iv.fn[0] = impl[0];
                    iv.fn[1] = iv.fn[2] = iv.fn[3] = UINT64_MAX;
                    r.verts.push_back(iv);
                }
                gid[j] = it.first->second;
            } else if (nb == 1) {
                uint64_t e[3];
                int ne = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != bnd[0]) e[ne++] = tet[c];
                std::sort(e, e + 3);
                auto it = on_face.try_emplace({e[0], e[1], e[2], impl[0], impl[1]}, r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 3;
                    iv.sv[0] = e[0];
                    iv.sv[1] = e[1];
                    iv.sv[2] = e[2];
                    iv.fn[0] = impl[0];
                    iv.fn[1] = impl[1];
                    iv.fn[2] = iv.fn[3] = UINT64_MAX;
                    r.verts.push_back(iv);
                }
                gid[j] = it.first->second;
            } else if (nb == 0) {
                IsoVert iv{};
                iv.tet = t;
                iv.tet_vert = j;
                iv.simplex_size = 4;
                for (int c = 0; c < 4; ++c) iv.sv[c] = tet[c];
                iv.fn[0] = impl[0];
                iv.fn[1] = impl[1];
                iv.fn[2] = impl[2];
                iv.fn[3] = UINT64_MAX;
                gid[j] = r.verts.size();
                r.verts.push_back(iv);
            } else {
                int c = 0;
                while (c == bnd[0] || c == bnd[1] || c == bnd[2]) ++c;
                auto it = on_vert.try_emplace(tet[c], r.verts.size());
                if (it.second) {
                    IsoVert iv{};
                    iv.tet = t;
                    iv.tet_vert = j;
                    iv.simplex_size = 1;
                    iv.sv[0] = tet[c];
                    iv.fn[0] = iv.fn[1] = iv.fn[2] = iv.fn[3] = UINT64_MAX;
                    r.verts.push_back(iv);
                }
                gid[j] = it.first->second;
            }
        }
        for (size_t fi = 0; fi < C.faces.size(); ++fi) {
            if (!iso_f[fi]) continue;
            const IAFace& f = C.faces[fi];
            std::vector<uint64_t> fvs;
            for (int v : f.vs) fvs.push_back(gid[v]);
            // func_index.first (src/extract_mesh.cpp:249,258).  A face on a tet
            // face whose group holds an implicit plane takes the smallest
            // implicit plane of the group (the reference indexes with the
            // boundary id there, outside the tet's CSR range).
            int sp = f.plane;
            if (sp <= 3) {
                int rep = C.grp[sp];
                for (int q = 4; q < C.nplanes; ++q)
                    if (C.grp[q] == rep) {
                        sp = q;
                        break;
                    }
            }
            uint64_t fn = r.func_in_tet[s0 + sp - 4];
            if (f.neg == NONE) {
                // compute_iso_face_key (src/extract_mesh.h:68-92)
                uint64_t mn = fvs[0], mx = fvs[0];
                size_t mp = 0;
                for (size_t q = 1; q < fvs.size(); ++q) {
                    if (fvs[q] < mn) {
                        mn = fvs[q];
                        mp = q;
                    } else if (fvs[q] > mx)
                        mx = fvs[q];
                }
                uint64_t sm = mx + 1;
                for (size_t q = 0; q < fvs.size(); ++q)
                    if (q != mp && fvs[q] < sm) sm = fvs[q];
                auto it = face_on_bndry.try_emplace({mn, sm, mx}, r.faces.size());
                if (it.second) {
                    IsoFace F;
                    F.vs = fvs;
                    F.tet_faces.push_back({t, fi});
                    F.f0 = fn;
                    F.f1 = UINT64_MAX;
                    r.faces.push_back(F);
                } else {
                    r.faces[it.first->second].tet_faces.push_back({t, fi});
                }
            } else {
                IsoFace F;
                F.vs = fvs;
                F.tet_faces.push_back({t, fi});
                F.f0 = fn;
                F.f1 = UINT64_MAX;
                r.faces.push_back(F);
            }
        }
    }
    double t4 = now();
    r.timings.push_back({"extract mesh", t4 - t3});

    // src/extract_mesh.cpp:1446-1538 with src/extract_mesh.h:111-159
    r.pts.resize(r.verts.size());
    for (size_t i = 0; i < r.verts.size(); ++i) {
        const IsoVert& iv = r.verts[i];
        auto P = [&](int q) { return pts[iv.sv[q]]; };
        if (iv.simplex_size == 1) {
            r.pts[i] = P(0);
        } else if (iv.simplex_size == 2) {
            double f1 = fv(iv.sv[0], iv.fn[0]), f2 = fv(iv.sv[1], iv.fn[0]);
            double b0 = f2 / (f2 - f1);
            double b1 = 1 - b0;
            for (int a = 0; a < 3; ++a) r.pts[i][a] = b0 * P(0)[a] + b1 * P(1)[a];
        } else if (iv.simplex_size == 3) {
            double p1[3], p2[3];
            for (int q = 0; q < 3; ++q) {
                p1[q] = fv(iv.sv[q], iv.fn[0]);
                p2[q] = fv(iv.sv[q], iv.fn[1]);
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
                a1[q] = fv(iv.sv[q], iv.fn[0]);
                a2[q] = fv(iv.sv[q], iv.fn[1]);
                a3[q] = fv(iv.sv[q], iv.fn[2]);
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
    if (!(flags & F_KEEP_CUTS)) r.ia_cuts.clear(), r.ia_cuts.shrink_to_fit();
}

}  // namespace oracle
