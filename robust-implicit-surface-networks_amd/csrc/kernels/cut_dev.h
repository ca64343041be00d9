// Per-tet cell complexes on the device: implicit arrangement (IA) and
// material interface (MI), one thread per tetrahedron, state in fixed-size
// thread-private arrays (capacities are template parameters chosen per
// active-function bucket).
//
// Both follow the incremental-insertion contract of the reference's external
// simplicial_arrangement library (compute_arrangement /
// compute_material_interface, called at src/implicit_arrangement.cpp:283 and
// src/material_interface.cpp:327):
//   IA: planes 0..3 are the tet faces, implicit planes 4.. are inserted in
//       ascending function order; a vertex is a sorted plane triple; a split
//       face keeps its id for the positive part and appends the negative part;
//       a split cell keeps its id for the positive side, appends the negative
//       side and appends one face on the new plane; a plane containing an
//       existing face only joins that face's coincidence group.
//   MI: labels 0..3 are the tet faces, materials 4.. are inserted in ascending
//       order starting from the tet filled with material 4; each insertion
//       clips faces to where the new material loses, merges the tet-face
//       parts it wins into one face per tet face, adds one face per crossed
//       cell and drops fully dominated cells and unreferenced vertices.
// The local numbering produced here is the one the rest of the pipeline (and
// the host graph stages) rely on.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "exact_dev.h"

namespace rin {
namespace dev {

constexpr uint8_t NIL = 0xFF;

template <int MV_, int MF_, int MFV_, int MP_, class Idx_ = uint8_t>
struct Caps {
    static constexpr int MV = MV_;    // vertices
    static constexpr int MF = MF_;    // faces
    static constexpr int MFV = MFV_;  // vertices per face
    static constexpr int MP = MP_;    // labels (4 + active functions)
    using Idx = Idx_;                 // vertex / face / cell / label ids
    static constexpr Idx NIL = Idx(~Idx(0));
    // cells (IA): 8 per label with 8-bit ids, 64 per label with 16-bit ids
    static constexpr int MC = sizeof(Idx) == 1 ? (MP * 8 < 254 ? MP * 8 : 254) : (MP * 64 < 65534 ? MP * 64 : 65534);
};

template <class C>
struct Complex {
    using I = typename C::Idx;
    int nv, nf, nc;
    bool dup, err;
    I vl[C::MV][4];          // IA: 3 plane ids (+NIL); MI: 4 labels
    I fa[C::MF], fb[C::MF], fc[C::MF], fn[C::MF];  // IA: plane,pos,neg ; MI: pos,neg,-
    I fv[C::MF][C::MFV];
    I fe[C::MF][C::MFV];     // third plane / label of edge fv[i] -> fv[i+1]
    I cells[C::MP];          // MI: material label per cell
    I grp[C::MP], ori[C::MP];
};

// The insertion's working arrays that scale with the capacities (the rest
// are a few dozen bytes).  Kernels keep them, like the complex, in LDS or in
// device memory, so the per-lane private segment stays small.
template <class C>
struct InsertTemps {
    using I = typename C::Idx;
    static constexpr int PE = 4 * C::MFV * 4;  // MI: tet-face edges won by the new material
    int8_t s[C::MV];                           // sign of the new plane / material at each vertex
    I cu[C::MV], cw[C::MV], cx[C::MV];         // crossing edges (u < w) and their new vertex
    I pp[PE], pa[PE], pb[PE], pl[PE];          // (tet face, from, to, edge label)
    I remap[C::MV];                            // vertex renumbering after an MI insertion
    bool used[C::MV];                          // vertices still referenced
    bool cpos[C::MP], cneg[C::MP], consumed[C::MP], split[C::MP];
};

template <class T>
__host__ __device__ inline void sort3(T& a, T& b, T& c) {
    T t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
}
template <class T>
__host__ __device__ inline void sort4(T* v) {
    for (int i = 1; i < 4; ++i)
        for (int j = i; j > 0 && v[j - 1] > v[j]; --j) {
            T t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
}

// ------------------------------------------------------------------ IA

// the tet before any insertion: 4 corners, 4 faces, 1 cell; err when k
// planes exceed the label capacity
template <class C>
__host__ __device__ void ia_init(Complex<C>& X, int k) {
    X.err = false;
    X.dup = false;
    X.nv = 4;
    const unsigned char corner[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int v = 0; v < 4; ++v) {
        for (int q = 0; q < 3; ++q) X.vl[v][q] = corner[v][q];
        X.vl[v][3] = C::NIL;
    }
    const unsigned char cyc[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
    X.nf = 4;
    for (int i = 0; i < 4; ++i) {
        X.fa[i] = i;
        X.fb[i] = 0;
        X.fc[i] = C::NIL;
        X.fn[i] = 3;
        for (int q = 0; q < 3; ++q) {
            X.fv[i][q] = cyc[i][q];
            X.fe[i][q] = cyc[i][(q + 2) % 3];
        }
    }
    X.nc = 1;
    for (int p = 0; p < C::MP; ++p) {
        X.grp[p] = p < 4 ? p : C::NIL;
        X.ori[p] = 1;
    }
    if (4 + k > C::MP) X.err = true;
}

// sign of plane j at every vertex (exact)
template <class C>
__host__ __device__ void ia_classify(const double (*P)[4], int j, Complex<C>& X, int8_t* s) {
    for (int v = 0; v < X.nv; ++v) s[v] = (int8_t)orient_ia(P, X.vl[v][0], X.vl[v][1], X.vl[v][2], P[j], X.err);
}

// inserts plane Pid given its sign s[v] at every vertex; the result depends
// on the complex and the signs only
template <class C>
__host__ __device__ void ia_insert(Complex<C>& X, typename C::Idx Pid, int8_t* s, InsertTemps<C>& T) {
    using I = typename C::Idx;
    {
        bool any_pos = false, any_neg = false;
        for (int v = 0; v < X.nv; ++v) {
            any_pos |= s[v] > 0;
            any_neg |= s[v] < 0;
        }
        int cop = -1;
        for (int f = 0; f < X.nf && cop < 0; ++f) {
            bool allz = true;
            for (int q = 0; q < X.fn[f]; ++q) allz &= s[X.fv[f][q]] == 0;
            if (allz) cop = f;
        }
        if (cop >= 0) {
            const I cell = X.fb[cop];
            int sv = 0;
            for (int g = 0; g < X.nf && sv == 0; ++g) {
                if (X.fb[g] != cell && X.fc[g] != cell) continue;
                for (int q = 0; q < X.fn[g]; ++q)
                    if (s[X.fv[g][q]] != 0) {
                        sv = s[X.fv[g][q]];
                        break;
                    }
            }
            const I pl = X.fa[cop];
            X.grp[Pid] = X.grp[pl];
            X.ori[Pid] = (sv > 0) ? X.ori[pl] : (I)(1 - X.ori[pl]);
            X.dup = true;
            return;
        }
        X.grp[Pid] = Pid;
        X.ori[Pid] = 1;
        if (!any_pos || !any_neg) return;
    }
    // one new vertex per edge with a strict sign change
    I* cu = T.cu;
    I* cw = T.cw;
    I* cx = T.cx;
    int ncross = 0;
    for (int f = 0; f < X.nf; ++f) {
        const int n = X.fn[f];
        for (int i = 0; i < n; ++i) {
            I u = X.fv[f][i], w = X.fv[f][(i + 1) % n];
            if (s[u] * s[w] >= 0) continue;
            I a = u < w ? u : w, b = u < w ? w : u;
            bool have = false;
            for (int q = 0; q < ncross; ++q) have |= (cu[q] == a && cw[q] == b);
            if (have) continue;
            if (X.nv >= C::MV) {
                X.err = true;
                return;
            }
            I t0 = X.fa[f], t1 = X.fe[f][i], t2 = Pid;
            sort3(t0, t1, t2);
            X.vl[X.nv][0] = t0;
            X.vl[X.nv][1] = t1;
            X.vl[X.nv][2] = t2;
            X.vl[X.nv][3] = C::NIL;
            s[X.nv] = 0;
            cu[ncross] = a;
            cw[ncross] = b;
            cx[ncross] = (I)X.nv;
            ++ncross;
            ++X.nv;
        }
    }
    auto cross_of = [&](I u, I w) -> I {
        I a = u < w ? u : w, b = u < w ? w : u;
        for (int q = 0; q < ncross; ++q)
            if (cu[q] == a && cw[q] == b) return cx[q];
        return C::NIL;
    };

    // split faces: positive part keeps the id, negative part appended
    const int nf0 = X.nf;
    for (int f = 0; f < nf0; ++f) {
        const int n = X.fn[f];
        bool fp = false, fq = false;
        for (int q = 0; q < n; ++q) {
            fp |= s[X.fv[f][q]] > 0;
            fq |= s[X.fv[f][q]] < 0;
        }
        if (!(fp && fq)) continue;
        if (X.nf >= C::MF) {
            X.err = true;
            return;
        }
        I av[2 * C::MFV], ae[2 * C::MFV];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            I u = X.fv[f][i], w = X.fv[f][(i + 1) % n];
            av[m] = u;
            ae[m] = X.fe[f][i];
            ++m;
            if (s[u] * s[w] < 0) {
                av[m] = cross_of(u, w);
                ae[m] = X.fe[f][i];
                ++m;
            }
        }
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            const int dst = (sgn > 0) ? f : X.nf;
            int idx[2 * C::MFV], ni = 0;
            for (int i = 0; i < m; ++i)
                if (s[av[i]] * sgn >= 0) idx[ni++] = i;
            if (ni > C::MFV) {
                X.err = true;
                return;
            }
            I nvv[C::MFV], nee[C::MFV];
            for (int q = 0; q < ni; ++q) {
                int ia = idx[q], ib = idx[(q + 1) % ni];
                nvv[q] = av[ia];
                nee[q] = ((ia + 1) % m == ib) ? ae[ia] : Pid;
            }
            for (int q = 0; q < ni; ++q) {
                X.fv[dst][q] = nvv[q];
                X.fe[dst][q] = nee[q];
            }
            X.fn[dst] = (I)ni;
            if (sgn < 0) {
                X.fa[dst] = X.fa[f];
                X.fb[dst] = X.fb[f];
                X.fc[dst] = X.fc[f];
            }
        }
        ++X.nf;
    }

    // split cells: positive side keeps the id, negative side appended,
    // one face on the new plane per split cell
    const int nc0 = X.nc;
    for (int c = 0; c < nc0; ++c) {
        bool cp = false, cq = false;
        for (int g = 0; g < X.nf; ++g) {
            if (X.fb[g] != c && X.fc[g] != c) continue;
            for (int q = 0; q < X.fn[g]; ++q) {
                cp |= s[X.fv[g][q]] > 0;
                cq |= s[X.fv[g][q]] < 0;
            }
        }
        if (!(cp && cq)) continue;
        if (X.nf >= C::MF || X.nc >= C::MC) {
            X.err = true;
            return;
        }
        const I c2 = (I)X.nc++;
        for (int g = 0; g < X.nf; ++g) {
            if (X.fb[g] != c && X.fc[g] != c) continue;
            bool gn = false;
            for (int q = 0; q < X.fn[g]; ++q) gn |= s[X.fv[g][q]] < 0;
            if (!gn) continue;
            if (X.fb[g] == c) X.fb[g] = c2;
            if (X.fc[g] == c) X.fc[g] = c2;
        }
        I ea[C::MFV], eb[C::MFV], el[C::MFV];
        int ne = 0;
        for (int g = 0; g < X.nf; ++g) {
            if (X.fb[g] != c && X.fc[g] != c) continue;
            const int n = X.fn[g];
            for (int i = 0; i < n; ++i) {
                I a = X.fv[g][i], b = X.fv[g][(i + 1) % n];
                if (s[a] != 0 || s[b] != 0) continue;
                if (ne >= C::MFV) {
                    X.err = true;
                    return;
                }
                if (X.fb[g] == c) {
                    ea[ne] = b;
                    eb[ne] = a;
                } else {
                    ea[ne] = a;
                    eb[ne] = b;
                }
                el[ne] = X.fa[g];
                ++ne;
            }
        }
        if (ne < 3) {
            X.err = true;
            return;
        }
        const int F = X.nf++;
        X.fa[F] = Pid;
        X.fb[F] = (I)c;
        X.fc[F] = c2;
        int cur = 0, cnt = 0;
        for (;;) {
            X.fv[F][cnt] = ea[cur];
            X.fe[F][cnt] = el[cur];
            ++cnt;
            int nxt = -1;
            for (int q = 0; q < ne; ++q)
                if (ea[q] == eb[cur]) {
                    nxt = q;
                    break;
                }
            if (nxt < 0 || cnt > ne) {
                X.err = true;
                return;
            }
            cur = nxt;
            if (cur == 0) break;
        }
        if (cnt != ne) {
            X.err = true;
            return;
        }
        X.fn[F] = (I)cnt;
    }
}

template <class C>
__host__ __device__ void ia_cut(const double (*P)[4], int k, Complex<C>& X, InsertTemps<C>& T) {
    ia_init(X, k);
    if (X.err) return;
    for (int j = 0; j < k; ++j) {
        ia_classify(P, j, X, T.s);
        if (X.err) return;
        ia_insert(X, (typename C::Idx)(4 + j), T.s, T);
        if (X.err) return;
    }
}

// host forms (the table builders): per-thread working arrays
template <class C>
void ia_insert(Complex<C>& X, typename C::Idx Pid, int8_t* s) {
    static thread_local InsertTemps<C> T;
    ia_insert(X, Pid, s, T);
}
template <class C>
void ia_cut(const double (*P)[4], int k, Complex<C>& X) {
    static thread_local InsertTemps<C> T;
    ia_cut(P, k, X, T);
}

// record layout: include/rin/cut_record.h; returns words written or -1
template <class C>
__host__ __device__ int write_record_ia(const Complex<C>& X, int k, uint16_t* out, int cap) {
    int p = 8;
    const int L = 4 + k;
    auto put = [&](uint16_t w) {
        if (p < cap) out[p] = w;
        ++p;
    };
    for (int v = 0; v < X.nv; ++v) {
        put(X.vl[v][0]);
        put(X.vl[v][1]);
        put(X.vl[v][2]);
        put(0xFFFF);
    }
    for (int f = 0; f < X.nf; ++f) {
        put(X.fa[f]);
        put(X.fb[f]);
        put(X.fc[f] == C::NIL ? 0xFFFF : X.fc[f]);
        put(X.fn[f]);
        for (int q = 0; q < X.fn[f]; ++q) put(X.fv[f][q]);
    }
    if (X.dup) {
        for (int l = 0; l < L; ++l) put(X.grp[l]);
        for (int l = 0; l < L; ++l) put(X.ori[l]);
    }
    if (p > cap) return -1;
    out[0] = 1;
    out[1] = (uint16_t)L;
    out[2] = (uint16_t)X.nv;
    out[3] = (uint16_t)X.nf;
    out[4] = (uint16_t)X.nc;
    out[5] = X.dup ? 1 : 0;
    out[6] = (uint16_t)p;
    out[7] = 0;
    return p;
}

// ------------------------------------------------------------------ MI

// chain directed edges (ea->eb, label el) into one cycle from edge 0
template <int CAP, class T>
__host__ __device__ inline int chain_edges(const T* ea, const T* eb, const T* el, int ne, T* ov, T* oe) {
    if (ne <= 0 || ne > CAP) return -1;
    int cur = 0, cnt = 0;
    for (;;) {
        ov[cnt] = ea[cur];
        oe[cnt] = el[cur];
        ++cnt;
        int nxt = -1;
        for (int q = 0; q < ne; ++q)
            if (ea[q] == eb[cur]) {
                nxt = q;
                break;
            }
        if (nxt < 0 || cnt > ne) return -1;
        cur = nxt;
        if (cur == 0) break;
    }
    return cnt == ne ? cnt : -1;
}

// the tet filled with material 4 (the first of k materials); err when k
// exceeds the label capacity
template <class C>
__host__ __device__ void mi_init(Complex<C>& X, int k) {
    X.err = false;
    X.dup = false;
    for (int p = 0; p < C::MP; ++p) X.grp[p] = p < 4 ? p : C::NIL;
    if (4 + k > C::MP || k < 1) {
        X.err = k > 0;
        X.nv = X.nf = X.nc = 0;
        return;
    }
    X.nv = 4;
    for (int i = 0; i < 4; ++i) {
        int n = 0;
        for (int q = 0; q < 4; ++q)
            if (q != i) X.vl[i][n++] = q;
        X.vl[i][3] = 4;
    }
    const unsigned char cyc[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
    X.nf = 4;
    for (int i = 0; i < 4; ++i) {
        X.fa[i] = i;
        X.fb[i] = 4;
        X.fc[i] = C::NIL;
        X.fn[i] = 3;
        for (int q = 0; q < 3; ++q) {
            X.fv[i][q] = cyc[i][q];
            X.fe[i][q] = cyc[i][(q + 2) % 3];
        }
    }
    X.nc = 1;
    X.cells[0] = 4;
    X.grp[4] = 4;
}

// sign of material j against each vertex's materials (exact): > 0 where j wins
template <class C>
__host__ __device__ void mi_classify(const double (*P)[4], int j, Complex<C>& X, int8_t* s) {
    for (int v = 0; v < X.nv; ++v) {
        int lab[4] = {X.vl[v][0], X.vl[v][1], X.vl[v][2], X.vl[v][3]};
        s[v] = (int8_t)orient_mi(P, lab, P[j], X.err);
    }
}

// inserts material m given its sign s[v] at every vertex; the result depends
// on the complex and the signs only
template <class C>
__host__ __device__ void mi_insert(Complex<C>& X, typename C::Idx m, int8_t* s, InsertTemps<C>& T) {
    using I = typename C::Idx;
    constexpr int PE = InsertTemps<C>::PE;
    const int nv0 = X.nv;
    bool any_pos = false;
    for (int v = 0; v < nv0; ++v) any_pos |= s[v] > 0;
    bool* cpos = T.cpos;
    bool* cneg = T.cneg;
    for (int l = 0; l < C::MP; ++l) cpos[l] = cneg[l] = false;
    for (int f = 0; f < X.nf; ++f)
        for (int side = 0; side < 2; ++side) {
            I lab = side ? X.fb[f] : X.fa[f];
            if (lab < 4) continue;
            for (int q = 0; q < X.fn[f]; ++q) {
                cpos[lab] |= s[X.fv[f][q]] > 0;
                cneg[lab] |= s[X.fv[f][q]] < 0;
            }
        }
    if (!any_pos) {
        X.grp[m] = m;
        for (int c = 0; c < X.nc; ++c) {
            I L = X.cells[c];
            if (!cpos[L] && !cneg[L]) {
                X.grp[m] = X.grp[L];
                X.dup = true;
                break;
            }
        }
        return;
    }
    X.grp[m] = m;
    bool* consumed = T.consumed;
    bool* split = T.split;
    for (int l = 0; l < C::MP; ++l) {
        consumed[l] = cpos[l] && !cneg[l];
        split[l] = cpos[l] && cneg[l];
    }
    I* cu = T.cu;
    I* cw = T.cw;
    I* cx = T.cx;
    int ncross = 0;
    for (int f = 0; f < X.nf; ++f) {
        const int n = X.fn[f];
        for (int i = 0; i < n; ++i) {
            I u = X.fv[f][i], w = X.fv[f][(i + 1) % n];
            if (s[u] * s[w] >= 0) continue;
            I a = u < w ? u : w, b = u < w ? w : u;
            bool have = false;
            for (int q = 0; q < ncross; ++q) have |= (cu[q] == a && cw[q] == b);
            if (have) continue;
            if (X.nv >= C::MV) {
                X.err = true;
                return;
            }
            I t[4] = {X.fa[f], X.fb[f], X.fe[f][i], m};
            sort4(t);
            for (int q = 0; q < 4; ++q) X.vl[X.nv][q] = t[q];
            s[X.nv] = 0;
            cu[ncross] = a;
            cw[ncross] = b;
            cx[ncross] = (I)X.nv;
            ++ncross;
            ++X.nv;
        }
    }
    auto cross_of = [&](I u, I w) -> I {
        I a = u < w ? u : w, b = u < w ? w : u;
        for (int q = 0; q < ncross; ++q)
            if (cu[q] == a && cw[q] == b) return cx[q];
        return C::NIL;
    };
    // tet-face edges won by m, per tet face: (plane, from, to, label)
    I* pp = T.pp;
    I* pa = T.pa;
    I* pb = T.pb;
    I* pl = T.pl;
    int npe = 0;
    auto add_plus = [&](I plane, I a, I b, I l) {
        if (npe >= PE) {
            X.err = true;
            return;
        }
        pp[npe] = plane;
        pa[npe] = a;
        pb[npe] = b;
        pl[npe] = l;
        ++npe;
    };
    // clip faces in place, keeping order
    int w = 0;
    const int nf0 = X.nf;
    for (int f = 0; f < nf0; ++f) {
        const int n = X.fn[f];
        bool fp = false, fq = false;
        for (int q = 0; q < n; ++q) {
            fp |= s[X.fv[f][q]] > 0;
            fq |= s[X.fv[f][q]] < 0;
        }
        const I P0 = X.fa[f], N0 = X.fb[f];
        I nv_[C::MFV], ne_[C::MFV];
        int nn = -1;
        I oP = P0, oN = N0;
        if (fq && !fp) {
            nn = n;
            for (int q = 0; q < n; ++q) {
                nv_[q] = X.fv[f][q];
                ne_[q] = X.fe[f][q];
            }
        } else if (fq && fp) {
            I av[2 * C::MFV], ae[2 * C::MFV];
            int mm = 0;
            for (int i = 0; i < n; ++i) {
                I u = X.fv[f][i], ww = X.fv[f][(i + 1) % n];
                av[mm] = u;
                ae[mm] = X.fe[f][i];
                ++mm;
                if (s[u] * s[ww] < 0) {
                    av[mm] = cross_of(u, ww);
                    ae[mm] = X.fe[f][i];
                    ++mm;
                }
            }
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                if (sgn > 0 && P0 >= 4) break;
                int idx[2 * C::MFV], ni = 0;
                for (int i = 0; i < mm; ++i)
                    if (s[av[i]] * sgn >= 0) idx[ni++] = i;
                if (ni > C::MFV) {
                    X.err = true;
                    return;
                }
                const I cutlab = sgn < 0 ? m : N0;
                for (int q = 0; q < ni; ++q) {
                    int ia = idx[q], ib = idx[(q + 1) % ni];
                    I lab = ((ia + 1) % mm == ib) ? ae[ia] : cutlab;
                    if (sgn < 0) {
                        nv_[q] = av[ia];
                        ne_[q] = lab;
                    } else {
                        add_plus(P0, av[ia], av[ib], lab);
                    }
                }
                if (sgn < 0) nn = ni;
            }
        } else if (fp && !fq) {
            if (P0 < 4)
                for (int i = 0; i < n; ++i) add_plus(P0, X.fv[f][i], X.fv[f][(i + 1) % n], X.fe[f][i]);
        } else {
            const bool pc = P0 >= 4 && consumed[P0];
            const bool nc = consumed[N0];
            if (P0 < 4) {
                if (nc) {
                    for (int i = 0; i < n; ++i) add_plus(P0, X.fv[f][i], X.fv[f][(i + 1) % n], X.fe[f][i]);
                } else {
                    nn = n;
                    for (int q = 0; q < n; ++q) {
                        nv_[q] = X.fv[f][q];
                        ne_[q] = X.fe[f][q];
                    }
                }
            } else if (pc && nc) {
            } else if (nc) {
                // m becomes the positive side: reverse the cycle
                nn = n;
                for (int q = 0; q < n; ++q) {
                    nv_[q] = X.fv[f][(n - q) % n];
                    ne_[q] = X.fe[f][(2 * n - 1 - q) % n];
                }
                oP = m;
                oN = P0;
            } else {
                nn = n;
                for (int q = 0; q < n; ++q) {
                    nv_[q] = X.fv[f][q];
                    ne_[q] = X.fe[f][q];
                }
                if (pc) oP = m;
            }
        }
        if (X.err) return;
        if (nn >= 0) {
            X.fa[w] = oP;
            X.fb[w] = oN;
            X.fc[w] = C::NIL;
            X.fn[w] = (I)nn;
            for (int q = 0; q < nn; ++q) {
                X.fv[w][q] = nv_[q];
                X.fe[w][q] = ne_[q];
            }
            ++w;
        }
    }
    X.nf = w;
    // one merged face per tet face won by m
    for (int i = 0; i < 4; ++i) {
        I ea[C::MFV * 2], eb[C::MFV * 2], el[C::MFV * 2];
        int ne = 0, any = 0;
        for (int q = 0; q < npe; ++q) {
            if (pp[q] != i) continue;
            ++any;
            bool shared = false;
            for (int r = 0; r < npe; ++r)
                if (pp[r] == i && pa[r] == pb[q] && pb[r] == pa[q]) {
                    shared = true;
                    break;
                }
            if (shared) continue;
            if (ne >= 2 * C::MFV) {
                X.err = true;
                return;
            }
            ea[ne] = pa[q];
            eb[ne] = pb[q];
            el[ne] = pl[q];
            ++ne;
        }
        if (!any) continue;
        I vs[C::MFV * 2], es[C::MFV * 2];
        int n = chain_edges<2 * C::MFV>(ea, eb, el, ne, vs, es);
        if (n < 0 || X.nf >= C::MF) {
            X.err = true;
            return;
        }
        const int F = X.nf++;
        int cnt = 0;
        for (int q = 0; q < n; ++q) {
            int prev = (q + n - 1) % n;
            if (vs[q] >= 4 && es[prev] == es[q]) continue;
            if (cnt >= C::MFV) {
                X.err = true;
                return;
            }
            X.fv[F][cnt] = vs[q];
            X.fe[F][cnt] = es[q];
            ++cnt;
        }
        X.fa[F] = i;
        X.fb[F] = m;
        X.fc[F] = C::NIL;
        X.fn[F] = (I)cnt;
    }
    // one face (m | X) per cell crossed by f_m = f_X
    for (int c = 0; c < X.nc; ++c) {
        const I L = X.cells[c];
        if (!split[L]) continue;
        I ea[C::MFV], eb[C::MFV], el[C::MFV];
        int ne = 0;
        for (int g = 0; g < X.nf; ++g) {
            if (X.fa[g] != L && X.fb[g] != L) continue;
            const int n = X.fn[g];
            const I other = (X.fa[g] == L) ? X.fb[g] : X.fa[g];
            for (int i = 0; i < n; ++i) {
                I a = X.fv[g][i], b = X.fv[g][(i + 1) % n];
                if (s[a] != 0 || s[b] != 0) continue;
                if (ne >= C::MFV) {
                    X.err = true;
                    return;
                }
                if (X.fb[g] == L) {
                    ea[ne] = b;
                    eb[ne] = a;
                } else {
                    ea[ne] = a;
                    eb[ne] = b;
                }
                el[ne] = other;
                ++ne;
            }
        }
        if (X.nf >= C::MF) {
            X.err = true;
            return;
        }
        const int F = X.nf;
        int n = chain_edges<C::MFV>(ea, eb, el, ne, X.fv[F], X.fe[F]);
        if (n < 0) {
            X.err = true;
            return;
        }
        X.fa[F] = m;
        X.fb[F] = L;
        X.fc[F] = C::NIL;
        X.fn[F] = (I)n;
        ++X.nf;
    }
    // cells: survivors in order, then m
    int nc = 0;
    for (int c = 0; c < X.nc; ++c)
        if (!consumed[X.cells[c]]) X.cells[nc++] = X.cells[c];
    X.cells[nc++] = m;
    X.nc = nc;
    // corners won by m change label; unreferenced vertices disappear
    for (int c = 0; c < 4; ++c)
        if (s[c] > 0) {
            int n = 0;
            for (int q = 0; q < 4; ++q)
                if (q != c) X.vl[c][n++] = q;
            X.vl[c][3] = m;
        }
    I* remap = T.remap;
    bool* used = T.used;
    for (int v = 0; v < X.nv; ++v) used[v] = v < 4;
    for (int f = 0; f < X.nf; ++f)
        for (int q = 0; q < X.fn[f]; ++q) used[X.fv[f][q]] = true;
    int nvn = 0;
    for (int v = 0; v < X.nv; ++v) {
        if (!used[v]) {
            remap[v] = C::NIL;
            continue;
        }
        remap[v] = (I)nvn;
        if (nvn != v)
            for (int q = 0; q < 4; ++q) X.vl[nvn][q] = X.vl[v][q];
        ++nvn;
    }
    X.nv = nvn;
    for (int f = 0; f < X.nf; ++f)
        for (int q = 0; q < X.fn[f]; ++q) X.fv[f][q] = remap[X.fv[f][q]];
}

template <class C>
__host__ __device__ void mi_cut(const double (*P)[4], int k, Complex<C>& X, InsertTemps<C>& T) {
    mi_init(X, k);
    if (X.err || k < 1) return;
    for (int j = 1; j < k; ++j) {
        mi_classify(P, j, X, T.s);
        if (X.err) return;
        mi_insert(X, (typename C::Idx)(4 + j), T.s, T);
        if (X.err) return;
    }
}

template <class C>
void mi_insert(Complex<C>& X, typename C::Idx m, int8_t* s) {
    static thread_local InsertTemps<C> T;
    mi_insert(X, m, s, T);
}
template <class C>
void mi_cut(const double (*P)[4], int k, Complex<C>& X) {
    static thread_local InsertTemps<C> T;
    mi_cut(P, k, X, T);
}

template <class C>
__host__ __device__ int write_record_mi(const Complex<C>& X, int k, uint16_t* out, int cap) {
    int p = 8;
    const int L = 4 + k;
    auto put = [&](uint16_t w) {
        if (p < cap) out[p] = w;
        ++p;
    };
    for (int v = 0; v < X.nv; ++v)
        for (int q = 0; q < 4; ++q) put(X.vl[v][q]);
    for (int f = 0; f < X.nf; ++f) {
        put(X.fa[f]);
        put(X.fb[f]);
        put(0xFFFF);
        put(X.fn[f]);
        for (int q = 0; q < X.fn[f]; ++q) put(X.fv[f][q]);
    }
    for (int c = 0; c < X.nc; ++c) put(X.cells[c]);
    if (X.dup)
        for (int l = 0; l < L; ++l) put(X.grp[l]);
    if (p > cap) return -1;
    out[0] = 2;
    out[1] = (uint16_t)L;
    out[2] = (uint16_t)X.nv;
    out[3] = (uint16_t)X.nf;
    out[4] = (uint16_t)X.nc;
    out[5] = X.dup ? 1 : 0;
    out[6] = (uint16_t)p;
    out[7] = 0;
    return p;
}

}  // namespace dev
}  // namespace rin
