// Wave-cooperative material-interface cut: a group of G lanes builds one
// tet's complex in LDS.  Every insertion runs the steps of mi_insert
// (cut_dev.h) and produces the same complex, numbering included; the work of
// each step is spread over the lanes:
//   * vertex classification (the exact predicates): one vertex per lane;
//   * per-label sign flags: one face per lane, LDS atomics;
//   * crossing vertices: strict sign-change edges listed in (face, edge)
//     order by a prefix sum, the first occurrence of each edge creating its
//     vertex (an LDS hash finds it);
//   * face clipping: each face's output (kept side, reversed copy or none)
//     and its tet-face edges won by the new material, placed in face order by
//     prefix sums;
//   * the merged face of each tet face won by the new material: one lane per
//     tet face chains its edges;
//   * the new face of each split cell: lanes file the cell's zero edges with
//     their (face, position) order, one lane per split cell chains them;
//   * vertex compaction: reference flags by face lanes, ranks by a prefix sum.
// Blocks are single wavefronts, so __syncthreads() is a wave barrier.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "cut_wave.h"

namespace rin {
namespace dev {

template <class C>
struct WaveMI {
    static constexpr int MV = C::MV, MF = C::MF, MFV = C::MFV, MP = C::MP;
    static constexpr int CAND = 2 * C::MF;
    static constexpr int HT = CAND <= 64 ? 128 : (CAND <= 128 ? 256 : (CAND <= 512 ? 1024 : 2048));
    static constexpr int PE = 4 * MFV * 4;  // tet-face edges won by the new material
    static constexpr int TE = 2 * MFV;      // edges of one merged tet face
    double P[MP - 4][4];
    uint8_t vl[MV][4];
    int8_t s[MV];
    uint8_t fa[MF], fb[MF], fn[MF];
    alignas(8) uint8_t fv[MF][MFV];
    alignas(8) uint8_t fe[MF][MFV];
    uint8_t cells[MP], grp[MP];
    uint32_t lflag[MP];  // per label: bit 0 a positive vertex on its faces, bit 1 a negative one
    uint8_t fflag[MF];   // per face: bit 0 a vertex with s > 0, bit 1 s < 0
    uint16_t fcand[MF];  // first crossing candidate of each face
    uint16_t fout[MF];   // the clipped face's index, or 0xFFFF
    uint16_t fpe[MF];    // the face's first tet-face edge
    uint8_t used[MV], remap[MV];
    uint32_t tface[4];   // merged tet faces: bit 0 when the tet face has edges won by m
    // working arrays of stages that never overlap
    struct Cand {
        uint32_t ht[HT];
        uint16_t ckey[CAND];
        uint8_t cedge[CAND], cfirst[CAND], cvid[CAND];
        uint16_t cface[CAND];
    };
    struct Clip {
        uint8_t fa2[MF], fb2[MF], fn2[MF];
        alignas(8) uint8_t fv2[MF][MFV];
        alignas(8) uint8_t fe2[MF][MFV];
    };
    struct Split {
        uint8_t ea[4][TE], eb[4][TE], el[4][TE];  // merged tet faces' edges
        uint8_t tv[MP][MFV], tb[MP][MFV], te[MP][MFV];
        uint16_t tk[MP][MFV];  // collection order of a split cell's edges
        uint32_t ccnt[MP];
        uint8_t crow[MP];      // row of each split label, or 0xFF
    };
    union {
        Cand c;
        Split e;
    } u;
    Clip k2;  // the clip's output faces (live while the candidates are)
    uint8_t pp[PE], pa[PE], pb[PE], pl[PE];
    uint8_t pshr[PE];  // the edge's reverse is also there (left out of the merged face)
    int nv, nf, nc, dup, err;
};

// vertex signs of the insertion under way (bit masks when ids fit 64 bits)
template <class C>
struct MISigns {
    static constexpr bool kMask = C::MV <= 64;
    uint64_t pos = 0, neg = 0;
    __device__ int8_t get(const WaveMI<C>& X, uint8_t v) const {
        if constexpr (kMask) {
            const unsigned b = v & 63u;
            return (int8_t)((int)((pos >> b) & 1u) - (int)((neg >> b) & 1u));
        } else {
            return X.s[v];
        }
    }
};

template <class C>
__device__ inline int mi_load_face(const WaveMI<C>& X, const MISigns<C>& S, int f, uint8_t (&fvr)[C::MFV],
                                   int8_t (&sr)[C::MFV]) {
    const int n = X.fn[f];
    load_row<C::MFV>(X.fv[f], fvr);
#pragma unroll
    for (int i = 0; i < C::MFV; ++i) sr[i] = i < n ? S.get(X, fvr[i]) : (int8_t)0;
    return n;
}

// chains the edges (ea -> eb, label el; ne of them, in collection order) into
// one cycle from edge 0, always to the first edge in collection order that
// continues it (chain_edges of cut_dev.h); returns the length or -1
template <int CAP>
__device__ inline int wave_chain(const uint8_t* ea, const uint8_t* eb, const uint8_t* el, int ne, uint8_t* ov,
                                 uint8_t* oe) {
    if (ne <= 0 || ne > CAP) return -1;
    int cur = 0, cnt = 0;
    for (;;) {
        ov[cnt] = ea[cur];
        oe[cnt] = el[cur];
        ++cnt;
        int nxt = -1;
        const uint8_t b = eb[cur];
        for (int q = 0; q < ne; ++q)
            if (ea[q] == b) {
                nxt = q;
                break;
            }
        if (nxt < 0 || cnt > ne) return -1;
        cur = nxt;
        if (cur == 0) break;
    }
    return cnt == ne ? cnt : -1;
}

// Runs the insertion of materials 1..k-1 (X.P[0..k-1] loaded by the caller)
// into the tet filled with material 0 (label 4); on return X.err is set when
// a capacity or the predicates' exponent range was exceeded.
// start: materials already in the complex (1, or 3 when the caller has
// loaded the complex after the first three, which depends on their signs only)
template <class C, int G = 64>
__device__ __forceinline__ void wave_mi_cut(WaveMI<C>& X, int k, int start = 1) {
    using W = Grp<G>;
    constexpr int MFV = C::MFV, MV = C::MV, MF = C::MF, MP = C::MP;
    constexpr int CAND = WaveMI<C>::CAND, HT = WaveMI<C>::HT, PE = WaveMI<C>::PE, TE = WaveMI<C>::TE;
    const int lane = W::lane();
    // the tet filled with material 4 (mi_init)
    if (start == 1 && lane < 4) {
        const int i = lane;
        int n = 0;
        for (int q = 0; q < 4; ++q)
            if (q != i) X.vl[i][n++] = (uint8_t)q;
        X.vl[i][3] = 4;
        const uint8_t cyc[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
        X.fa[i] = (uint8_t)i;
        X.fb[i] = 4;
        X.fn[i] = 3;
        for (int q = 0; q < 3; ++q) {
            X.fv[i][q] = cyc[i][q];
            X.fe[i][q] = cyc[i][(q + 2) % 3];
        }
    }
    if (start == 1 && lane < MP) X.grp[lane] = lane < 4 ? (uint8_t)lane : (lane == 4 ? (uint8_t)4 : NIL);
    if (start == 1 && lane == 0) {
        X.nv = 4;
        X.nf = 4;
        X.nc = 1;
        X.cells[0] = 4;
        X.dup = 0;
        X.err = (4 + k > MP || k < 1) ? 1 : 0;
    }
    __syncthreads();
    if (X.err) return;
    auto fail = [&]() {
        if (lane == 0) X.err = 1;
        __syncthreads();
    };
    for (int j = start; j < k; ++j) {
        const uint8_t m = (uint8_t)(4 + j);
        const int nv0 = X.nv, nf0 = X.nf, nc0 = X.nc;

        // classify every vertex: > 0 where material j wins
        bool e = false, ap = false;
        MISigns<C> S;
        if constexpr (MISigns<C>::kMask) {
            uint64_t pm = 0, nm = 0;
            for (int r0 = 0; r0 < nv0; r0 += G) {
                const int v = r0 + lane;
                int sv = 0;
                if (v < nv0) {
                    const int lab[4] = {X.vl[v][0], X.vl[v][1], X.vl[v][2], X.vl[v][3]};
                    sv = orient_mi(X.P, lab, X.P[j], e);
                    X.s[v] = (int8_t)sv;
                }
                pm |= W::ballot(sv > 0) << r0;
                nm |= W::ballot(sv < 0) << r0;
            }
            S.pos = pm;
            S.neg = nm;
            ap = pm != 0;
        } else {
            for (int v = lane; v < nv0; v += G) {
                const int lab[4] = {X.vl[v][0], X.vl[v][1], X.vl[v][2], X.vl[v][3]};
                const int sv = orient_mi(X.P, lab, X.P[j], e);
                X.s[v] = (int8_t)sv;
                ap |= sv > 0;
            }
        }
        if (lane < MP) X.lflag[lane] = 0;
        const bool any_err = W::any(e), any_pos = W::any(ap);
        if (any_err) return fail();
        __syncthreads();

        // each face's sign flags, and per label the signs met on its faces
        for (int f = lane; f < nf0; f += G) {
            uint8_t fvr[MFV];
            int8_t sr[MFV];
            (void)mi_load_face<C>(X, S, f, fvr, sr);
            const uint8_t fl = sign_flags<MFV>(sr);
            X.fflag[f] = fl;
            if (fl) {
                if (X.fa[f] >= 4) atomicOr(&X.lflag[X.fa[f]], (uint32_t)fl);
                if (X.fb[f] >= 4) atomicOr(&X.lflag[X.fb[f]], (uint32_t)fl);
            }
        }
        __syncthreads();
        if (!any_pos) {
            // m wins nowhere: it joins the group of a cell it ties everywhere
            if (lane == 0) {
                X.grp[m] = m;
                for (int c = 0; c < nc0; ++c) {
                    const uint8_t L = X.cells[c];
                    if (X.lflag[L] == 0) {
                        X.grp[m] = X.grp[L];
                        X.dup = 1;
                        break;
                    }
                }
            }
            __syncthreads();
            continue;
        }
        if (lane == 0) X.grp[m] = m;

        // crossing vertices: strict sign-change edges in (face, edge) order,
        // the first occurrence of each edge creating its vertex
        for (int h = lane; h < HT; h += G) X.u.c.ht[h] = 0xFFFFFFFFu;
        uint32_t ncand = 0;
        for (int base = 0; base < nf0; base += G) {
            const int f = base + lane;
            uint8_t fvr[MFV];
            int8_t sr[MFV];
            int n = 0;
            uint32_t c = 0;
            const bool mixed = f < nf0 && X.fflag[f] == 3;
            if (mixed) {
                n = mi_load_face<C>(X, S, f, fvr, sr);
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    const int sw = cyc_next<MFV>(sr, i, n);
                    c += i < n && sr[i] * sw < 0;
                }
            }
            static_assert(MFV < 16, "crossings per face fit four bits");
            uint32_t o = ncand + W::template excl<4>(c);
            if (mixed) {
                X.fcand[f] = (uint16_t)o;
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    if (i >= n) continue;
                    const int sw = cyc_next<MFV>(sr, i, n);
                    const uint8_t vw = cyc_next<MFV>(fvr, i, n);
                    if (sr[i] * sw >= 0) continue;
                    if (o < (uint32_t)CAND) {
                        const uint8_t uu = fvr[i];
                        X.u.c.ckey[o] = (uint16_t)((uu < vw ? uu : vw) << 8 | (uu < vw ? vw : uu));
                        X.u.c.cface[o] = (uint16_t)f;
                        X.u.c.cedge[o] = (uint8_t)i;
                    }
                    ++o;
                }
            }
            ncand += W::template sum<4>(c);
        }
        if (ncand > (uint32_t)CAND) return fail();
        __syncthreads();
        auto home = [&](uint32_t key) { return (key * 2654435761u) >> 16 & (uint32_t)(HT - 1); };
        auto first_of = [&](uint32_t key) {
            uint32_t h = home(key);
            for (int probe = 0; probe < HT && X.u.c.ht[h] >> 16 != key; ++probe) h = (h + 1) & (uint32_t)(HT - 1);
            return X.u.c.ht[h] & 0xFFFFu;
        };
        for (int c = lane; c < (int)ncand; c += G) {
            const uint32_t key = X.u.c.ckey[c], packed = key << 16 | (uint32_t)c;
            uint32_t h = home(key);
            for (int probe = 0; probe < HT; ++probe) {
                const uint32_t cur = atomicCAS(&X.u.c.ht[h], 0xFFFFFFFFu, packed);
                if (cur == 0xFFFFFFFFu) break;
                if (cur >> 16 == key) {
                    atomicMin(&X.u.c.ht[h], packed);
                    break;
                }
                h = (h + 1) & (uint32_t)(HT - 1);
            }
        }
        __syncthreads();
        for (int c = lane; c < (int)ncand; c += G) X.u.c.cfirst[c] = first_of(X.u.c.ckey[c]) == (uint32_t)c;
        __syncthreads();
        const uint32_t nnew = wave_prefix<G, 1>(
            (int)ncand, [&](int c) { return (uint32_t)X.u.c.cfirst[c]; },
            [&](int c, uint32_t o) {
                if (X.u.c.cfirst[c]) X.u.c.cvid[c] = (uint8_t)(nv0 + o);
            });
        if (nv0 + (int)nnew > MV) return fail();
        __syncthreads();
        for (int c = lane; c < (int)ncand; c += G) {
            if (X.u.c.cfirst[c]) {
                const int f = X.u.c.cface[c];
                uint8_t t[4] = {X.fa[f], X.fb[f], X.fe[f][X.u.c.cedge[c]], m};
                sort4(t);
                const int id = X.u.c.cvid[c];
                for (int q = 0; q < 4; ++q) X.vl[id][q] = t[q];
                X.s[id] = 0;
            } else {
                X.u.c.cvid[c] = X.u.c.cvid[first_of(X.u.c.ckey[c])];
            }
        }
        __syncthreads();

        // clip the faces, keeping their order: each face's output and its
        // tet-face edges won by m, placed by prefix sums
        auto consumed = [&](uint8_t L) { return X.lflag[L] == 1u; };
        // per face: emits (0 / 1) and the number of tet-face edges it adds
        auto plan = [&](int f, uint32_t& emit, uint32_t& npe) {
            const uint8_t fl = X.fflag[f];
            const bool fp = fl & 1, fq = fl & 2;
            const uint8_t P0 = X.fa[f], N0 = X.fb[f];
            const int n = X.fn[f];
            emit = 0;
            npe = 0;
            if (fq && !fp) {
                emit = 1;
            } else if (fq && fp) {
                emit = 1;
                if (P0 < 4) {
                    // the positive side: its kept positions, crossings included
                    uint8_t fvr[MFV];
                    int8_t sr[MFV];
                    (void)mi_load_face<C>(X, S, f, fvr, sr);
#pragma unroll
                    for (int i = 0; i < MFV; ++i) {
                        if (i >= n) continue;
                        const int sw = cyc_next<MFV>(sr, i, n);
                        npe += sr[i] >= 0;
                        npe += sr[i] * sw < 0;  // the crossing vertex (sign 0)
                    }
                }
            } else if (fp && !fq) {
                if (P0 < 4) npe = (uint32_t)n;
            } else {
                const bool pc = P0 >= 4 && consumed(P0);
                const bool nc = consumed(N0);
                if (P0 < 4) {
                    if (nc)
                        npe = (uint32_t)n;
                    else
                        emit = 1;
                } else if (!(pc && nc)) {
                    emit = 1;
                }
            }
        };
        uint32_t nout = 0, npe_all = 0;
        for (int base = 0; base < nf0; base += G) {
            const int f = base + lane;
            uint32_t emit = 0, npe = 0;
            if (f < nf0) plan(f, emit, npe);
            const uint32_t oo = nout + W::template excl<1>(emit), op = npe_all + W::template excl<4>(npe);
            if (f < nf0) {
                X.fout[f] = emit ? (uint16_t)oo : (uint16_t)0xFFFF;
                X.fpe[f] = (uint16_t)op;
            }
            nout += W::template sum<1>(emit);
            npe_all += W::template sum<4>(npe);
        }
        if (npe_all > (uint32_t)PE) return fail();
        __syncthreads();
        bool ferr = false;
        for (int f = lane; f < nf0; f += G) {
            const uint8_t fl = X.fflag[f];
            const bool fp = fl & 1, fq = fl & 2;
            const uint8_t P0 = X.fa[f], N0 = X.fb[f];
            uint8_t fvr[MFV], fer[MFV];
            int8_t sr[MFV];
            const int n = mi_load_face<C>(X, S, f, fvr, sr);
            load_row<MFV>(X.fe[f], fer);
            const int w = X.fout[f];
            const int pe0 = X.fpe[f];
            auto keep = [&](uint8_t oP, uint8_t oN) {
                X.k2.fa2[w] = oP;
                X.k2.fb2[w] = oN;
                X.k2.fn2[w] = (uint8_t)n;
#pragma unroll
                for (int q = 0; q < MFV; ++q) {
                    X.k2.fv2[w][q] = fvr[q];
                    X.k2.fe2[w][q] = fer[q];
                }
            };
            auto plus_all = [&]() {
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    if (i >= n) continue;
                    X.pp[pe0 + i] = P0;
                    X.pa[pe0 + i] = fvr[i];
                    X.pb[pe0 + i] = cyc_next<MFV>(fvr, i, n);
                    X.pl[pe0 + i] = fer[i];
                }
            };
            if (fq && !fp) {
                keep(P0, N0);
            } else if (fq && fp) {
                uint8_t cr[MFV];
                uint32_t ci = X.fcand[f];
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    cr[i] = 0;
                    if (i < n) {
                        const int sw = cyc_next<MFV>(sr, i, n);
                        if (sr[i] * sw < 0) cr[i] = X.u.c.cvid[ci++];
                    }
                }
                // the negative side stays, its cut edge labelled m
                const int nn = split_side_regs<MFV>(fvr, fer, sr, cr, n, -1, m, X.k2.fv2[w], X.k2.fe2[w]);
                if (nn < 0) {
                    ferr = true;
                    continue;
                }
                X.k2.fa2[w] = P0;
                X.k2.fb2[w] = N0;
                X.k2.fn2[w] = (uint8_t)nn;
                if (P0 < 4) {
                    // the positive side's edges go to the tet face, its cut
                    // edge labelled N0
                    const int np = split_side_regs<MFV>(fvr, fer, sr, cr, n, +1, N0, X.pa + pe0, X.pl + pe0);
                    if (np < 0) {
                        ferr = true;
                        continue;
                    }
                    for (int q = 0; q < np; ++q) {
                        X.pp[pe0 + q] = P0;
                        X.pb[pe0 + q] = X.pa[pe0 + (q + 1 == np ? 0 : q + 1)];
                    }
                }
            } else if (fp && !fq) {
                if (P0 < 4) plus_all();
            } else {
                const bool pc = P0 >= 4 && consumed(P0);
                const bool nc = consumed(N0);
                if (P0 < 4) {
                    if (nc)
                        plus_all();
                    else
                        keep(P0, N0);
                } else if (pc && nc) {
                } else if (nc) {
                    // m becomes the positive side: the cycle reversed
                    X.k2.fa2[w] = m;
                    X.k2.fb2[w] = P0;
                    X.k2.fn2[w] = (uint8_t)n;
#pragma unroll
                    for (int q = 0; q < MFV; ++q) {
                        if (q >= n) continue;
                        // vertex (n - q) % n, edge label (2n - 1 - q) % n
                        const int iv = q == 0 ? 0 : n - q;
                        const int ie = n - 1 - q;
                        uint8_t a = 0, l = 0;
#pragma unroll
                        for (int r = 0; r < MFV; ++r) {
                            a = r == iv ? fvr[r] : a;
                            l = r == ie ? fer[r] : l;
                        }
                        X.k2.fv2[w][q] = a;
                        X.k2.fe2[w][q] = l;
                    }
                } else {
                    keep(pc ? m : P0, N0);
                }
            }
        }
        if (W::any(ferr)) return fail();
        __syncthreads();
        // the clipped faces replace the old ones
        for (int w = lane; w < (int)nout; w += G) {
            X.fa[w] = X.k2.fa2[w];
            X.fb[w] = X.k2.fb2[w];
            X.fn[w] = X.k2.fn2[w];
            uint64_t rv, re;
            __builtin_memcpy(&rv, X.k2.fv2[w], 8 * (MFV / 8));
            if constexpr (MFV == 8) {
                __builtin_memcpy(&re, X.k2.fe2[w], 8);
                __builtin_memcpy(X.fv[w], &rv, 8);
                __builtin_memcpy(X.fe[w], &re, 8);
            } else {
                for (int q = 0; q < MFV; ++q) {
                    X.fv[w][q] = X.k2.fv2[w][q];
                    X.fe[w][q] = X.k2.fe2[w][q];
                }
            }
        }
        __syncthreads();

        // one merged face per tet face won by m: its edges in order, less the
        // edges whose reverse is also there, chained
        const int npe = (int)npe_all;
        if (lane < 4) X.tface[lane] = 0;
        __syncthreads();
        for (int q = lane; q < npe; q += G) {
            const uint8_t pq = X.pp[q], aq = X.pa[q], bq = X.pb[q];
            atomicOr(&X.tface[pq], 1u);
            bool shared = false;
            for (int r = 0; r < npe && !shared; ++r) shared = X.pp[r] == pq && X.pa[r] == bq && X.pb[r] == aq;
            X.pshr[q] = shared ? 1 : 0;
        }
        __syncthreads();
        // face indices of the merged faces in tet-face order
        int nf1 = (int)nout;
        {
            int add = 0;
            for (int i = 0; i < 4; ++i) add += (int)(X.tface[i] & 1u);
            if (nf1 + add > MF) return fail();
        }
        bool terr = false;
        if (lane < 4 && (X.tface[lane] & 1u)) {
            const int i = lane;
            int F = nf1;
            for (int q = 0; q < i; ++q) F += (int)(X.tface[q] & 1u);
            uint8_t* ea = X.u.e.ea[i];
            uint8_t* eb = X.u.e.eb[i];
            uint8_t* el = X.u.e.el[i];
            int ne = 0;
            for (int q = 0; q < npe && !terr; ++q) {
                if (X.pp[q] != i || X.pshr[q]) continue;
                if (ne >= TE) {
                    terr = true;
                    break;
                }
                ea[ne] = X.pa[q];
                eb[ne] = X.pb[q];
                el[ne] = X.pl[q];
                ++ne;
            }
            uint8_t vs[TE], es[TE];
            const int n = terr ? -1 : wave_chain<TE>(ea, eb, el, ne, vs, es);
            if (n < 0) {
                terr = true;
            } else {
                int cnt = 0;
                for (int q = 0; q < n; ++q) {
                    const int prev = (q + n - 1) % n;
                    if (vs[q] >= 4 && es[prev] == es[q]) continue;
                    if (cnt >= MFV) {
                        terr = true;
                        break;
                    }
                    X.fv[F][cnt] = vs[q];
                    X.fe[F][cnt] = es[q];
                    ++cnt;
                }
                X.fa[F] = (uint8_t)i;
                X.fb[F] = m;
                X.fn[F] = (uint8_t)cnt;
            }
        }
        if (W::any(terr)) return fail();
        for (int i = 0; i < 4; ++i) nf1 += (int)(X.tface[i] & 1u);
        __syncthreads();

        // one face (m | L) per split cell L, in cell order: lanes file the
        // cell's edges between zero vertices with their (face, position)
        // order, a lane per split cell chains them
        if (lane == 0) {
            for (int L = 0; L < MP; ++L) X.u.e.crow[L] = 0xFF;  // m and the tet faces: no row
            int r = 0;
            for (int c = 0; c < nc0; ++c) {
                const uint8_t L = X.cells[c];
                X.u.e.crow[L] = X.lflag[L] == 3u ? (uint8_t)r++ : (uint8_t)0xFF;
            }
            X.u.e.ccnt[MP - 1] = (uint32_t)r;  // the split count, read below
        }
        __syncthreads();
        const int nsplit = (int)X.u.e.ccnt[MP - 1];
        if (nf1 + nsplit > MF) return fail();
        __syncthreads();
        for (int r = lane; r < MP; r += G) X.u.e.ccnt[r] = 0;
        __syncthreads();
        bool cerr = false;
        for (int g = lane; g < nf1; g += G) {
            const uint8_t A = X.fa[g], B = X.fb[g];
            const int ra = (A >= 4 && A < MP) ? (int)X.u.e.crow[A] : 0xFF;
            const int rb = (B >= 4 && B < MP) ? (int)X.u.e.crow[B] : 0xFF;
            if (ra == 0xFF && rb == 0xFF) continue;
            uint8_t fvr[MFV];
            int8_t sr[MFV];
            const int n = mi_load_face<C>(X, S, g, fvr, sr);
#pragma unroll
            for (int i = 0; i < MFV; ++i) {
                if (i >= n) continue;
                const uint8_t a = fvr[i], b = cyc_next<MFV>(fvr, i, n);
                if (sr[i] != 0 || cyc_next<MFV>(sr, i, n) != 0) continue;
                for (int side = 0; side < 2; ++side) {
                    // side 0: the face has L as its positive label (a -> b),
                    // side 1: as its negative label (b -> a)
                    const int row = side ? rb : ra;
                    if (row == 0xFF) continue;
                    const uint32_t slot = atomicAdd(&X.u.e.ccnt[row], 1u);
                    if (slot >= (uint32_t)MFV) {
                        cerr = true;
                        continue;
                    }
                    X.u.e.tv[row][slot] = side ? b : a;
                    X.u.e.tb[row][slot] = side ? a : b;
                    X.u.e.te[row][slot] = side ? A : B;
                    X.u.e.tk[row][slot] = (uint16_t)(g << 4 | i);
                }
            }
        }
        if (W::any(cerr)) return fail();
        __syncthreads();
        for (int c = lane; c < nc0; c += G) {
            const uint8_t L = X.cells[c];
            const int row = X.u.e.crow[L];
            if (row == 0xFF) continue;
            const int ne = (int)X.u.e.ccnt[row];
            // the row in collection order (a few entries: insertion sort)
            uint8_t ea[MFV], eb[MFV], el[MFV];
            uint16_t ek[MFV];
            for (int q = 0; q < ne; ++q) {
                uint16_t kq = X.u.e.tk[row][q];
                uint8_t aq = X.u.e.tv[row][q], bq = X.u.e.tb[row][q], lq = X.u.e.te[row][q];
                int p = q;
                while (p > 0 && ek[p - 1] > kq) {
                    ek[p] = ek[p - 1];
                    ea[p] = ea[p - 1];
                    eb[p] = eb[p - 1];
                    el[p] = el[p - 1];
                    --p;
                }
                ek[p] = kq;
                ea[p] = aq;
                eb[p] = bq;
                el[p] = lq;
            }
            const int F = nf1 + row;
            const int n = wave_chain<MFV>(ea, eb, el, ne, X.fv[F], X.fe[F]);
            if (n < 0) {
                cerr = true;
                continue;
            }
            X.fa[F] = m;
            X.fb[F] = L;
            X.fn[F] = (uint8_t)n;
        }
        if (W::any(cerr)) return fail();
        const int nf2 = nf1 + nsplit;
        const int nv1 = nv0 + (int)nnew;
        __syncthreads();

        // cells: survivors in order, then m; corners won by m change label
        if (lane == 0) {
            int nc = 0;
            for (int c = 0; c < nc0; ++c)
                if (!consumed(X.cells[c])) X.cells[nc++] = X.cells[c];
            X.cells[nc++] = m;
            X.nc = nc;
        }
        if (lane < 4 && S.get(X, (uint8_t)lane) > 0) {
            int n = 0;
            for (int q = 0; q < 4; ++q)
                if (q != lane) X.vl[lane][n++] = (uint8_t)q;
            X.vl[lane][3] = m;
        }
        // unreferenced vertices disappear, the others keep their order
        for (int v = lane; v < nv1; v += G) X.used[v] = v < 4 ? 1 : 0;
        __syncthreads();
        for (int f = lane; f < nf2; f += G) {
            const int n = X.fn[f];
            for (int q = 0; q < n; ++q) X.used[X.fv[f][q]] = 1;
        }
        __syncthreads();
        const uint32_t nvn = wave_prefix<G, 1>(
            nv1, [&](int v) { return (uint32_t)X.used[v]; },
            [&](int v, uint32_t o) { X.remap[v] = X.used[v] ? (uint8_t)o : (uint8_t)NIL; });
        __syncthreads();
        // move the vertex labels down: read every row first, then write
        constexpr int R = (MV + G - 1) / G;
        uint32_t row[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int v = r * G + lane;
            row[r] = 0;
            if (v < nv1) __builtin_memcpy(&row[r], X.vl[v], 4);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int v = r * G + lane;
            if (v < nv1 && X.used[v] && X.remap[v] != v) __builtin_memcpy(X.vl[X.remap[v]], &row[r], 4);
        }
        for (int f = lane; f < nf2; f += G) {
            const int n = X.fn[f];
            for (int q = 0; q < n; ++q) X.fv[f][q] = X.remap[X.fv[f][q]];
        }
        if (lane == 0) {
            X.nv = (int)nvn;
            X.nf = nf2;
        }
        __syncthreads();
    }
}

// Writes the record (include/rin/cut_record.h layout, write_record_mi) with
// the group and returns its length, or -1 when it exceeds cap; niv / nif /
// nfv are the tet's interface-vertex, face and face-corner counts
// (mi_counts), uniform over the group.
template <class C, int G = 64>
__device__ __forceinline__ int wave_mi_record(WaveMI<C>& X, int k, uint16_t* out, int cap, uint32_t& niv, uint32_t& nif,
                              uint32_t& nfv) {
    using W = Grp<G>;
    const int lane = W::lane();
    const int L = 4 + k, nv = X.nv, nf = X.nf, nc = X.nc;
    // face word offsets: fpe is free after the insertions
    static_assert(4 + C::MFV < 16, "face words fit four bits");
    const uint32_t fwords = wave_prefix<G, 4>(
        nf, [&](int f) { return (uint32_t)(4 + X.fn[f]); }, [&](int f, uint32_t o) { X.fpe[f] = (uint16_t)o; });
    const int total = 8 + 4 * nv + (int)fwords + nc + (X.dup ? L : 0);
    for (int v = lane; v < nv; v += G) X.used[v] = 0;
    __syncthreads();
    uint32_t cf = 0, cfv = 0;
    for (int f = lane; f < nf; f += G) {
        if (X.fa[f] <= 3) continue;
        ++cf;
        cfv += X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) X.used[X.fv[f][q]] = 1;
    }
    nif = W::sum(cf);
    nfv = W::sum(cfv);
    __syncthreads();
    uint32_t cv = 0;
    for (int v = lane; v < nv; v += G) cv += X.used[v];
    niv = W::sum(cv);
    if (total > cap) return -1;
    for (int v = lane; v < nv; v += G) {
        uint16_t* o = out + 8 + 4 * v;
        for (int q = 0; q < 4; ++q) o[q] = X.vl[v][q];
    }
    const int fb0 = 8 + 4 * nv;
    for (int f = lane; f < nf; f += G) {
        uint16_t* o = out + fb0 + X.fpe[f];
        o[0] = X.fa[f];
        o[1] = X.fb[f];
        o[2] = 0xFFFF;
        o[3] = X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) o[4 + q] = X.fv[f][q];
    }
    const int cb0 = fb0 + (int)fwords;
    for (int c = lane; c < nc; c += G) out[cb0 + c] = X.cells[c];
    if (X.dup)
        for (int l = lane; l < L; l += G) out[cb0 + nc + l] = X.grp[l];
    if (lane == 0) {
        out[0] = 2;
        out[1] = (uint16_t)L;
        out[2] = (uint16_t)nv;
        out[3] = (uint16_t)nf;
        out[4] = (uint16_t)nc;
        out[5] = X.dup ? 1 : 0;
        out[6] = (uint16_t)total;
        out[7] = 0;
    }
    return total;
}

}  // namespace dev
}  // namespace rin
