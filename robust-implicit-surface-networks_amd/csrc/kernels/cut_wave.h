// Wave-cooperative implicit-arrangement cut: one 64-lane wavefront builds one
// tet's complex, held in LDS.  Every insertion runs the steps of ia_cut
// (cut_dev.h) and produces the same numbering; the work of each step is
// spread over the lanes:
//   * vertex classification (the exact predicates): one vertex per lane;
//   * crossing vertices: the strict sign-change edges are listed in
//     (face, edge) order by a prefix sum over faces (a convex face has at
//     most two), the first occurrence of each edge creates its vertex, and a
//     second prefix sum numbers the new vertices in that order;
//   * face splits: one face per lane, negative parts appended in face order;
//   * cell splits: flags by LDS atomics, split cells numbered in cell order,
//     one new face per split cell chained by one lane.
// Blocks are exactly one wavefront, so __syncthreads() is a wave barrier.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "cut_dev.h"

namespace rin {
namespace dev {

// C: the thread-cut capacities (vertices, faces, vertices per face, labels);
// cells are bounded as in ia_cut (8 per label, at most 254)
template <class C>
struct WaveIA {
    static constexpr int MV = C::MV, MF = C::MF, MFV = C::MFV, MP = C::MP;
    static constexpr int MC = C::MP * 8 < 254 ? C::MP * 8 : 254;
    static constexpr int CAND = 2 * C::MF;
    double P[MP - 4][4];
    uint8_t vl[MV][3];
    int8_t s[MV];
    uint8_t fa[MF], fb[MF], fc[MF], fn[MF];
    alignas(8) uint8_t fv[MF][MFV];  // rows of 8 entries load as one word
    alignas(8) uint8_t fe[MF][MFV];
    uint8_t grp[MP], ori[MP];
    uint8_t fflag[MF];  // bit 0: a vertex with s > 0, bit 1: s < 0
    uint16_t fcand[MF];
    uint16_t frank[MF];
    // candidate edges by key: key << 16 | smallest candidate index
    static constexpr int HT = CAND <= 64 ? 128 : (CAND <= 128 ? 256 : (CAND <= 512 ? 1024 : 2048));
    // rows per split cell (new-face edges): an insertion splits at most the
    // cells already there, and up to 4 planes cut a tet into at most 15
    static constexpr int ROWS = MP <= 8 ? 16 : 64;
    // Working arrays of stages that never overlap share their storage: the
    // crossing edges live from the candidate listing to the face splits, the
    // cell splits after those, and the record pass after the last insertion
    // (each stage starts after a group barrier that ends the one before).
    struct Cand {
        uint32_t ht[HT];
        uint16_t ckey[CAND];
        uint16_t cface[CAND];
        uint8_t cedge[CAND], cfirst[CAND], cvid[CAND];
    };
    struct Cells {
        uint32_t cflag[MC];
        uint16_t crank[MC];
        uint8_t tv[ROWS][MFV], te[ROWS][MFV], tb[ROWS][MFV];
        uint16_t tk[ROWS][MFV];  // collection order of a split cell's edges
        uint32_t ccnt[ROWS];     // edges collected per split cell
    };
    struct Rec {
        uint16_t foff[MF];
        uint8_t isov[MV];
    };
    union {
        Cand c;
        Cells e;
        Rec r;
    } u;
    int nv, nf, nc, dup, err;
};

// Lane-group primitives: a group of G lanes (16, 32 or the whole 64-lane
// wavefront) works on one tet, so a wavefront can carry 64 / G tets.
template <int G>
struct Grp {
    static_assert(G == 16 || G == 32 || G == 64, "group size");
    __device__ static int lane() { return threadIdx.x & (G - 1); }
    __device__ static bool any(bool p) {
        const uint64_t b = __ballot(p);
        if (G == 64) return b != 0;
        return ((b >> ((threadIdx.x & 63) & ~(G - 1))) & ((1ull << (G & 63)) - 1)) != 0;
    }
    // the group's lanes' predicate bits, group lane i at bit i
    __device__ static uint64_t ballot(bool p) {
        const uint64_t b = __ballot(p);
        if (G == 64) return b;
        return (b >> ((threadIdx.x & 63) & ~(G - 1))) & ((1ull << (G & 63)) - 1);
    }
    // exclusive prefix and sum over the group.  B > 0: x < 2^B on every lane,
    // and bit b of the result comes from a ballot of bit b (a mask count, no
    // cross-lane data movement); B = 0: any x, by log2(G) lane shuffles.
    template <int B = 0>
    __device__ static uint32_t excl(uint32_t x) {
        const int l = lane();
        if constexpr (B > 0) {
            const uint64_t below = (1ull << l) - 1u;  // the group's lanes before this one
            uint32_t r = 0;
#pragma unroll
            for (int b = 0; b < B; ++b) r += (uint32_t)__popcll(ballot((x >> b) & 1u) & below) << b;
            return r;
        } else {
            uint32_t v = x;
            for (int d = 1; d < G; d <<= 1) {
                const uint32_t o = __shfl_up(v, d, G);
                if (l >= d) v += o;
            }
            return v - x;
        }
    }
    template <int B = 0>
    __device__ static uint32_t sum(uint32_t x) {
        if constexpr (B > 0) {
            uint32_t r = 0;
#pragma unroll
            for (int b = 0; b < B; ++b) r += (uint32_t)__popcll(ballot((x >> b) & 1u)) << b;
            return r;
        } else {
            for (int d = G / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, G);
            return x;
        }
    }
    __device__ static int min(int x) {
        for (int d = G / 2; d > 0; d >>= 1) {
            const int o = __shfl_xor(x, d, G);
            x = o < x ? o : x;
        }
        return x;
    }
};

// exclusive prefix of count(i) over items 0..n-1, item i on group lane
// i % G; store(i, offset) runs on that same lane; returns the total (uniform
// over the group).  B > 0: every count is below 2^B.
template <int G, int B = 0, class F, class H>
__device__ inline uint32_t wave_prefix(int n, F count, H store) {
    uint32_t base = 0;
    for (int r0 = 0; r0 < n; r0 += G) {
        const int i = r0 + Grp<G>::lane();
        const uint32_t c = i < n ? count(i) : 0u;
        const uint32_t pre = Grp<G>::template excl<B>(c);
        if (i < n) store(i, base + pre);
        base += Grp<G>::template sum<B>(c);
    }
    return base;
}

// The vertex signs of the insertion under way.  When vertex ids fit in 64
// bits, every lane of the group holds them as two bit masks (positive,
// negative) built by ballots after the classification, so a face's signs
// need no LDS reads; otherwise they are read from X.s.
template <class C>
struct SignView {
    static constexpr bool kMask = C::MV <= 64;
    uint64_t pos = 0, neg = 0;
    __device__ int8_t get(const WaveIA<C>& X, uint8_t v) const {
        if constexpr (kMask) {
            const unsigned b = v & 63u;
            return (int8_t)((int)((pos >> b) & 1u) - (int)((neg >> b) & 1u));
        } else {
            return X.s[v];
        }
    }
};

// A face row into registers: one 8-byte LDS read when rows are 8 entries
template <int MFV>
__device__ inline void load_row(const uint8_t (&row)[MFV], uint8_t (&r)[MFV]) {
    if constexpr (MFV == 8) {
        uint64_t w;
        __builtin_memcpy(&w, row, 8);
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (uint8_t)(w >> (8 * i));
    } else {
#pragma unroll
        for (int i = 0; i < MFV; ++i) r[i] = row[i];
    }
}

// A face's vertex ids and their signs in registers.  Entries past the
// vertex count hold sign 0.  Returns the vertex count.
template <class C>
__device__ inline int load_face(const WaveIA<C>& X, const SignView<C>& S, int f, uint8_t (&fvr)[C::MFV],
                                int8_t (&sr)[C::MFV]) {
    const int n = X.fn[f];
    load_row<C::MFV>(X.fv[f], fvr);
#pragma unroll
    for (int i = 0; i < C::MFV; ++i) sr[i] = i < n ? S.get(X, fvr[i]) : (int8_t)0;
    return n;
}

// bit 0: some vertex with s > 0, bit 1: some with s < 0
template <int MFV>
__device__ inline uint8_t sign_flags(const int8_t (&sr)[MFV]) {
    uint8_t fl = 0;
#pragma unroll
    for (int i = 0; i < MFV; ++i) fl |= sr[i] > 0 ? 1 : (sr[i] < 0 ? 2 : 0);
    return fl;
}

// The entry after position i of a face's cycle of n entries held in
// registers (i is an unrolled constant, so both candidates are fixed
// registers and the choice is one select, not a scan of the row).
template <int MFV, class T>
__device__ inline T cyc_next(const T (&r)[MFV], int i, int n) {
    return i + 1 == n ? r[0] : r[i + 1 < MFV ? i + 1 : 0];
}

// One side (sgn = +1 / -1) of a face split, streamed over the face's cycle
// with the crossing vertices inserted; the cycle starts at the first kept
// position and an edge between two kept positions that were not adjacent
// lies on the new plane.  The face comes in registers: its vertex ids, edge
// labels and vertex signs (fvr, fer, sr), and the crossing vertex after each
// position (cr, valid where the edge changes sign strictly).  Returns the
// vertex count, or -1 past MFV.
template <int MFV>
__device__ inline int split_side_regs(const uint8_t (&fvr)[MFV], const uint8_t (&fer)[MFV], const int8_t (&sr)[MFV],
                                      const uint8_t (&cr)[MFV], int n, int sgn, uint8_t pid, uint8_t* dv, uint8_t* de) {
    int cnt = 0, p = 0, first = -1, prev = -1;
    uint8_t pv = 0, pe = 0;
    bool over = false;
    auto visit = [&](uint8_t v, int sv, uint8_t e) {
        if (sv * sgn >= 0) {
            if (prev >= 0) {
                if (cnt >= MFV)
                    over = true;
                else {
                    dv[cnt] = pv;
                    de[cnt] = (prev + 1 == p) ? pe : pid;
                }
                ++cnt;
            } else {
                first = p;
            }
            prev = p;
            pv = v;
            pe = e;
        }
        ++p;
    };
#pragma unroll
    for (int i = 0; i < MFV; ++i) {
        if (i < n) {
            const int sw = cyc_next<MFV>(sr, i, n);
            visit(fvr[i], sr[i], fer[i]);
            if (sr[i] * sw < 0) visit(cr[i], 0, fer[i]);
        }
    }
    if (prev >= 0) {
        if (cnt >= MFV)
            over = true;
        else {
            dv[cnt] = pv;
            de[cnt] = ((prev + 1) % p == first) ? pe : pid;
        }
        ++cnt;
    }
    return over ? -1 : cnt;
}

// Runs the insertion of k planes (X.P[0..k-1], loaded by the caller) into
// the tet; on return X.err is set when a capacity or the predicates'
// exponent range was exceeded.
// start: planes already inserted (0, or 2 when the caller has loaded the
// complex after the first two, which depends on their signs only).  Always
// inlined: as a called function (several kernels instantiate it) the calling
// convention would save and restore the live registers through scratch memory
// on every call.
template <class C, int G = 64>
__device__ __forceinline__ void wave_ia_cut(WaveIA<C>& X, int k, int start = 0) {
    using W = Grp<G>;
    const int lane = W::lane();
    if (start == 0 && lane < 4) {
        const int i = lane;
        for (int q = 0; q < 3; ++q) X.vl[i][q] = (uint8_t)(q < i ? q : q + 1);  // corner i: planes != i
        X.fa[i] = (uint8_t)i;
        X.fb[i] = 0;
        X.fc[i] = NIL;
        X.fn[i] = 3;
    }
    if (start == 0 && lane == 0) {
        const uint8_t cyc[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
        for (int i = 0; i < 4; ++i)
            for (int q = 0; q < 3; ++q) {
                X.fv[i][q] = cyc[i][q];
                X.fe[i][q] = cyc[i][(q + 2) % 3];
            }
        X.nv = 4;
        X.nf = 4;
        X.nc = 1;
        X.dup = 0;
        X.err = (4 + k > C::MP) ? 1 : 0;
    }
    if (start == 0 && lane < C::MP) {
        X.grp[lane] = lane < 4 ? (uint8_t)lane : NIL;
        X.ori[lane] = 1;
    }
    __syncthreads();
    if (X.err) return;
    auto fail = [&]() {
        if (lane == 0) X.err = 1;
        __syncthreads();
    };
    for (int j = start; j < k; ++j) {
        const uint8_t pid = (uint8_t)(4 + j);
        const int nv0 = X.nv, nf0 = X.nf, nc0 = X.nc;

        // classify every vertex against plane j
        bool e = false, ap = false, an = false;
        SignView<C> S;
        if constexpr (SignView<C>::kMask) {
            uint64_t pm = 0, nm = 0;
            for (int r0 = 0; r0 < nv0; r0 += G) {
                const int v = r0 + lane;
                int sv = 0;
                if (v < nv0) {
                    sv = orient_ia(X.P, X.vl[v][0], X.vl[v][1], X.vl[v][2], X.P[j], e);
                    X.s[v] = (int8_t)sv;
                }
                pm |= W::ballot(sv > 0) << r0;
                nm |= W::ballot(sv < 0) << r0;
            }
            S.pos = pm;
            S.neg = nm;
            ap = pm != 0;
            an = nm != 0;
        } else {
            for (int v = lane; v < nv0; v += G) {
                const int sv = orient_ia(X.P, X.vl[v][0], X.vl[v][1], X.vl[v][2], X.P[j], e);
                X.s[v] = (int8_t)sv;
                ap |= sv > 0;
                an |= sv < 0;
            }
        }
        const bool any_err = W::any(e), any_pos = W::any(ap), any_neg = W::any(an);
        if (any_err) return fail();
        __syncthreads();

        // each face's sign flags (kept for the face and cell splits), and a
        // plane containing an existing face joins that face's group
        // (the smallest such face: the first round with one, its first lane)
        int cop = INT_MAX;
        for (int r0 = 0; r0 < nf0; r0 += G) {
            const int f = r0 + lane;
            bool on = false;
            if (f < nf0) {
                uint8_t fvr[C::MFV];
                int8_t sr[C::MFV];
                (void)load_face<C>(X, S, f, fvr, sr);
                const uint8_t fl = sign_flags<C::MFV>(sr);
                X.fflag[f] = fl;
                on = fl == 0;
            }
            const uint64_t b = W::ballot(on);
            if (b && cop == INT_MAX) cop = r0 + __builtin_ctzll(b);
        }
        if (cop != INT_MAX) {
            if (lane == 0) {
                const uint8_t cell = X.fb[cop];
                int sv = 0;
                for (int g = 0; g < nf0 && sv == 0; ++g) {
                    if (X.fb[g] != cell && X.fc[g] != cell) continue;
                    for (int q = 0; q < X.fn[g]; ++q)
                        if (X.s[X.fv[g][q]] != 0) {
                            sv = X.s[X.fv[g][q]];
                            break;
                        }
                }
                const uint8_t pl = X.fa[cop];
                X.grp[pid] = X.grp[pl];
                X.ori[pid] = (sv > 0) ? X.ori[pl] : (uint8_t)(1 - X.ori[pl]);
                X.dup = 1;
            }
            __syncthreads();
            continue;
        }
        if (lane == 0) {
            X.grp[pid] = pid;
            X.ori[pid] = 1;
        }
        if (!any_pos || !any_neg) {
            __syncthreads();
            continue;
        }

        // strict sign-change edges in (face, edge) order: each lane loads its
        // face's vertex ids and signs in parallel, counts and files them
        constexpr int MFV = C::MFV, CAND = WaveIA<C>::CAND, HT = WaveIA<C>::HT;
        for (int h = lane; h < HT; h += G) X.u.c.ht[h] = 0xFFFFFFFFu;
        // only faces with both signs have strict sign changes; the same pass
        // ranks them for the face splits
        uint32_t ncand = 0, nmix = 0;
        for (int base = 0; base < nf0; base += G) {
            const int f = base + lane;
            uint8_t fvr[MFV];
            int8_t sr[MFV];
            int n = 0;
            uint32_t c = 0;
            const bool mixed = f < nf0 && X.fflag[f] == 3;
            if (mixed) {
                n = load_face<C>(X, S, f, fvr, sr);
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    const int sw = cyc_next<MFV>(sr, i, n);
                    c += i < n && sr[i] * sw < 0;
                }
            }
            static_assert(MFV < 16, "crossings per face fit four bits");
            uint32_t o = ncand + W::template excl<4>(c);
            const uint32_t m = nmix + W::template excl<1>(mixed ? 1u : 0u);
            if (mixed) {
                X.frank[f] = (uint16_t)m;
                X.fcand[f] = (uint16_t)o;
#pragma unroll
                for (int i = 0; i < MFV; ++i) {
                    if (i >= n) continue;
                    const int sw = cyc_next<MFV>(sr, i, n);
                    const uint8_t vw = cyc_next<MFV>(fvr, i, n);
                    if (sr[i] * sw >= 0) continue;
                    if (o < (uint32_t)CAND) {
                        const uint8_t u = fvr[i];
                        X.u.c.ckey[o] = (uint16_t)((u < vw ? u : vw) << 8 | (u < vw ? vw : u));
                        X.u.c.cface[o] = (uint16_t)f;
                        X.u.c.cedge[o] = (uint8_t)i;
                    }
                    ++o;
                }
            }
            ncand += W::template sum<4>(c);
            nmix += W::template sum<1>(mixed ? 1u : 0u);
        }
        if (ncand > (uint32_t)CAND) return fail();
        __syncthreads();
        // the first candidate of each edge: an LDS hash from key to the
        // smallest candidate index (keys are 16 bits, indices below CAND)
        auto home = [&](uint32_t key) { return (key * 2654435761u) >> 16 & (uint32_t)(HT - 1); };
        // smallest candidate index of a key already inserted
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
        if (nv0 + (int)nnew > C::MV) return fail();
        __syncthreads();
        for (int c = lane; c < (int)ncand; c += G) {
            if (X.u.c.cfirst[c]) {
                const int f = X.u.c.cface[c];
                uint8_t t0 = X.fa[f], t1 = X.fe[f][X.u.c.cedge[c]], t2 = pid;
                sort3(t0, t1, t2);
                const int id = X.u.c.cvid[c];
                X.vl[id][0] = t0;
                X.vl[id][1] = t1;
                X.vl[id][2] = t2;
                X.s[id] = 0;
            } else {
                X.u.c.cvid[c] = X.u.c.cvid[first_of(X.u.c.ckey[c])];
            }
        }
        __syncthreads();

        // face splits: the positive part keeps the id, the negative part is
        // appended in face order
        if (nf0 + (int)nmix > C::MF) return fail();
        bool ferr = false;
        for (int f = lane; f < nf0; f += G) {
            if (X.fflag[f] != 3) continue;
            // the face in registers: ids, labels and signs load in parallel,
            // then both sides are streamed without further LDS reads but
            // the crossing vertices; the positive side overwrites the row
            constexpr int MFV = C::MFV;
            uint8_t fvr[MFV], fer[MFV], cr[MFV];
            int8_t sr[MFV];
            const int n = load_face<C>(X, S, f, fvr, sr);
            load_row<MFV>(X.fe[f], fer);
            uint32_t ci = X.fcand[f];
#pragma unroll
            for (int i = 0; i < MFV; ++i) {
                cr[i] = 0;
                if (i < n) {
                    const int sw = cyc_next<MFV>(sr, i, n);
                    if (sr[i] * sw < 0) cr[i] = X.u.c.cvid[ci++];
                }
            }
            const int dn = nf0 + X.frank[f];
            const int nneg = split_side_regs<MFV>(fvr, fer, sr, cr, n, -1, pid, X.fv[dn], X.fe[dn]);
            const int npos = split_side_regs<MFV>(fvr, fer, sr, cr, n, +1, pid, X.fv[f], X.fe[f]);
            if (nneg < 0 || npos < 0) {
                ferr = true;
                continue;
            }
            X.fn[dn] = (uint8_t)nneg;
            X.fa[dn] = X.fa[f];
            X.fb[dn] = X.fb[f];
            X.fc[dn] = X.fc[f];
            X.fn[f] = (uint8_t)npos;
        }
        if (W::any(ferr)) return fail();
        const int nf1 = nf0 + (int)nmix;
        for (int c = lane; c < nc0; c += G) X.u.e.cflag[c] = 0;
        __syncthreads();

        // cell splits: the positive side keeps the id, the negative side and
        // one face on the new plane are appended in cell order
        // a face's flags after the splits follow from its flags before: a
        // split face's positive part keeps its positive vertices and gains
        // crossing vertices of sign 0 (flags 1), its negative part (appended
        // past nf0) the negative ones (flags 2)
        for (int g = lane; g < nf1; g += G) {
            uint8_t fl = 2;
            if (g < nf0) {
                fl = X.fflag[g];
                if (fl == 3) fl = 1;
            }
            X.fflag[g] = fl;
            if (fl) {
                atomicOr(&X.u.e.cflag[X.fb[g]], (uint32_t)fl);
                if (X.fc[g] != NIL) atomicOr(&X.u.e.cflag[X.fc[g]], (uint32_t)fl);
            }
        }
        __syncthreads();
        const uint32_t nsplit = wave_prefix<G, 1>(
            nc0, [&](int c) { return (uint32_t)(X.u.e.cflag[c] == 3); },
            [&](int c, uint32_t o) { X.u.e.crank[c] = X.u.e.cflag[c] == 3 ? (uint16_t)o : (uint16_t)0xFFFF; });
        if (nf1 + (int)nsplit > C::MF || nc0 + (int)nsplit > WaveIA<C>::MC) return fail();
        __syncthreads();
        for (int g = lane; g < nf1; g += G) {
            if (!(X.fflag[g] & 2)) continue;
            const uint8_t b = X.fb[g], c = X.fc[g];
            if (X.u.e.crank[b] != 0xFFFF) X.fb[g] = (uint8_t)(nc0 + X.u.e.crank[b]);
            if (c != NIL && X.u.e.crank[c] != 0xFFFF) X.fc[g] = (uint8_t)(nc0 + X.u.e.crank[c]);
        }
        __syncthreads();
        // the new face of each split cell, bounded by the edges on the plane
        // of the faces on its positive side (fb == c: edge taken b -> a) and
        // negative side (fc == c: a -> b).  Lanes take faces and file each
        // such edge under its cell's row (the cell's split rank) with its
        // (face, position) order; a lane per split cell then chains its row
        // from the earliest edge, always to the earliest edge that continues
        // it: the order of a scan of the faces one by one.
        if (nsplit > (uint32_t)WaveIA<C>::ROWS) return fail();
        for (int c = lane; c < (int)nsplit; c += G) X.u.e.ccnt[c] = 0;
        __syncthreads();
        bool cerr = false;
        for (int g = lane; g < nf1; g += G) {
            const uint8_t b = X.fb[g], cc = X.fc[g];
            const int rb = (b < nc0 && X.u.e.crank[b] != 0xFFFF) ? (int)X.u.e.crank[b] : -1;
            const int rc = (cc != NIL && cc < nc0 && X.u.e.crank[cc] != 0xFFFF) ? (int)X.u.e.crank[cc] : -1;
            if (rb < 0 && rc < 0) continue;
            uint8_t fvr[C::MFV];
            int8_t sr[C::MFV];
            const int n = load_face<C>(X, S, g, fvr, sr);
#pragma unroll
            for (int i = 0; i < C::MFV; ++i) {
                if (i >= n) continue;
                const uint8_t w = cyc_next<C::MFV>(fvr, i, n);
                const int sw = cyc_next<C::MFV>(sr, i, n);
                const uint8_t a = fvr[i];
                if (sr[i] != 0 || sw != 0) continue;
                for (int side = 0; side < 2; ++side) {
                    const int row = side ? rc : rb;
                    if (row < 0) continue;
                    const uint32_t slot = atomicAdd(&X.u.e.ccnt[row], 1u);
                    if (slot >= (uint32_t)C::MFV) {
                        cerr = true;
                        continue;
                    }
                    X.u.e.tv[row][slot] = side ? a : w;
                    X.u.e.tb[row][slot] = side ? w : a;
                    X.u.e.te[row][slot] = X.fa[g];
                    X.u.e.tk[row][slot] = (uint16_t)(g << 4 | i);
                }
            }
        }
        if (W::any(cerr)) return fail();
        __syncthreads();
        for (int c = lane; c < nc0 && !cerr; c += G) {
            if (X.u.e.crank[c] == 0xFFFF) continue;
            const int row = X.u.e.crank[c];
            const int ne = (int)X.u.e.ccnt[row];
            if (ne < 3) {
                cerr = true;
                break;
            }
            // the row in registers; chaining then needs no LDS round trips
            constexpr int MFV = C::MFV;
            uint8_t ea[MFV], eb[MFV], el[MFV];
            uint16_t ek[MFV];
#pragma unroll
            for (int q = 0; q < MFV; ++q) {
                ea[q] = X.u.e.tv[row][q];
                eb[q] = X.u.e.tb[row][q];
                el[q] = X.u.e.te[row][q];
                ek[q] = q < ne ? X.u.e.tk[row][q] : (uint16_t)0xFFFF;
            }
            int start = 0;
            uint16_t kmin = ek[0];
#pragma unroll
            for (int q = 1; q < MFV; ++q)
                if (ek[q] < kmin) {
                    kmin = ek[q];
                    start = q;
                }
            const int F = nf1 + row;
            int cur = start, cnt = 0;
            for (;;) {
                if (cnt >= ne) {  // the chain does not close within the row
                    cerr = true;
                    break;
                }
                uint8_t a = 0, b = 0, l = 0;
#pragma unroll
                for (int q = 0; q < MFV; ++q)
                    if (q == cur) {
                        a = ea[q];
                        b = eb[q];
                        l = el[q];
                    }
                X.fv[F][cnt] = a;
                X.fe[F][cnt] = l;
                ++cnt;
                int nxt = -1;
                uint16_t knxt = 0xFFFF;
#pragma unroll
                for (int q = 0; q < MFV; ++q)
                    if (q < ne && ea[q] == b && (nxt < 0 || ek[q] < knxt)) {
                        nxt = q;
                        knxt = ek[q];
                    }
                if (nxt < 0) {
                    cerr = true;
                    break;
                }
                cur = nxt;
                if (cur == start) break;
            }
            if (cerr || cnt != ne) {
                cerr = true;
                break;
            }
            X.fn[F] = (uint8_t)cnt;
            X.fa[F] = pid;
            X.fb[F] = (uint8_t)c;
            X.fc[F] = (uint8_t)(nc0 + row);
        }
        if (W::any(cerr)) return fail();
        __syncthreads();
        if (lane == 0) {
            X.nv = nv0 + (int)nnew;
            X.nf = nf1 + (int)nsplit;
            X.nc = nc0 + (int)nsplit;
        }
        __syncthreads();
    }
}

// Writes the record (include/rin/cut_record.h layout) with the whole wave and
// returns its length, or -1 when it exceeds cap; niv / nif / nfv are the
// tet's iso-vertex, iso-face and iso-face-corner counts (uniform).  When the
// slot has room, the word offset of every face within the record follows the
// record (nf words past its length), so that readers need not walk the
// variable-length face list.
template <class C, int G = 64>
__device__ __forceinline__ int wave_ia_record(WaveIA<C>& X, int k, uint16_t* out, int cap, uint32_t& niv, uint32_t& nif,
                              uint32_t& nfv) {
    using W = Grp<G>;
    const int lane = W::lane();
    const int L = 4 + k, nv = X.nv, nf = X.nf;
    static_assert(4 + C::MFV < 16, "face words fit four bits");
    const uint32_t fwords = wave_prefix<G, 4>(
        nf, [&](int f) { return (uint32_t)(4 + X.fn[f]); }, [&](int f, uint32_t o) { X.u.r.foff[f] = (uint16_t)o; });
    const int total = 8 + 4 * nv + (int)fwords + (X.dup ? 2 * L : 0);
    for (int v = lane; v < nv; v += G) X.u.r.isov[v] = 0;
    __syncthreads();
    uint32_t cf = 0, cfv = 0;
    for (int f = lane; f < nf; f += G) {
        bool is = X.fa[f] > 3;
        if (!is && X.dup) {
            const uint8_t rep = X.grp[X.fa[f]];
            for (int q = 4; q < L; ++q) is |= X.grp[q] == rep;
        }
        if (!is) continue;
        ++cf;
        cfv += X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) X.u.r.isov[X.fv[f][q]] = 1;
    }
    nif = W::sum(cf);
    nfv = W::sum(cfv);
    __syncthreads();
    uint32_t cv = 0;
    for (int v = lane; v < nv; v += G) cv += X.u.r.isov[v];
    niv = W::sum(cv);
    if (total > cap) return -1;
    for (int v = lane; v < nv; v += G) {
        uint16_t* o = out + 8 + 4 * v;
        o[0] = X.vl[v][0];
        o[1] = X.vl[v][1];
        o[2] = X.vl[v][2];
        o[3] = 0xFFFF;
    }
    const int fb0 = 8 + 4 * nv;
    for (int f = lane; f < nf; f += G) {
        uint16_t* o = out + fb0 + X.u.r.foff[f];
        o[0] = X.fa[f];
        o[1] = X.fb[f];
        o[2] = X.fc[f] == NIL ? 0xFFFF : X.fc[f];
        o[3] = X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) o[4 + q] = X.fv[f][q];
    }
    if (X.dup)
        for (int l = lane; l < L; l += G) {
            out[fb0 + fwords + l] = X.grp[l];
            out[fb0 + fwords + L + l] = X.ori[l];
        }
    if (total + nf <= cap)
        for (int f = lane; f < nf; f += G) out[total + f] = (uint16_t)(fb0 + X.u.r.foff[f]);
    if (lane == 0) {
        out[0] = 1;
        out[1] = (uint16_t)L;
        out[2] = (uint16_t)nv;
        out[3] = (uint16_t)nf;
        out[4] = (uint16_t)X.nc;
        out[5] = X.dup ? 1 : 0;
        out[6] = (uint16_t)total;
        out[7] = 0;
    }
    return total;
}

}  // namespace dev
}  // namespace rin
