// Batched form of the implicit-arrangement hot path.
//
//  * Tets are bucketed by active-function count k.  k = 1 (most active tets)
//    never runs the cut: its complex depends only on the four corner signs,
//    so an 81-entry table built from the cut itself gives counts, iso
//    vertices, iso faces and the record; each block stages the table in LDS.
//  * k = 2 goes by a sign-vector table; 3..4 and 5..8 run per-bucket cut
//    kernels sized for that bucket: lane groups per tet with the complex in
//    LDS, or one lane per tet (persistent grids pulling 64-tet chunks from a
//    work queue); a small autotuner times the candidates on the run's own
//    buckets.
//  * Every stage reads its element count from device memory and checks it
//    against the workspace capacity, so the whole path is captured once as a
//    hipGraph and replayed; the run's closing kernel publishes the sizes and the
//    overflow flags into mapped host memory (on overflow the workspace grows
//    and the graph is re-captured).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>

#include "cut_wave.h"
#include "kernels_ia.h"

namespace rin {

namespace {

using CK2 = dev::Caps<16, 24, 6, 6>;
// CMed (k 3..4) and CBig (k 5..8) come from kernels_ia.h

// words of a record slot: the largest record a bucket's capacity allows
// (header, vertices, faces with their cycles, coincidence groups)
template <class C>
constexpr uint32_t rec_words() {
    return 8 + 4 * C::MV + C::MF * (4 + C::MFV) + 2 * C::MP;
}

constexpr uint32_t K1_FLAG = 0x80000000u;
constexpr uint32_t ERR_OVERFLOW = 8;

// Counters in the workspace (device).  Sums that many workgroups add to
// go to kSpread words 16 apart (separate cache lines, so the adds do not all
// queue on one address) and are folded into their counter at the end:
// record words read by the extraction, and per cut bucket (k = 3..4, 5..8)
// the active functions and the record words written.
constexpr int kSpreadBase = 64, kSpread = 64, kSpreadStride = 16;
enum SpreadCounter { SP_RECW = 0, SP_CUTK3, SP_CUTW3, SP_CUTK5, SP_CUTW5, SP_N };
constexpr int kQueueBase = 40;  // work counters of the thread-form cuts
__host__ __device__ inline uint32_t spread_at(int counter, uint32_t slot) {
    return kSpreadBase + ((uint32_t)counter * kSpread + slot % kSpread) * kSpreadStride;
}
enum Ctr : int {
    C_ACTIVE = 0, C_FIT, C_VO, C_FO, C_FV, C_VERTS, C_FACES, C_FACEVERTS, C_B1, C_B2, C_B3, C_B5, C_BX, C_DEGEN,
    C_Q2, C_Q3, C_Q5, C_ABORT, C_TETFACES, C_RECW, C_DFAIL, C_CUTK3, C_CUTW3, C_CUTK5, C_CUTW5, C_E3, C_E5, C_E3B, C_N
};

struct K1Entry {
    uint8_t niv, nif, nfv, valid;
    uint8_t vj[4];
    uint8_t vt[4];  // 0..3 corner, 4 + local edge index
    uint8_t fl, fb, fn, pad;
    uint8_t fv[4];
    uint16_t rec_off, rec_len;
};
static_assert(sizeof(K1Entry) == 24, "k1 entry layout");

struct K1Table {
    K1Entry e[81];
    std::vector<uint16_t> words;
};

// one complex per sign pattern, from the cut itself (host build of the
// device algorithm)
const K1Table& k1_table() {
    static K1Table T;
    static std::once_flag once;
    std::call_once(once, [] {
        std::memset(T.e, 0, sizeof(T.e));
        static dev::Complex<CK2> X;
        for (int pat = 0; pat < 81; ++pat) {
            double P[1][4];
            int np = 0, nn = 0;
            for (int c = 0, q = pat; c < 4; ++c, q /= 3) {
                P[0][c] = (double)(q % 3 - 1);
                np += P[0][c] > 0;
                nn += P[0][c] < 0;
            }
            K1Entry& E = T.e[pat];
            if (np == 4 || nn == 4) continue;
            dev::ia_cut<CK2>(P, 1, X);
            if (X.err) throw std::runtime_error("k1 table: cut failed");
            uint16_t rec[512];
            const int len = dev::write_record_ia<CK2>(X, 1, rec, 512);
            E.valid = 1;
            E.rec_off = (uint16_t)T.words.size();
            E.rec_len = (uint16_t)len;
            T.words.insert(T.words.end(), rec, rec + len);
            bool iso_v[CK2::MV] = {false};
            int isof = -1, niso_f = 0;
            for (int f = 0; f < X.nf; ++f) {
                bool is = X.fa[f] > 3;
                if (!is && X.dup) is = X.grp[4] == X.grp[X.fa[f]];
                if (!is) continue;
                ++niso_f;
                isof = f;
                for (int q = 0; q < X.fn[f]; ++q) iso_v[X.fv[f][q]] = true;
            }
            if (niso_f > 1) throw std::runtime_error("k1 table: more than one iso face");
            int idx[CK2::MV];
            int n = 0;
            for (int j = 0; j < X.nv; ++j) {
                if (!iso_v[j]) continue;
                if (n >= 4) throw std::runtime_error("k1 table: more than four iso vertices");
                idx[j] = n;
                E.vj[n] = (uint8_t)j;
                int bnd[3], nb = 0;
                for (int q = 0; q < 3; ++q)
                    if (X.vl[j][q] <= 3) bnd[nb++] = X.vl[j][q];
                if (nb == 3) {
                    int c = 0;
                    while (c == bnd[0] || c == bnd[1] || c == bnd[2]) ++c;
                    E.vt[n] = (uint8_t)c;
                } else if (nb == 2) {
                    int l[2], m = 0;
                    for (int c = 0; c < 4; ++c)
                        if (c != bnd[0] && c != bnd[1]) l[m++] = c;
                    E.vt[n] = (uint8_t)(4 + dev::local_edge(l[0], l[1]));
                } else {
                    throw std::runtime_error("k1 table: unexpected vertex type");
                }
                ++n;
            }
            E.niv = (uint8_t)n;
            E.nif = (uint8_t)niso_f;
            if (isof >= 0) {
                E.fl = (uint8_t)isof;
                E.fb = X.fc[isof] == dev::NIL ? (uint8_t)(1 + X.fa[isof]) : 0;
                E.fn = X.fn[isof];
                for (int q = 0; q < X.fn[isof]; ++q) E.fv[q] = (uint8_t)idx[X.fv[isof][q]];
                E.nfv = E.fn;
            }
        }
    });
    return T;
}

// k = 2: the first plane's complex depends on its corner signs only (the
// k = 1 table); inserting the second plane depends on that complex and the
// second plane's sign at each of its vertices only.  Those signs are the
// second plane's corner signs and, at each vertex where the first plane
// crosses a tet edge, a sign fixed by the corner signs unless they differ
// across that edge (then one exact predicate).  Every sign vector the
// corner signs allow is enumerated once on the host, running the insertion
// itself; a tet then needs at most four predicates and one table read.

struct K2Table {
    K2Pat pat[81];
    uint32_t nslots = 0;
    std::vector<uint32_t> off;  // word offset of each slot's record, or ~0u
    std::vector<uint32_t> cnt;  // iso vertices | iso faces << 8 | iso-face corners << 16
    std::vector<uint16_t> words;
    // The complex after both planes, per slot (byte offset, or ~0u): the state
    // the lane-group cuts of 3..8 functions continue from.  Bytes: nv, nf, nc,
    // dup; grp[0..5]; ori[0..5]; per vertex its plane triple; per face fa, fb,
    // fc, fn, fv[6], fe[6].
    std::vector<uint32_t> snap_off;
    std::vector<uint8_t> snap;
};


const K2Table& k2_table() {
    static K2Table T;
    static std::once_flag once;
    std::call_once(once, [] {
        std::memset(T.pat, 0, sizeof(T.pat));
        static dev::Complex<CK2> X1, X2;
        auto first_plane = [&](int pat) {
            double P[2][4] = {};
            for (int c = 0, q = pat; c < 4; ++c, q /= 3) P[0][c] = (double)(q % 3 - 1);
            dev::ia_init<CK2>(X1, 2);
            int8_t sg[CK2::MV];
            dev::ia_classify<CK2>(P, 0, X1, sg);
            dev::ia_insert<CK2>(X1, 4, sg);
            if (X1.err) throw std::runtime_error("k2 table: first insertion failed");
        };
        uint32_t base = 0;
        for (int pat = 0; pat < 81; ++pat) {
            int np = 0, nn = 0;
            for (int c = 0, q = pat; c < 4; ++c, q /= 3) {
                np += q % 3 == 2;
                nn += q % 3 == 0;
            }
            K2Pat& K = T.pat[pat];
            if (np == 4 || nn == 4) continue;
            first_plane(pat);
            K.valid = 1;
            K.nv = (uint8_t)X1.nv;
            K.base = base;
            for (int v = 4; v < X1.nv; ++v) {
                // labels {f1, f2, 4}: the edge joins the two corners other than f1, f2
                int n = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != X1.vl[v][0] && c != X1.vl[v][1]) K.e[v - 4][n++] = (uint8_t)c;
            }
            uint32_t p3 = 1;
            for (int v = 0; v < X1.nv; ++v) p3 *= 3;
            base += p3;
        }
        T.nslots = base;
        T.off.assign(base, K2_NONE);
        T.cnt.assign(base, 0);
        T.snap_off.assign(base, K2_NONE);
        for (int pat = 0; pat < 81; ++pat) {
            const K2Pat& K = T.pat[pat];
            if (!K.valid) continue;
            first_plane(pat);
            const int ncr = K.nv - 4;
            for (int sb = 0; sb < 81; ++sb) {
                int sB[4];
                for (int c = 0, q = sb; c < 4; ++c, q /= 3) sB[c] = q % 3 - 1;
                int forced[4], nfree = 0;
                for (int q = 0; q < ncr; ++q) {
                    const int a = sB[K.e[q][0]], b = sB[K.e[q][1]];
                    forced[q] = (a * b < 0) ? 2 : (a == b ? a : (a != 0 ? a : b));
                    nfree += forced[q] == 2;
                }
                int p3f = 1;
                for (int q = 0; q < nfree; ++q) p3f *= 3;
                for (int fc = 0; fc < p3f; ++fc) {
                    int8_t sg[CK2::MV];
                    for (int c = 0; c < 4; ++c) sg[c] = (int8_t)sB[c];
                    for (int q = 0, r = fc; q < ncr; ++q) {
                        if (forced[q] == 2) {
                            sg[4 + q] = (int8_t)(r % 3 - 1);
                            r /= 3;
                        } else {
                            sg[4 + q] = (int8_t)forced[q];
                        }
                    }
                    uint32_t slot = K.base, w = 1;
                    for (int v = 0; v < K.nv; ++v, w *= 3) slot += (uint32_t)(sg[v] + 1) * w;
                    X2 = X1;
                    dev::ia_insert<CK2>(X2, 5, sg);
                    if (X2.err) continue;  // the cut fails the same way; the device reports it
                    uint16_t rec[512];
                    const int len = dev::write_record_ia<CK2>(X2, 2, rec, 512);
                    if (len < 0) throw std::runtime_error("k2 table: record overflow");
                    bool iso_v[CK2::MV] = {false};
                    uint32_t nif = 0, nfv = 0, niv = 0;
                    for (int f = 0; f < X2.nf; ++f) {
                        bool is = X2.fa[f] > 3;
                        if (!is && X2.dup) is = X2.grp[4] == X2.grp[X2.fa[f]] || X2.grp[5] == X2.grp[X2.fa[f]];
                        if (!is) continue;
                        ++nif;
                        nfv += X2.fn[f];
                        for (int q = 0; q < X2.fn[f]; ++q) iso_v[X2.fv[f][q]] = true;
                    }
                    for (int v = 0; v < X2.nv; ++v) niv += iso_v[v];
                    T.off[slot] = (uint32_t)T.words.size();
                    T.cnt[slot] = niv | nif << 8 | nfv << 16;
                    T.words.insert(T.words.end(), rec, rec + len);
                    static_assert(CK2::MP == 6 && CK2::MFV == 6, "k2 snapshot layout");
                    T.snap_off[slot] = (uint32_t)T.snap.size();
                    const uint8_t head[4] = {(uint8_t)X2.nv, (uint8_t)X2.nf, (uint8_t)X2.nc, (uint8_t)X2.dup};
                    T.snap.insert(T.snap.end(), head, head + 4);
                    for (int p = 0; p < 6; ++p) T.snap.push_back(X2.grp[p]);
                    for (int p = 0; p < 6; ++p) T.snap.push_back(X2.ori[p]);
                    for (int v = 0; v < X2.nv; ++v)
                        for (int q = 0; q < 3; ++q) T.snap.push_back(X2.vl[v][q]);
                    for (int f = 0; f < X2.nf; ++f) {
                        const uint8_t fh[4] = {X2.fa[f], X2.fb[f], X2.fc[f], X2.fn[f]};
                        T.snap.insert(T.snap.end(), fh, fh + 4);
                        for (int q = 0; q < 6; ++q) T.snap.push_back(q < X2.fn[f] ? X2.fv[f][q] : (uint8_t)0);
                        for (int q = 0; q < 6; ++q) T.snap.push_back(q < X2.fn[f] ? X2.fe[f][q] : (uint8_t)0);
                    }
                }
            }
        }
    });
    return T;
}

// Extraction plans of the k = 2 table records: what the record pass would
// derive from each record, precomputed once.  A plan is bytes:
//   [niv] [nif], then per iso vertex in vertex order [j] [plane 0..2] (the
//   vertex's plane triple), then per iso face in face order
//   [face] [function slot] [boundary code] [n] [n iso-vertex ranks],
// where the function slot is the supporting plane (or, for a tet face, the
// implicit plane of its coincidence group) minus 4 and the boundary code is
// 1 + the tet face for faces on the tet boundary, else 0.
struct K2Plans {
    std::vector<uint32_t> off;  // per k = 2 slot: byte offset, or K2_PLAN_NONE
    std::vector<uint8_t> bytes;
};
constexpr uint32_t K2_PLAN_NONE = 0xFFFFFFFFu;

const K2Plans& k2_plans() {
    static const K2Plans P = [] {
        const K2Table& K2 = k2_table();
        K2Plans Q;
        Q.off.assign(K2.nslots, K2_PLAN_NONE);
        for (uint32_t slot = 0; slot < K2.nslots; ++slot) {
            if (K2.off[slot] == K2_NONE) continue;
            const uint16_t* r = K2.words.data() + K2.off[slot];
            if (r[0] != 1) continue;
            const int L = r[1], nv = r[2], nf = r[3];
            const bool dup = r[5] & 1;
            const uint32_t len = r[6];
            const uint16_t* grp = r + len - 2 * L;
            std::vector<uint8_t> isov(nv, 0);
            std::vector<const uint16_t*> iso_faces;
            std::vector<int> iso_f;
            const uint16_t* F = r + 8 + 4 * nv;
            for (int f = 0; f < nf; ++f) {
                const int n = F[3];
                bool iso = F[0] > 3;
                if (!iso && dup)
                    for (int q = 4; q < L && !iso; ++q) iso = grp[q] == grp[F[0]];
                if (iso) {
                    for (int q = 0; q < n; ++q) isov[F[4 + q]] = 1;
                    iso_faces.push_back(F);
                    iso_f.push_back(f);
                }
                F += 4 + n;
            }
            std::vector<int> rank(nv, -1);
            int niv = 0;
            for (int j = 0; j < nv; ++j)
                if (isov[j]) rank[j] = niv++;
            if (niv > 255 || iso_faces.size() > 255) throw std::logic_error("k2 plan: record too large");
            Q.off[slot] = (uint32_t)Q.bytes.size();
            Q.bytes.push_back((uint8_t)niv);
            Q.bytes.push_back((uint8_t)iso_faces.size());
            for (int j = 0; j < nv; ++j)
                if (isov[j]) {
                    Q.bytes.push_back((uint8_t)j);
                    for (int q = 0; q < 3; ++q) Q.bytes.push_back((uint8_t)r[8 + 4 * j + q]);
                }
            for (size_t i = 0; i < iso_faces.size(); ++i) {
                const uint16_t* G = iso_faces[i];
                int sp = G[0];
                if (sp <= 3)
                    for (int q = 4; q < L; ++q)
                        if (grp[q] == grp[sp]) {
                            sp = q;
                            break;
                        }
                Q.bytes.push_back((uint8_t)iso_f[i]);
                Q.bytes.push_back((uint8_t)(sp - 4));
                Q.bytes.push_back(G[2] == 0xFFFF ? (uint8_t)(1 + G[0]) : (uint8_t)0);
                Q.bytes.push_back((uint8_t)G[3]);
                for (int q = 0; q < G[3]; ++q) Q.bytes.push_back((uint8_t)rank[G[4 + q]]);
            }
        }
        return Q;
    }();
    return P;
}

__device__ inline int bucket_of(int k) { return k <= 1 ? 0 : (k == 2 ? 1 : (k <= 4 ? 2 : (k <= 8 ? 3 : 4))); }

// per tet: active count and bucket histogram
__global__ void k_filter_count(GridView g, const uint64_t* __restrict__ pos, const uint64_t* __restrict__ neg,
                               uint64_t full, uint32_t* __restrict__ kcnt, uint32_t* __restrict__ isact,
                               uint32_t* __restrict__ ctr) {
    __shared__ uint32_t hist[8];
    if (threadIdx.x < 8) hist[threadIdx.x] = 0;
    __syncthreads();
    GRID_STRIDE(t, g.T) {
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint64_t ap = pos[v[0]] & pos[v[1]] & pos[v[2]] & pos[v[3]];
        const uint64_t an = neg[v[0]] & neg[v[1]] & neg[v[2]] & neg[v[3]];
        const uint64_t m = full & ~(ap | an);
        const uint32_t k = (uint32_t)__popcll(m);
        kcnt[t] = k;
        isact[t] = k != 0;
        if (k) atomicAdd(&hist[bucket_of((int)k)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 5 && hist[threadIdx.x]) atomicAdd(&ctr[C_B1 + threadIdx.x], hist[threadIdx.x]);
}

struct FillArgs {
    GridView g;
    const uint64_t *pos, *neg;
    uint64_t full;
    const uint32_t *csr_off, *rank;
    uint32_t *csr_off_w, *rank_w;  // written by the cube form
    uint8_t* fit;
    uint32_t *act, *rec_loc, *cnt;
    uint32_t* act_off;  // CSR offset by active rank
    uint8_t* act_k;
    const K1Entry* k1;
    uint32_t* list[3];      // buckets k=2, 3..4, 5..8
    uint64_t* lts[3];       // per slot: CSR offset << 32 | tet
    uint32_t cap[3];
    uint32_t rec_base[3];   // word offsets of each bucket's slots
    uint32_t slot[3];       // words per slot
    uint32_t* ctr;
    uint32_t* errs;
    uint32_t cap_active, cap_fit;
};

__global__ void k_fill_batched(FillArgs A) {
    __shared__ K1Entry tab[81];
    for (int i = threadIdx.x; i < 81; i += blockDim.x) tab[i] = A.k1[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint64_t T = A.g.T;
    // whole waves iterate together so the bucket appends can be aggregated
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < T; base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = base + threadIdx.x;
        uint64_t m = 0;
        uint32_t v[4] = {0, 0, 0, 0};
        if (t < T) {
            dev::tet_corners(A.g, t, v);
            const uint64_t ap = A.pos[v[0]] & A.pos[v[1]] & A.pos[v[2]] & A.pos[v[3]];
            const uint64_t an = A.neg[v[0]] & A.neg[v[1]] & A.neg[v[2]] & A.neg[v[3]];
            m = A.full & ~(ap | an);
        }
        const int k = __popcll(m);
        int b = -1;
        uint32_t r = 0, o0 = 0;
        if (k) {
            r = A.rank[t];
            uint32_t o = A.csr_off[t];
            o0 = o;
            if (r >= A.cap_active || o + (uint32_t)k > A.cap_fit) {
                A.errs[lane] |= ERR_OVERFLOW;
                A.ctr[C_ABORT] = 1;
            } else {
                A.act[r] = (uint32_t)t;
                A.act_k[r] = (uint8_t)k;
                A.act_off[r] = o;
                for (uint64_t mm = m; mm; mm &= mm - 1) A.fit[o++] = (uint8_t)(__ffsll((unsigned long long)mm) - 1);
                if (k == 1) {
                    const uint32_t f = __ffsll((unsigned long long)m) - 1;
                    int pat = 0, w = 1;
                    for (int c = 0; c < 4; ++c, w *= 3) {
                        const int s = ((A.pos[v[c]] >> f) & 1) ? 2 : (((A.neg[v[c]] >> f) & 1) ? 0 : 1);
                        pat += s * w;
                    }
                    const K1Entry& E = tab[pat];
                    A.cnt[3 * r] = E.niv;
                    A.cnt[3 * r + 1] = E.nif;
                    A.cnt[3 * r + 2] = E.nfv;
                    A.rec_loc[r] = K1_FLAG | (uint32_t)pat;
                } else {
                    b = k == 2 ? 0 : (k <= 4 ? 1 : (k <= 8 ? 2 : -1));
                    if (b < 0) A.errs[lane] |= 1;  // more active functions than the device cut supports
                }
            }
        }
        // wave-aggregated appends to the bucket lists
        for (int q = 0; q < 3; ++q) {
            const uint64_t mask = __ballot(b == q);
            if (!mask) continue;
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t start = 0;
            if (lane == leader) start = atomicAdd(&A.ctr[C_Q2 + q], (uint32_t)__popcll(mask));
            start = __shfl(start, leader, 64);
            if (b == q) {
                const uint32_t slot = start + (uint32_t)__popcll(mask & lt);
                if (slot < A.cap[q]) {
                    A.list[q][slot] = r;
                    A.lts[q][slot] = (uint64_t)o0 << 32 | (uint32_t)t;
                    A.rec_loc[r] = A.rec_base[q] + slot * A.slot[q];
                } else {
                    A.errs[lane] |= ERR_OVERFLOW;
                    A.ctr[C_ABORT] = 1;
                }
            }
        }
    }
}

// per vertex: sign masks; the count of exactly-zero values goes to ctr
// (1024 lanes a workgroup: on the 101^3-vertex, 20-function grid with cold
// caches 37.6 us against 40.6 us at 256; streaming loads, as ordinary loads
// took 72 us when the caches held other kernels' dirty lines)
__global__ void __launch_bounds__(1024) k_signs_count(const double* __restrict__ vals, uint64_t V, uint32_t F, uint64_t* __restrict__ pos,
                              uint64_t* __restrict__ neg, uint32_t* __restrict__ ctr) {
    uint32_t z = 0;
    GRID_STRIDE(v, V) {
        uint64_t p = 0, n = 0;
        // eight function rows in flight per lane: the pass is bound by the
        // number of loads outstanding, not by their issue
        for (uint32_t f0 = 0; f0 < F; f0 += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                x[u] = f0 + u < F ? __builtin_nontemporal_load(&vals[(uint64_t)(f0 + u) * V + v]) : 1.0;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                p |= (uint64_t)(x[u] > 0) << (f0 + u);
                n |= (uint64_t)(x[u] < 0) << (f0 + u);
                z += x[u] == 0;
            }
        }
        pos[v] = p;
        neg[v] = n;
    }
    for (int d = 32; d > 0; d >>= 1) z += __shfl_xor(z, d, 64);
    if ((threadIdx.x & 63) == 0 && z) atomicAdd(&ctr[C_DEGEN], z);
}

// ---- regular grids: one cube (five tets) per thread
//
// The count pass and the fill pass each load a cube's eight corner sign
// masks once for its five tets and need no per-tet intermediate arrays: the
// count pass leaves per tile of kCubeTile cubes its active functions, active
// tets and bucket sizes, one workgroup scans those, and each fill workgroup
// places its tets in the CSR, the active list and the bucket lists from its
// tile's prefixes, in tet order and without atomics.

// cubes per tile (one workgroup): 256, 512 or 1024, picked per plan by the
// autotuner (the results do not depend on it)
constexpr int kCubeTile = 512;  // the default

// pat (optional): per tet with one active function, its corner sign pattern
// (sum of s_c 3^c, s_c = 0 negative, 1 zero, 2 positive), else 0
__device__ inline void cube_masks(const GridView& g, uint32_t cube, const uint64_t* __restrict__ pos,
                                  const uint64_t* __restrict__ neg, uint64_t full, uint32_t c[8], uint64_t m[5],
                                  bool& even, uint8_t* pat = nullptr) {
    const uint32_t R = (uint32_t)g.res, N = R + 1;
    const uint32_t i = cube / (R * R), rem = cube - i * (R * R), j = rem / R, k = rem - j * R;
    const uint32_t b = i * N * N + j * N + k;
    c[0] = b;
    c[1] = b + N * N;
    c[2] = b + N * N + N;
    c[3] = b + N;
    for (int a = 0; a < 4; ++a) c[4 + a] = c[a] + 1;
    even = ((i + j + k) & 1u) == 0;
    uint64_t P[8], Q[8];
    for (int a = 0; a < 8; ++a) {
        P[a] = pos[c[a]];
        Q[a] = neg[c[a]];
    }
    for (int q = 0; q < 5; ++q) {
        const uint8_t* t = even ? dev::kCubeTetEven[q] : dev::kCubeTetOdd[q];
        const uint64_t ap = P[t[0]] & P[t[1]] & P[t[2]] & P[t[3]];
        const uint64_t an = Q[t[0]] & Q[t[1]] & Q[t[2]] & Q[t[3]];
        m[q] = full & ~(ap | an);
        if (pat) {
            pat[q] = 0;
            if (__popcll(m[q]) == 1) {
                const uint32_t f = __ffsll((unsigned long long)m[q]) - 1;
                int p = 0, w = 1;
                for (int a = 0; a < 4; ++a, w *= 3)
                    p += (((P[t[a]] >> f) & 1) ? 2 : (((Q[t[a]] >> f) & 1) ? 0 : 1)) * w;
                pat[q] = (uint8_t)p;
            }
        }
    }
}

// Per tile of cubes: active functions, active tets and tets per bucket
// (k = 1, 2, 3..4, 5..8, > 8).  No atomics: the tile scan below turns the
// tile sums into prefixes (the CSR offsets and the bucket-list slots of each
// tile) and writes the totals, so the bucket lists come out in tet order.
constexpr int kTileFields = 8;
enum TileField { TF_FIT = 0, TF_ACTIVE, TF_B1, TF_B2, TF_B3, TF_B5, TF_BX };
struct TileSums {
    uint32_t v[kTileFields];
};

// A tet with 3..8 active functions, listed by the count pass with its place
// inside its tile of cubes: with the tile's prefixes (after the tile scan) it
// gives the tet's active rank and bucket slot, so its cut can run beside the
// fill pass instead of after it.
struct EarlyEntry {
    uint32_t t, tile;
    uint32_t lo;   // active functions of the tile's tets before it
    uint32_t lrb;  // active tets before it | tets of its bucket list before it << 16
};

// early3, early5 (nullable): lists of the k = 3..4 and 5..8 tets, at most
// early_cap3 / early_cap5 entries, counted in ctr[C_E3] / ctr[C_E5] (zero at
// the start of a run).  The k = 3..4 list holds the k = 4 tets from its front
// (ctr[C_E3] of them) and the k = 3 tets from its back (ctr[C_E3B]), so that a
// wavefront of the cut mostly holds tets of one k (two insertions past the
// two-plane start, or one) and the longer ones start first.  A tile's entries
// are contiguous and in tet order (the back part reversed).
template <int TILE>
__global__ void __launch_bounds__(TILE) k_cube_count(GridView g, const uint64_t* __restrict__ pos,
                                                          const uint64_t* __restrict__ neg, uint64_t full,
                                                          TileSums* __restrict__ tile_sums, EarlyEntry* early3,
                                                          EarlyEntry* early5, uint32_t early_cap3,
                                                          uint32_t early_cap5, uint32_t* ctr) {
    const uint32_t ncubes = (uint32_t)(g.T / 5);
    const uint32_t tile = blockIdx.x;
    const uint32_t cube = tile * TILE + threadIdx.x;
    uint32_t v[kTileFields] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t m[5] = {0, 0, 0, 0, 0};
    uint32_t n4 = 0;  // k = 4 tets (of the k = 3..4 bucket)
    if (cube < ncubes) {
        uint32_t c[8];
        bool even;
        cube_masks(g, cube, pos, neg, full, c, m, even);
        for (int q = 0; q < 5; ++q) {
            const int k = __popcll(m[q]);
            v[TF_FIT] += (uint32_t)k;
            v[TF_ACTIVE] += k != 0;
            if (k) v[TF_B1 + bucket_of(k)]++;
            n4 += k == 4;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // tiles holding listed tets rank them inside the tile: exclusive scans of
    // (active functions, active tets) packed as in the fill pass, and of the
    // listed counts in 16-bit fields (k = 4 | k = 3 << 16 | k = 5..8 << 32;
    // a tile holds at most 5 TILE tets); one atomic add per list part and
    // tile reserves the tile's entries
    const uint64_t want = (early3 ? (uint64_t)n4 | (uint64_t)(v[TF_B3] - n4) << 16 : 0ull) |
                          (early5 ? (uint64_t)v[TF_B5] << 32 : 0ull);
    if ((early3 || early5) && __syncthreads_or(want != 0)) {
        __shared__ uint32_t esum[TILE / 64], ebase[3];
        __shared__ uint64_t ebk[TILE / 64];
        constexpr int kAsBits = 13;
        const uint32_t packed = v[TF_FIT] << kAsBits | v[TF_ACTIVE];
        uint32_t inc = packed;
        uint64_t binc = want;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            const uint64_t ob = __shfl_up(binc, d, 64);
            if (lane >= d) {
                inc += o;
                binc += ob;
            }
        }
        if (lane == 63) {
            esum[wave] = inc;
            ebk[wave] = binc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t tot = 0;
            for (int w = 0; w < TILE / 64; ++w) tot += ebk[w];
            const uint32_t t4 = (uint32_t)tot & 0xffffu, t3 = (uint32_t)(tot >> 16) & 0xffffu,
                           t5 = (uint32_t)(tot >> 32);
            ebase[0] = t4 ? atomicAdd(ctr + C_E3, t4) : 0u;
            ebase[1] = t5 ? atomicAdd(ctr + C_E5, t5) : 0u;
            ebase[2] = t3 ? atomicAdd(ctr + C_E3B, t3) : 0u;
        }
        uint32_t ex = inc - packed;
        uint64_t bx = binc - want;
        for (int w = 0; w < wave; ++w) {
            ex += esum[w];
            bx += ebk[w];
        }
        __syncthreads();
        uint32_t lo = ex >> kAsBits, lr = ex & ((1u << kAsBits) - 1u);
        // the tile's k = 4, k = 3 and k = 5..8 tets before this one
        uint32_t b4 = (uint32_t)bx & 0xffffu, b3 = (uint32_t)(bx >> 16) & 0xffffu, b5 = (uint32_t)(bx >> 32);
        for (int q = 0; q < 5; ++q) {
            const int k = __popcll(m[q]);
            const uint32_t t = cube * 5u + (uint32_t)q;
            if (k >= 3 && k <= 4 && early3) {
                const uint32_t rank = b3 + b4;  // place in the tile's bucket list
                if (k == 4) {
                    const uint32_t at = ebase[0] + b4++;
                    if (at < early_cap3) early3[at] = {t, tile, lo, lr | rank << 16};
                } else {
                    const uint32_t back = ebase[2] + b3++;
                    if (back < early_cap3) early3[early_cap3 - 1 - back] = {t, tile, lo, lr | rank << 16};
                }
            } else if (k >= 5 && k <= 8 && early5) {
                const uint32_t at = ebase[1] + b5;
                if (at < early_cap5) early5[at] = {t, tile, lo, lr | b5 << 16};
                ++b5;
            }
            lo += (uint32_t)k;
            lr += k != 0;
        }
    }
    __shared__ uint32_t red[TILE / 64][kTileFields];
    for (int x = 0; x < kTileFields - 1; ++x) {
        for (int d = 32; d > 0; d >>= 1) v[x] += __shfl_xor(v[x], d, 64);
        if (lane == 0) red[wave][x] = v[x];
    }
    __syncthreads();
    if (threadIdx.x < kTileFields) {
        uint32_t t = 0;
        if (threadIdx.x < kTileFields - 1)
            for (int w = 0; w < TILE / 64; ++w) t += red[w][threadIdx.x];
        tile_sums[tile].v[threadIdx.x] = t;
    }
}

// Exclusive prefixes of the tile sums, in place, and the totals: CSR size,
// active tets, the bucket histogram and the bucket-list sizes.  One
// workgroup, 1024 tiles per pass, transposed through LDS: wave f scans field
// f, each lane over 16 consecutive tiles.
constexpr int kTileScanThreads = 1024;

__global__ void __launch_bounds__(kTileScanThreads) k_tile_scan(TileSums* __restrict__ ts, uint32_t n,
                                                                 uint32_t* __restrict__ ctr) {
    constexpr int P = kTileScanThreads + kTileScanThreads / 16;  // a pad word every 16 against bank conflicts
    __shared__ uint32_t sm[kTileFields][P];
    __shared__ uint32_t carry[kTileFields];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < kTileFields) carry[threadIdx.x] = 0;
    for (uint32_t base = 0; base < n; base += kTileScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const int pi = threadIdx.x + threadIdx.x / 16;
        TileSums x = {};
        if (i < n) x = ts[i];
        for (int f = 0; f < kTileFields; ++f) sm[f][pi] = x.v[f];
        __syncthreads();
        if (wave < kTileFields) {
            const int f = wave, e0 = 16 * lane, p0 = e0 + e0 / 16;
            uint32_t sum = 0;
            for (int e = 0; e < 16; ++e) sum += sm[f][p0 + e];
            uint32_t inc = sum;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(inc, d, 64);
                if (lane >= d) inc += o;
            }
            uint32_t run = carry[f] + inc - sum;
            for (int e = 0; e < 16; ++e) {
                const uint32_t v = sm[f][p0 + e];
                sm[f][p0 + e] = run;
                run += v;
            }
            const uint32_t total = __shfl(inc, 63, 64);
            if (lane == 0) carry[f] += total;
        }
        __syncthreads();
        if (i < n) {
            TileSums y;
            for (int f = 0; f < kTileFields; ++f) y.v[f] = sm[f][pi];
            ts[i] = y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctr[C_FIT] = carry[TF_FIT];
        ctr[C_ACTIVE] = carry[TF_ACTIVE];
        ctr[C_B1] = carry[TF_B1];
        ctr[C_B2] = carry[TF_B2];
        ctr[C_B3] = carry[TF_B3];
        ctr[C_B5] = carry[TF_B5];
        ctr[C_BX] = carry[TF_BX];
        ctr[C_Q2] = carry[TF_B2];
        ctr[C_Q3] = carry[TF_B3];
        ctr[C_Q5] = carry[TF_B5];
    }
}

// tile_sums: exclusive prefixes (k_tile_scan)
template <int TILE>
__global__ void __launch_bounds__(TILE) k_cube_fill(FillArgs A, const TileSums* __restrict__ tile_sums) {
    __shared__ K1Entry tab[81];
    __shared__ uint32_t wsum[TILE / 64];
    __shared__ uint32_t wbk[TILE / 64];  // per-wave bucket counts, 10 bits per list
    __shared__ uint32_t s_off[5 * TILE], s_rank[5 * TILE];  // per-tet CSR entries of the tile
    for (int i = threadIdx.x; i < 81; i += blockDim.x) tab[i] = A.k1[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tile = blockIdx.x;
    const TileSums pre = tile_sums[tile];
    const uint32_t ncubes = (uint32_t)(A.g.T / 5);
    const uint32_t cube = tile * TILE + threadIdx.x;
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t m[5] = {0, 0, 0, 0, 0};
    bool even = true;
    uint8_t pat[5] = {0, 0, 0, 0, 0};
    if (cube < ncubes) cube_masks(A.g, cube, A.pos, A.neg, A.full, c, m, even, pat);
    // per lane: active functions, active tets, and tets per bucket list
    // (k = 2, 3..4, 5..8), packed ks << 11 | as and 10 bits per bucket
    uint32_t ks = 0, as = 0, bp = 0;
    for (int q = 0; q < 5; ++q) {
        const int k = __popcll(m[q]);
        ks += (uint32_t)k;
        as += k != 0;
        if (k >= 2 && k <= 8) bp += 1u << (10 * (k == 2 ? 0 : (k <= 4 ? 1 : 2)));
    }
    // bucket slots: the tile's prefix, the waves before this one, the lanes
    // before this one (a wave holds at most 320 tets of a list)
    uint32_t binc = bp;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(binc, d, 64);
        if (lane >= d) binc += o;
    }
    if (lane == 63) wbk[wave] = binc;
    // block exclusive scan of (ks, as) packed as ks << kAsBits | as (a tile
    // holds at most 5 TILE active tets and 64 times as many functions)
    constexpr int kAsBits = 13;
    static_assert(5 * TILE < (1 << kAsBits) && 5 * 64 * TILE < (1 << (32 - kAsBits)), "tile packing");
    const uint32_t packed = ks << kAsBits | as;
    uint32_t inc = packed;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t ex = inc - packed;
    for (int w = 0; w < wave; ++w) ex += wsum[w];
    uint32_t o = pre.v[TF_FIT] + (ex >> kAsBits), r = pre.v[TF_ACTIVE] + (ex & ((1u << kAsBits) - 1u));
    uint32_t bnext[3];
    for (int qq = 0; qq < 3; ++qq) {
        uint32_t b = pre.v[TF_B2 + qq] + (((binc - bp) >> (10 * qq)) & 1023u);
        for (int w = 0; w < wave; ++w) b += (wbk[w] >> (10 * qq)) & 1023u;
        bnext[qq] = b;
    }
    for (int q = 0; q < 5; ++q) {
        const uint64_t t = (uint64_t)cube * 5 + q;
        const uint64_t mq = m[q];
        const int k = __popcll(mq);
        s_off[5 * threadIdx.x + q] = o;
        s_rank[5 * threadIdx.x + q] = r;
        if (k) {
            if (r >= A.cap_active || o + (uint32_t)k > A.cap_fit) {
                A.errs[lane] |= ERR_OVERFLOW;
                A.ctr[C_ABORT] = 1;
            } else {
                A.act[r] = (uint32_t)t;
                A.act_k[r] = (uint8_t)k;
                A.act_off[r] = o;
                uint32_t oo = o;
                for (uint64_t mm = mq; mm; mm &= mm - 1) A.fit[oo++] = (uint8_t)(__ffsll((unsigned long long)mm) - 1);
                if (k == 1) {
                    const K1Entry& E = tab[pat[q]];
                    A.cnt[3 * r] = E.niv;
                    A.cnt[3 * r + 1] = E.nif;
                    A.cnt[3 * r + 2] = E.nfv;
                    A.rec_loc[r] = K1_FLAG | (uint32_t)pat[q];
                } else if (k <= 8) {
                    const int qq = k == 2 ? 0 : (k <= 4 ? 1 : 2);
                    const uint32_t slot = bnext[qq]++;
                    if (slot < A.cap[qq]) {
                        A.list[qq][slot] = r;
                        A.lts[qq][slot] = (uint64_t)o << 32 | (uint32_t)t;
                        A.rec_loc[r] = A.rec_base[qq] + slot * A.slot[qq];
                    } else {
                        A.errs[lane] |= ERR_OVERFLOW;
                        A.ctr[C_ABORT] = 1;
                    }
                } else {
                    A.errs[lane] |= 1;  // more active functions than the device cut supports
                }
            }
        }
        o += (uint32_t)k;
        r += k != 0;
    }
    // the tile's per-tet CSR entries, coalesced
    __syncthreads();
    const uint64_t t0 = (uint64_t)tile * (5 * TILE), T = A.g.T;
    for (int i = threadIdx.x; i < 5 * TILE; i += TILE)
        if (t0 + i < T) {
            A.csr_off_w[t0 + i] = s_off[i];
            A.rank_w[t0 + i] = s_rank[i];
        }
}

struct CutArgs {
    GridView g;
    const double* vals;
    const uint32_t *act, *csr_off, *list, *rec_loc;
    const uint64_t* lts;  // per slot: CSR offset << 32 | tet
    const uint8_t *act_k, *fit;
    const uint32_t* n;   // bucket size (device)
    const uint32_t* abort;
    uint32_t* queue;     // work counter
    uint32_t cap, slot;
    uint32_t F;          // functions (bounds the speculative corner-value loads)
    uint16_t* rec;
    uint32_t *cnt, *errs;
    uint32_t* ctr;       // for the work statistics (spread counters)
    int stat_k, stat_w;  // spread counters of this bucket's functions and words (-1: none)
    uint8_t* ws;         // thread form: one LaneCut slot per lane of the grid
    // early form (beside the fill pass): the count pass's list and the tile
    // prefixes (tf: the tile field of the bucket's slots); the tet's
    // functions come from its corner sign masks
    const EarlyEntry* early;
    const uint32_t *n_early, *n_back;  // the list's front and back parts (n_back nullable)
    const TileSums* tiles;
    const uint64_t *pos, *neg;
    uint64_t full;
    uint32_t rec_base, cap_active;
    int tf;
    K2Snap k2s;  // lane-group forms: the complexes after two planes (off null: none)
};

// persistent per-bucket cut: each wave takes 64 bucket entries at a time
template <class C>
__global__ void __launch_bounds__(64) k_cut_bucket(CutArgs A) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = *A.abort ? 0u : min(*A.n, A.cap);
    const uint64_t V = A.g.V;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(A.queue, 64u);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const uint32_t i = base + lane;
        if (i >= n) continue;
        const uint32_t r = A.list[i];
        const int k = A.act_k[r];
        const uint32_t t = A.act[r];
        uint32_t v[4];
        dev::tet_corners(A.g, t, v);
        const uint32_t s0 = A.csr_off[t];
        double P[C::MP - 4][4];
        for (int q = 0; q < k; ++q) {
            const uint32_t f = A.fit[s0 + q];
            for (int c = 0; c < 4; ++c) P[q][c] = A.vals[(uint64_t)f * V + v[c]];
        }
        const uint64_t slot = (uint64_t)blockIdx.x * 64 + lane;
        dev::Complex<C>& X = LaneCut<C>::X(A.ws, slot);
        dev::ia_cut<C>(P, k, X, LaneCut<C>::T(A.ws, slot));
        uint16_t* out = A.rec + A.rec_loc[r];
        if (X.err) {
            A.errs[lane] |= 1;
            out[0] = 0;
            A.cnt[3 * r] = A.cnt[3 * r + 1] = A.cnt[3 * r + 2] = 0;
            continue;
        }
        const int len = dev::write_record_ia<C>(X, k, out, (int)A.slot);
        if (len < 0) A.errs[lane] |= 4;
        if (A.stat_k >= 0 && len > 0) {
            atomicAdd(A.ctr + spread_at(A.stat_k, blockIdx.x * 64 + lane), (uint32_t)k);
            atomicAdd(A.ctr + spread_at(A.stat_w, blockIdx.x * 64 + lane), (uint32_t)len);
        }
        uint32_t a0, a1, a2;
        ia_counts<C>(X, a0, a1, a2);
        A.cnt[3 * r] = a0;
        A.cnt[3 * r + 1] = a1;
        A.cnt[3 * r + 2] = a2;
    }
}

// wave-cooperative cut: one lane group of G lanes per bucket entry (a block
// is a single wavefront carrying 64 / G entries), the complex in LDS; the
// group's lanes share the predicates and the splits of each insertion
//
// Early: the tets come from the count pass's list (EarlyEntry) and the tile
// prefixes, and nothing is read that the fill pass writes: the active rank
// and bucket slot are the tile's prefixes plus the tet's place in its tile,
// the functions are the set bits of the tet's corner sign masks (the CSR's
// order), and the record goes to the slot the fill pass assigns the tet.
template <class C, int G, bool Early = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_cut_wave(CutArgs A) {
    // 64 / G tets per wavefront, one lane group each
    constexpr int NG = 64 / G;
    __shared__ dev::WaveIA<C> Xs[NG];
    const int gid = threadIdx.x / G, lane = threadIdx.x & (G - 1);
    dev::WaveIA<C>& X = Xs[gid];
    const uint32_t i = blockIdx.x * NG + gid;
    uint32_t r, rec_at;
    int k;
    if constexpr (Early) {
        // the fill pass shares the CUs: its waves wait on memory, these issue
        // (the one-tet-a-wavefront k = 5..8 cut first)
        __builtin_amdgcn_s_setprio(G == 64 ? 3 : 2);
        // the list's front part, then its back part read from the end; a
        // list past the bucket's capacity is cut by no one (the fill pass
        // flags the overflow)
        const uint32_t nf = *A.n_early, nb = A.n_back ? *A.n_back : 0u;
        if (nf + nb > A.cap || i >= nf + nb) return;
        const EarlyEntry e = A.early[i < nf ? i : A.cap - 1 - (i - nf)];
        const TileSums& pre = A.tiles[e.tile];
        r = pre.v[TF_ACTIVE] + (e.lrb & 0xffffu);
        const uint32_t slot = pre.v[A.tf] + (e.lrb >> 16);
        if (r >= A.cap_active || slot >= A.cap) return;  // the fill pass flags the overflow
        rec_at = A.rec_base + slot * A.slot;
        uint32_t v[4];
        dev::tet_corners(A.g, e.t, v);
        const uint64_t m = A.full & ~((A.pos[v[0]] & A.pos[v[1]] & A.pos[v[2]] & A.pos[v[3]]) |
                                      (A.neg[v[0]] & A.neg[v[1]] & A.neg[v[2]] & A.neg[v[3]]));
        k = __popcll(m);
        if (lane < 4 * (C::MP - 4)) {
            // row q: the q-th active function (rows past k load function 0, unused)
            uint64_t mm = m;
            for (int q = lane >> 2; q > 0 && mm; --q) mm &= mm - 1;
            const uint32_t f = mm ? (uint32_t)(__ffsll((unsigned long long)mm) - 1) : 0u;
            const double x = A.vals[(uint64_t)f * A.g.V + v[lane & 3]];
            if (lane < 4 * k && 4 + k <= C::MP) X.P[lane >> 2][lane & 3] = x;
        }
    } else {
        const uint32_t n = *A.abort ? 0u : min(*A.n, A.cap);
        if (i >= n) return;
        r = A.list[i];
        const uint64_t ts = A.lts[i];
        k = A.act_k[r];
        rec_at = 0;
        // corner values: the loads need not wait for k (rows past the tet's
        // functions read the CSR's next entries, inside its head-room of 64,
        // with the function id clamped into range; their values go unused)
        if (lane < 4 * (C::MP - 4)) {
            const uint32_t t = (uint32_t)ts, s0 = (uint32_t)(ts >> 32);
            uint32_t v[4];
            dev::tet_corners(A.g, t, v);
            const uint32_t fr = A.fit[s0 + (lane >> 2)];
            const uint32_t f = fr < A.F ? fr : 0u;
            const double x = A.vals[(uint64_t)f * A.g.V + v[lane & 3]];
            if (lane < 4 * k && 4 + k <= C::MP) X.P[lane >> 2][lane & 3] = x;
        }
    }
    int start = 0;
    if (A.k2s.off) {
        __syncthreads();  // the rows, for the slot
        start = k2_start<C, G>(X, A.k2s, k) ? 2 : 0;
    }
    dev::wave_ia_cut<C, G>(X, k, start);
    uint16_t* out = A.rec + (Early ? rec_at : A.rec_loc[r]);
    if (X.err) {
        if (lane == 0) {
            atomicOr(A.errs, 1u);
            out[0] = 0;
            A.cnt[3 * r] = A.cnt[3 * r + 1] = A.cnt[3 * r + 2] = 0;
        }
        return;
    }
    uint32_t a0, a1, a2;
    const int len = dev::wave_ia_record<C, G>(X, k, out, (int)A.slot, a0, a1, a2);
    if (lane == 0) {
        if (len < 0) atomicOr(A.errs, 4u);
        if (A.stat_k >= 0 && len > 0) {
            // words written: the record, and the face offsets when they fit
            const uint32_t w = (uint32_t)len + ((uint32_t)(len + X.nf) <= A.slot ? (uint32_t)X.nf : 0u);
            atomicAdd(A.ctr + spread_at(A.stat_k, i), (uint32_t)k);
            atomicAdd(A.ctr + spread_at(A.stat_w, i), w);
        }
        A.cnt[3 * r] = a0;
        A.cnt[3 * r + 1] = a1;
        A.cnt[3 * r + 2] = a2;
    }
}

struct K2Args {
    GridView g;
    const double* vals;
    const uint64_t *pos, *neg;
    uint64_t full;  // the functions' bits
    const uint32_t *act, *csr_off, *list;
    const uint64_t* lts;  // per slot: CSR offset << 32 | tet
    const uint8_t* fit;
    const uint32_t* n;  // bucket size (device)
    const uint32_t* abort;
    uint32_t cap;
    const K2Pat* pat;
    const uint32_t *slot_off, *slot_cnt;
    uint32_t *rec_loc, *cnt, *errs;
    uint16_t* rec;
};

// the k = 2 table slot of tet t (corners v) whose active functions are
// fa < fb: the first plane's sign pattern picks the pattern, the second
// plane's signs at the corners and at the first plane's edge points the
// slot; K2_NONE when a sign pattern or point falls outside the table
__device__ __forceinline__ uint32_t k2_table_slot(const K2Args& A, const K2Pat* pats, const uint32_t* v, uint32_t fa,
                                                  uint32_t fb) {
    const uint64_t V = A.g.V;
    int pa = 0, sB[4];
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) {
        const uint64_t p = A.pos[v[c]], q = A.neg[v[c]];
        pa += (((p >> fa) & 1) ? 2 : (((q >> fa) & 1) ? 0 : 1)) * w;
        sB[c] = ((p >> fb) & 1) ? 1 : (((q >> fb) & 1) ? -1 : 0);
    }
    const K2Pat& K = pats[pa];
    uint32_t slot = K.base, w = 81;
    for (int c = 0, w3 = 1; c < 4; ++c, w3 *= 3) slot += (uint32_t)(sB[c] + 1) * w3;
    bool err = !K.valid;
    for (int q = 0; q + 4 < K.nv; ++q, w *= 3) {
        const int c1 = K.e[q][0], c2 = K.e[q][1];
        int sg;
        if (sB[c1] * sB[c2] < 0) {
            // second plane at the point of edge c1 c2 where the first vanishes
            sg = dev::orient_ia_edge(A.vals[(uint64_t)fa * V + v[c1]], A.vals[(uint64_t)fa * V + v[c2]],
                                     A.vals[(uint64_t)fb * V + v[c1]], A.vals[(uint64_t)fb * V + v[c2]], err);
        } else {
            sg = sB[c1] == sB[c2] ? sB[c1] : (sB[c1] != 0 ? sB[c1] : sB[c2]);
        }
        slot += (uint32_t)(sg + 1) * w;
    }
    return err ? K2_NONE : slot;
}

// points active tet r at table slot `slot` with its counts; false (flagged,
// zero counts) when the slot holds no record
__device__ __forceinline__ bool k2_point(const K2Args& A, uint32_t slot, uint32_t r) {
    if (slot == K2_NONE || A.slot_off[slot] == K2_NONE) {
        atomicOr(A.errs, (uint32_t)ERR_CAPACITY);
        A.cnt[3 * r] = A.cnt[3 * r + 1] = A.cnt[3 * r + 2] = 0;
        return false;
    }
    const uint32_t c3 = A.slot_cnt[slot];
    A.rec_loc[r] = K1_FLAG | (81u + slot);
    A.cnt[3 * r] = c3 & 0xFF;
    A.cnt[3 * r + 1] = (c3 >> 8) & 0xFF;
    A.cnt[3 * r + 2] = c3 >> 16;
    return true;
}

// k = 2 by table: sign vector of the second plane over the first plane's
// complex -> slot -> record and counts (no per-tet record is written)
__global__ void k_k2_lookup(K2Args A) {
    __shared__ K2Pat pats[81];
    for (int i = threadIdx.x; i < 81; i += blockDim.x) pats[i] = A.pat[i];
    __syncthreads();
    const uint32_t n = *A.abort ? 0u : min(*A.n, A.cap);
    GRID_STRIDE(i, n) {
        const uint32_t r = A.list[i];
        const uint32_t t = (uint32_t)A.lts[i];
        uint32_t v[4];
        dev::tet_corners(A.g, t, v);
        // the two active functions from the corner masks (the CSR's order,
        // one dependent load fewer than reading the CSR)
        const uint64_t m = A.full & ~((A.pos[v[0]] & A.pos[v[1]] & A.pos[v[2]] & A.pos[v[3]]) |
                                      (A.neg[v[0]] & A.neg[v[1]] & A.neg[v[2]] & A.neg[v[3]]));
        const uint32_t fa = (uint32_t)(__ffsll((unsigned long long)m) - 1);
        const uint32_t fb = (uint32_t)(__ffsll((unsigned long long)(m & (m - 1))) - 1);
        const uint32_t slot = k2_table_slot(A, pats, v, fa, fb);
        if (!k2_point(A, slot, r)) A.rec[A.rec_loc[r]] = 0;
    }
}

// The plain form's k = 1 and k = 2 tets by the same tables: the table
// record is copied into the tet's record slot (the plain form's extraction
// reads records there), with the tet's output counts.
struct PlainTab {
    const K1Entry* k1;
    const uint16_t* w1;  // k = 1 records (K1Entry::rec_off)
    const K2Pat* pat;
    const uint32_t *off2, *cnt2;
    const uint16_t* w2;  // k = 2 records (off2 per slot; a record's word 6 is its length)
    const uint32_t* snap_off;  // complexes after two planes (K2Table::snap_off, snap)
    const uint8_t* snap;
};

__global__ void k_table_plain(GridView g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                              const uint8_t* __restrict__ act_k, const uint32_t* __restrict__ csr_off,
                              const uint8_t* __restrict__ fit, const uint32_t* __restrict__ list, uint64_t n,
                              const uint32_t* __restrict__ rec_off, uint16_t* __restrict__ rec,
                              uint32_t* __restrict__ cnt, uint32_t* __restrict__ errs, PlainTab T) {
    __shared__ K1Entry k1[81];
    __shared__ K2Pat pats[81];
    for (int i = threadIdx.x; i < 81; i += blockDim.x) {
        k1[i] = T.k1[i];
        pats[i] = T.pat[i];
    }
    __syncthreads();
    const uint64_t V = g.V;
    GRID_STRIDE(i, n) {
        const uint32_t a = list[i];
        const int k = act_k[a];
        const uint32_t t = act[a], s0 = csr_off[t];
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t fa = fit[s0];
        double xa[4];
        int pa = 0;
        for (int c = 0, w = 1; c < 4; ++c, w *= 3) {
            xa[c] = vals[(uint64_t)fa * V + v[c]];
            pa += (xa[c] > 0 ? 2 : (xa[c] < 0 ? 0 : 1)) * w;
        }
        const uint16_t* src = nullptr;
        uint32_t len = 0, c0 = 0, c1 = 0, c2 = 0;
        bool err = false;
        if (k == 1) {
            const K1Entry& E = k1[pa];
            err = !E.valid;
            src = T.w1 + E.rec_off;
            len = E.rec_len;
            c0 = E.niv;
            c1 = E.nif;
            c2 = E.nfv;
        } else {
            const uint32_t fb = fit[s0 + 1];
            double xb[4];
            int sB[4];
            for (int c = 0; c < 4; ++c) {
                xb[c] = vals[(uint64_t)fb * V + v[c]];
                sB[c] = xb[c] > 0 ? 1 : (xb[c] < 0 ? -1 : 0);
            }
            const K2Pat& K = pats[pa];
            err = !K.valid;
            uint32_t slot = K.base, w = 81;
            for (int c = 0, w3 = 1; c < 4; ++c, w3 *= 3) slot += (uint32_t)(sB[c] + 1) * w3;
            for (int q = 0; q + 4 < K.nv && !err; ++q, w *= 3) {
                const int e1 = K.e[q][0], e2 = K.e[q][1];
                int sg;
                if (sB[e1] * sB[e2] < 0)
                    sg = dev::orient_ia_edge(xa[e1], xa[e2], xb[e1], xb[e2], err);
                else
                    sg = sB[e1] == sB[e2] ? sB[e1] : (sB[e1] != 0 ? sB[e1] : sB[e2]);
                slot += (uint32_t)(sg + 1) * w;
            }
            const uint32_t off = err ? K2_NONE : T.off2[slot];
            err = off == K2_NONE;
            if (!err) {
                src = T.w2 + off;
                len = src[6];
                const uint32_t c3 = T.cnt2[slot];
                c0 = c3 & 0xFF;
                c1 = (c3 >> 8) & 0xFF;
                c2 = c3 >> 16;
            }
        }
        uint16_t* dst = rec + rec_off[a];
        if (!err && len > rec_off[a + 1] - rec_off[a]) {
            flag_error(errs, ERR_RECORD);
            err = true;
        } else if (err) {
            flag_error(errs, ERR_CAPACITY);
        }
        if (err) {
            dst[0] = 0;
            cnt[3 * a] = cnt[3 * a + 1] = cnt[3 * a + 2] = 0;
            continue;
        }
        for (uint32_t q = 0; q < len; ++q) dst[q] = src[q];
        cnt[3 * a] = c0;
        cnt[3 * a + 1] = c1;
        cnt[3 * a + 2] = c2;
    }
}

struct ExtractArgs {
    GridView g;
    const uint32_t *act, *csr_off, *rec_loc;
    const uint32_t* act_off;   // CSR offset by active rank
    const uint8_t* fit;
    const uint16_t* rec;
    const K1Entry* k1;
    const uint16_t* tab;       // table records (k = 1 patterns, then k = 2 slots)
    const uint32_t* tab_off;
    const uint32_t* n_active;
    uint32_t cap_active;
    uint32_t* ctr;
    const uint32_t *voff, *foff, *fvoff;
    const uint32_t *tot;  // C_VO, C_FO, C_FV totals live at ctr
    uint32_t cap_vo, cap_fo, cap_fv;
    uint64_t* okey;
    uint32_t* oa;
    uint16_t* oj;
    uint32_t* fq_a;
    uint16_t* fq_l;
    uint8_t *fq_fn, *fq_bnd;
    uint32_t *fq_vs, *fve, *errs;
    const uint32_t* list[3];  // bucket lists (k = 2, 3..4, 5..8) of active ranks
    const uint64_t* lts[3];   // per slot: CSR offset << 32 | tet
    uint32_t list_cap[3];
    uint32_t slot_words[3];   // words per record slot of each bucket
    uint32_t fo_table[3];     // 1: the bucket's cut leaves face offsets after a record that fits
    const uint32_t* plan_off;  // k = 2 table records: extraction plans (k2_plans)
    const uint8_t* plan;
    uint32_t n_plans;
};

// Extraction in two passes so that waves stay uniform: the table pass
// writes the occurrences of every k = 1 tet (81 sign patterns in LDS), the
// record pass those of the tets on the bucket lists (k = 2..8), whose records
// vary in size.
__device__ inline bool extract_guard(const ExtractArgs& A) {
    if (A.ctr[C_ABORT]) return false;
    if (A.tot[C_VO] > A.cap_vo || A.tot[C_FO] > A.cap_fo || A.tot[C_FV] > A.cap_fv) {
        if (blockIdx.x == 0) {
            A.errs[threadIdx.x & 63] |= ERR_OVERFLOW;
            A.ctr[C_ABORT] = 1;
        }
        return false;
    }
    return true;
}

constexpr uint8_t kLocalEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

// The extraction passes as roles of a workgroup (blk: its index among the
// nblk workgroups of the role), so that one launch can run all three.
__device__ inline void extract_k1_role(const ExtractArgs& A, K1Entry* tab, uint32_t blk, uint32_t nblk) {
    for (int i = threadIdx.x; i < 81; i += blockDim.x) tab[i] = A.k1[i];
    __syncthreads();
    if (!extract_guard(A)) return;
    const uint64_t NA = min((uint64_t)*A.n_active, (uint64_t)A.cap_active);
    for (uint64_t a = (uint64_t)blk * blockDim.x + threadIdx.x; a < NA; a += (uint64_t)nblk * blockDim.x) {
        const uint32_t loc = A.rec_loc[a];
        if (!(loc & K1_FLAG) || (loc & ~K1_FLAG) >= 81) continue;
        const uint32_t t = A.act[a];
        uint32_t v[4];
        dev::tet_corners(A.g, t, v);
        const K1Entry& E = tab[loc & 0xFF];
        const uint32_t f = A.fit[A.act_off[a]];
        uint32_t o = A.voff[a];
        for (int i = 0; i < E.niv; ++i) {
            const int vt = E.vt[i];
            A.okey[o] = vt < 4 ? key_corner(v[vt])
                               : key_edge(dev::edge_id(A.g, t, v, kLocalEdge[vt - 4][0], kLocalEdge[vt - 4][1]), f);
            A.oa[o] = (uint32_t)a;
            A.oj[o] = E.vj[i];
            ++o;
        }
        if (E.nif) {
            const uint32_t q = A.foff[a], e = A.fvoff[a], vb = A.voff[a];
            A.fq_a[q] = (uint32_t)a;
            A.fq_l[q] = E.fl;
            A.fq_fn[q] = (uint8_t)f;
            A.fq_bnd[q] = E.fb;
            A.fq_vs[q] = e;
            for (int i = 0; i < E.fn; ++i) A.fve[e + i] = vb + E.fv[i];
        }
    }
}

__global__ void k_extract_k1(ExtractArgs A) {
    __shared__ K1Entry tab[81];
    extract_k1_role(A, tab, blockIdx.x, gridDim.x);
}

// k = 2 records (mostly sign-vector table records): one lane per tet
__device__ inline void extract_rec_role(const ExtractArgs& A, uint32_t blk, uint32_t nblk) {
    if (!extract_guard(A)) return;
    const uint32_t n0 = min(A.ctr[C_Q2], A.list_cap[0]);
    uint32_t recw = 0;  // record words read (throughput accounting)
    for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < n0; i += (uint64_t)nblk * blockDim.x) {
        const uint32_t a = A.list[0][i];
        const uint64_t ts = A.lts[0][i];
        const uint32_t loc = A.rec_loc[a];
        const uint32_t t = (uint32_t)ts, s0 = (uint32_t)(ts >> 32);
        uint32_t v[4];
        dev::tet_corners(A.g, t, v);
        const uint32_t ti = loc & ~K1_FLAG;
        const uint32_t po = ((loc & K1_FLAG) && ti >= 81 && ti - 81 < A.n_plans) ? A.plan_off[ti - 81] : K2_PLAN_NONE;
        if (po != K2_PLAN_NONE) {
            // a table record: stream its plan
            const uint8_t* P = A.plan + po;
            const int niv = P[0], nif = P[1];
            const uint8_t* pv = P + 2;
            const uint32_t vb = A.voff[a];
            for (int x = 0; x < niv; ++x) {
                const int j = pv[4 * x];
                const uint64_t key =
                    ia_vertex_key(A.g, t, v, a, j, pv[4 * x + 1], pv[4 * x + 2], pv[4 * x + 3], A.fit, s0);
                A.okey[vb + x] = key;
                A.oa[vb + x] = a;
                A.oj[vb + x] = (uint16_t)j;
            }
            const uint8_t* pf = pv + 4 * niv;
            uint32_t q = A.foff[a], e = A.fvoff[a];
            for (int x = 0; x < nif; ++x) {
                const int n = pf[3];
                A.fq_a[q] = a;
                A.fq_l[q] = pf[0];
                A.fq_fn[q] = A.fit[s0 + pf[1]];
                A.fq_bnd[q] = pf[2];
                A.fq_vs[q] = e;
                for (int c = 0; c < n; ++c) A.fve[e++] = vb + pf[4 + c];
                ++q;
                pf += 4 + n;
            }
            continue;
        }
        // general record (same layout and rules as the plain extraction); a
        // k = 2 complex has at most 16 vertices, so one 64-bit mask holds
        // its iso vertices
        const Rec R((loc & K1_FLAG) ? A.tab + A.tab_off[loc & ~K1_FLAG] : A.rec + loc);
        if (R.r[0] != 1) continue;
        if (R.nv > 64) {
            A.errs[threadIdx.x & 63] |= ERR_RECORD;
            continue;
        }
        recw += R.r[6];
        uint64_t iso = 0;
        const uint16_t* F = R.faces();
        for (int f = 0; f < R.nf; ++f) {
            const int n = F[3];
            if (ia_face_is_iso(R, F))
                for (int q = 0; q < n; ++q) iso |= 1ull << (F[4 + q] & 63);
            F += 4 + n;
        }
        uint32_t o = A.voff[a];
        for (int j = 0; j < R.nv; ++j) {
            if (!((iso >> j) & 1)) continue;
            const uint16_t* T3 = R.vert(j);
            const uint64_t key =
                ia_vertex_key(A.g, t, v, a, j, T3[0], T3[1], T3[2], A.fit, s0);
            A.okey[o] = key;
            A.oa[o] = (uint32_t)a;
            A.oj[o] = (uint16_t)j;
            ++o;
        }
        uint32_t q = A.foff[a], e = A.fvoff[a];
        const uint32_t vb = A.voff[a];
        F = R.faces();
        for (int f = 0; f < R.nf; ++f) {
            const int n = F[3];
            if (ia_face_is_iso(R, F)) {
                int sp = F[0];
                if (sp <= 3) {
                    const uint16_t* grp = R.groups_ia();
                    for (int p = 4; p < R.L; ++p)
                        if (grp[p] == grp[sp]) {
                            sp = p;
                            break;
                        }
                }
                A.fq_a[q] = (uint32_t)a;
                A.fq_l[q] = (uint16_t)f;
                A.fq_fn[q] = A.fit[s0 + sp - 4];
                A.fq_bnd[q] = F[2] == 0xFFFF ? (uint8_t)(1 + F[0]) : 0;
                A.fq_vs[q] = e;
                for (int i = 0; i < n; ++i) {
                    const int j = F[4 + i] & 63;
                    A.fve[e++] = vb + (uint32_t)__popcll(iso & ((1ull << j) - 1));
                }
                ++q;
            }
            F += 4 + n;
        }
    }
    for (int d = 32; d > 0; d >>= 1) recw += __shfl_xor(recw, d, 64);
    if ((threadIdx.x & 63) == 0 && recw)
        atomicAdd(A.ctr + spread_at(SP_RECW, blk * 4 + (threadIdx.x >> 6)), recw);
}

__global__ void k_extract_records(ExtractArgs A) { extract_rec_role(A, blockIdx.x, gridDim.x); }

// k = 3..8 records: one wave per tet, lanes over faces and vertices.  The
// wave cut leaves each record's face offsets after it, so lanes read their
// faces in place; a record without them (thread-form cut) gets its offsets
// from one lane's walk of the face list.  LDS holds only the offsets and
// the iso flags, so many waves stay resident to cover the loads.
// Occurrences and face rows come out in the order one lane per tet gives.
struct WaveExtractLds {
    uint16_t fo[CBig::MF];
    uint32_t isov[(CBig::MV + 31) / 32];
    uint8_t isof[CBig::MF];
};

__device__ inline void extract_wave_role(const ExtractArgs& A, WaveExtractLds& S, uint32_t blk) {
    uint16_t* fo = S.fo;
    uint32_t* isov = S.isov;
    uint8_t* isof = S.isof;
    if (!extract_guard(A)) return;
    const uint32_t n2 = min(A.ctr[C_Q5], A.list_cap[2]), n1 = min(A.ctr[C_Q3], A.list_cap[1]);
    const uint32_t i = blk;
    if (i >= n1 + n2) return;
    // the larger complexes first
    const int bq = i < n2 ? 2 : 1;
    const uint32_t a = bq == 2 ? A.list[2][i] : A.list[1][i - n2];
    const int lane = threadIdx.x;
    const uint32_t loc = A.rec_loc[a];
    const uint16_t* W = (loc & K1_FLAG) ? A.tab + A.tab_off[loc & ~K1_FLAG] : A.rec + loc;
    if (W[0] != 1) return;
    const uint32_t len = W[6];
    const int nv = W[2], nf = W[3], L = W[1];
    if (nv > CBig::MV || nf > CBig::MF || len < 8u + 4u * nv) {
        if (lane == 0) A.errs[0] |= ERR_RECORD;
        return;
    }
    for (int x = lane; x < (nv + 31) / 32; x += 64) isov[x] = 0;
    if (lane == 0) atomicAdd(A.ctr + spread_at(SP_RECW, blk), len);
    const bool dup = W[5] & 1;
    if (!(loc & K1_FLAG) && A.fo_table[bq] && len + nf <= A.slot_words[bq]) {
        for (int f = lane; f < nf; f += 64) fo[f] = W[len + f];
    } else if (lane == 0) {
        uint32_t off = 8 + 4 * nv;
        for (int f = 0; f < nf; ++f) {
            fo[f] = (uint16_t)off;
            off += off + 3 < len ? 4 + W[off + 3] : 4;
        }
    }
    __syncthreads();
    const uint16_t* grp = W + len - 2 * L;
    for (int f = lane; f < nf; f += 64) {
        const uint16_t* F = W + fo[f];
        bool iso = F[0] > 3;
        if (!iso && dup) {
            const uint16_t rep = grp[F[0]];
            for (int q = 4; q < L && !iso; ++q) iso = grp[q] == rep;
        }
        isof[f] = iso;
        if (iso)
            for (int q = 0; q < F[3]; ++q) atomicOr(&isov[F[4 + q] >> 5], 1u << (F[4 + q] & 31));
    }
    __syncthreads();
    auto vrank = [&](int j) {
        uint32_t r = 0;
        for (int w = 0; w < (j >> 5); ++w) r += __popc(isov[w]);
        return r + __popc(isov[j >> 5] & ((1u << (j & 31)) - 1u));
    };
    const uint32_t t = A.act[a];
    uint32_t v[4];
    dev::tet_corners(A.g, t, v);
    const uint32_t s0 = A.act_off[a];
    const uint32_t vb = A.voff[a];
    for (int j = lane; j < nv; j += 64) {
        if (!((isov[j >> 5] >> (j & 31)) & 1u)) continue;
        const uint16_t* T3 = W + 8 + 4 * j;
        const uint64_t key =
            ia_vertex_key(A.g, t, v, a, j, T3[0], T3[1], T3[2], A.fit, s0);
        const uint32_t o = vb + vrank(j);
        A.okey[o] = key;
        A.oa[o] = a;
        A.oj[o] = (uint16_t)j;
    }
    // iso faces in face order: 64 at a time, ordinals and corner offsets by
    // wave prefix sums
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t q = A.foff[a], e = A.fvoff[a];
    for (int base = 0; base < nf; base += 64) {
        const int f = base + lane;
        const bool iso = f < nf && isof[f];
        const uint16_t* F = W + fo[f < nf ? f : 0];
        const uint32_t n = iso ? F[3] : 0u;
        const uint64_t m = __ballot(iso);
        uint32_t inc = n;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t x = __shfl_up(inc, d, 64);
            if (lane >= d) inc += x;
        }
        if (iso) {
            const uint32_t qi = q + (uint32_t)__popcll(m & lt);
            uint32_t ei = e + inc - n;
            int sp = F[0];
            if (sp <= 3)
                for (int p = 4; p < L; ++p)
                    if (grp[p] == grp[sp]) {
                        sp = p;
                        break;
                    }
            A.fq_a[qi] = a;
            A.fq_l[qi] = (uint16_t)f;
            A.fq_fn[qi] = A.fit[s0 + sp - 4];
            A.fq_bnd[qi] = F[2] == 0xFFFF ? (uint8_t)(1 + F[0]) : 0;
            A.fq_vs[qi] = ei;
            for (uint32_t c = 0; c < n; ++c) A.fve[ei++] = vb + vrank(F[4 + c]);
        }
        q += (uint32_t)__popcll(m);
        e += __shfl(inc, 63, 64);
    }
}

__global__ void __launch_bounds__(64) k_extract_records_wave(ExtractArgs A) {
    __shared__ WaveExtractLds S;
    extract_wave_role(A, S, blockIdx.x);
}

// all three extraction passes in one launch of 64-lane workgroups: the
// lane-per-tet k = 2 records first, the wave-per-tet records of 3..8
// functions next, the k = 1 tets last
__global__ void __launch_bounds__(64) k_extract_all(ExtractArgs A, uint32_t nb_rec, uint32_t nb_wave, uint32_t nb_k1) {
    __shared__ WaveExtractLds S;
    __shared__ K1Entry tab[81];
    const uint32_t b = blockIdx.x;
    if (b < nb_rec)
        extract_rec_role(A, b, nb_rec);
    else if (b < nb_rec + nb_wave)
        extract_wave_role(A, S, b - nb_rec);
    else
        extract_k1_role(A, tab, b - nb_rec - nb_wave, nb_k1);
}

// the host-visible summary of a run: the counters, then the union of the
// error flags (one 64-lane block)
constexpr int kPubWords = 33;
static_assert(C_N <= 32, "the published summary holds 32 counters");
__device__ void publish_counters(const uint32_t* ctr, const uint32_t* errs, uint32_t* pub) {
    if (threadIdx.x < C_N) pub[threadIdx.x] = ctr[threadIdx.x];
    uint32_t e = errs[threadIdx.x];
    for (int d = 32; d > 0; d >>= 1) e |= __shfl_xor(e, d, 64);
    if (threadIdx.x == 0) pub[32] = e;
    __threadfence_system();
}

__global__ void k_publish(const uint32_t* ctr, const uint32_t* errs, uint32_t* pub) {
    publish_counters(ctr, errs, pub);
}

// The faces' emit pass and the run's closing work in one launch. Workgroups
// 1.. emit (each adds up the raw tile counts before its tile; the sizes come
// from the first pass, so no emitting workgroup reads a counter); workgroup 0
// adds up all the tiles (totals into the counters), folds the spread
// counters, closes the CSR and face offset arrays, publishes the counters
// and clears them for the next run.
__global__ void __launch_bounds__(256) k_face_emit_close(FaceTileArgs A, uint32_t* ctr, uint64_t T, uint32_t* csr_off,
                                                         uint32_t* errs, uint32_t* pub) {
    if (blockIdx.x > 0) {
        face_emit_role(A, blockIdx.x - 1);
        return;
    }
    __shared__ uint32_t part[3][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t nq = A.sizes[0];
    const uint32_t nt = (nq + kFaceTile - 1) / kFaceTile;
    uint32_t x[3] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nt; b += 256)
        for (int f = 0; f < 3; ++f) x[f] += A.tiles[4 * (uint64_t)b + f];
    for (int f = 0; f < 3; ++f) {
        for (int d = 32; d > 0; d >>= 1) x[f] += __shfl_xor(x[f], d, 64);
        if (lane == 0) part[f][wave] = x[f];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x == 0) {
            const uint32_t tot[3] = {part[0][0] + part[0][1] + part[0][2] + part[0][3],
                                     part[1][0] + part[1][1] + part[1][2] + part[1][3],
                                     part[2][0] + part[2][1] + part[2][2] + part[2][3]};
            ctr[C_FACES] = tot[0];
            ctr[C_TETFACES] = tot[1];
            ctr[C_FACEVERTS] = tot[2];
            csr_off[T] = ctr[C_FIT];
            // the totals are at most nq (faces, tet-face entries) and the
            // corner capacity, as in the first pass
            A.f_vstart[tot[0]] = tot[2];
            A.f_tfstart[tot[0]] = tot[1];
        }
        const int fold_to[SP_N] = {C_RECW, C_CUTK3, C_CUTW3, C_CUTK5, C_CUTW5};
        for (int c = 0; c < SP_N; ++c) {
            uint32_t v = ctr[spread_at(c, threadIdx.x)];
            for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
            if (threadIdx.x == 0) ctr[fold_to[c]] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) publish_counters(ctr, errs, pub);
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x < C_N) ctr[threadIdx.x] = 0;
        errs[threadIdx.x] = 0;
        for (int c = 0; c < SP_N; ++c) ctr[spread_at(c, threadIdx.x)] = 0;
    }
}

__global__ void k_zero_ctr(uint32_t* ctr, int n, uint32_t* errs) {
    if (threadIdx.x < (unsigned)n) ctr[threadIdx.x] = 0;
    if (threadIdx.x < 64) errs[threadIdx.x] = 0;
    for (int c = 0; c < SP_N; ++c)
        if (threadIdx.x < kSpread) ctr[spread_at(c, threadIdx.x)] = 0;
}


}  // namespace

// the k = 1 and k = 2 tables on each device, uploaded once per process and
// device (they stay until the process ends)
PlainTab plain_tables(hipStream_t s) {
    static std::mutex mu;
    static std::map<int, PlainTab> per_device;
    int dev = 0;
    RIN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(dev);
    if (it != per_device.end()) return it->second;
    const K1Table& K1 = k1_table();
    const K2Table& K2 = k2_table();
    auto up = [&](const void* src, size_t bytes) {
        void* p = nullptr;
        RIN_HIP(hipMalloc(&p, bytes ? bytes : 4));
        if (bytes) RIN_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
        return p;
    };
    PlainTab D;
    D.k1 = static_cast<const K1Entry*>(up(K1.e, sizeof(K1.e)));
    D.w1 = static_cast<const uint16_t*>(up(K1.words.data(), 2 * K1.words.size()));
    D.pat = static_cast<const K2Pat*>(up(K2.pat, sizeof(K2.pat)));
    D.off2 = static_cast<const uint32_t*>(up(K2.off.data(), 4 * K2.off.size()));
    D.cnt2 = static_cast<const uint32_t*>(up(K2.cnt.data(), 4 * K2.cnt.size()));
    D.w2 = static_cast<const uint16_t*>(up(K2.words.data(), 2 * K2.words.size()));
    D.snap_off = static_cast<const uint32_t*>(up(K2.snap_off.data(), 4 * K2.snap_off.size()));
    D.snap = static_cast<const uint8_t*>(up(K2.snap.data(), K2.snap.size()));
    RIN_HIP(hipStreamSynchronize(s));
    per_device[dev] = D;
    return D;
}

void plain_k2_snapshots(hipStream_t s, const void** pat, const uint32_t** off, const uint8_t** bytes) {
    const PlainTab T = plain_tables(s);
    *pat = T.pat;
    *off = T.snap_off;
    *bytes = T.snap;
}

void table_cut_plain(const GridView& g, const double* vals, const uint32_t* act, const uint8_t* act_k,
                     const uint32_t* csr_off, const uint8_t* fit, const uint32_t* list, uint64_t n,
                     const uint32_t* rec_off, uint16_t* rec, uint32_t* cnt, uint32_t* errs, hipStream_t s) {
    if (!n) return;
    const PlainTab T = plain_tables(s);
    hipLaunchKernelGGL(k_table_plain, dim3(nblk(n)), dim3(256), 0, s, g, vals, act, act_k, csr_off, fit, list, n,
                       rec_off, rec, cnt, errs, T);
    RIN_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ plan

struct Caps_ {
    uint32_t active = 0, fit = 0, b[3] = {0, 0, 0}, vo = 0, fo = 0, fv = 0;
};

struct Workspace {
    DevMem mem;  // every buffer (and every temporary of the captured launches) lives here
    Caps_ cap;
    uint32_t rec_base[3] = {0, 0, 0};
    // words per record slot: the largest record a bucket's capacity allows
    // (header, vertices, faces with their cycles, coincidence groups)
    uint32_t slot[3] = {rec_words<CK2>(), rec_words<CMed>(), rec_words<CBig>()};
    bool wave_cut[3] = {false, false, false};  // bucket cut by the wave form (records carry face offsets)
    DBuf<uint64_t> pos, neg;
    DBuf<uint32_t> kcnt, isact, csr_off, rank, act, rec_loc, cnt, list[3];
    DBuf<uint64_t> lts[3];  // per bucket slot: CSR offset << 32 | tet
    DBuf<uint32_t> act_off;  // CSR offset of each active tet, by active rank
    DBuf<TileSums> tile_sums;  // regular grids: per tile of cubes
    int cube_tile = kCubeTile;  // cubes per tile of the count and fill passes
    DBuf<EarlyEntry> early[3];  // regular grids: the k = 3..4 and 5..8 tets ([1], [2]), listed by the count pass
    DBuf<uint8_t> fit, act_k;
    DBuf<uint16_t> rec;
    DBuf<uint32_t> cv, cf, cfv;
    DBuf<uint64_t> okey;
    DBuf<uint32_t> oa;
    DBuf<uint16_t> oj;
    DBuf<uint32_t> fq_a, fq_vs, fve;
    DBuf<uint16_t> fq_l;
    DBuf<uint8_t> fq_fn, fq_bnd;
    DBuf<uint32_t> first, vid, vrank, first_occ, mo, fail;
    DBuf<uint32_t> tile_first;  // star path: first occurrences per tile of occurrences
    DBuf<uint32_t> star_head;  // star tables of the regular grid (k_dedup_star)
    DBuf<uint8_t> star_ent;
    // vertex dedup hash table (zero between runs) and each occurrence's slot
    DBuf<unsigned long long> hkey;
    DBuf<uint32_t> hmx, hslot;
    int hbits = 0;
    bool hash_dirty = false;  // a run stopped early: clear the table before the next
    // faces: per occurrence flags and partner, corner ids, per-tile counts
    DBuf<uint32_t> ffirst, fmatch, fvm, ftiles;
    DBuf<uint32_t> fsizes;  // face occurrences and corner entries, for the emit pass
    // outputs
    DBuf<uint32_t> v_tet, v_sv, f_vstart, f_verts, f_tfstart, f_tf_tet;
    DBuf<uint16_t> v_j, f_tf_local;
    DBuf<uint8_t> v_ss, v_fn, f_func;
    DBuf<double> xyz;
    DBuf<uint32_t> ctr, errs;
    DBuf<K1Entry> k1;
    DBuf<uint16_t> k1_words;  // table records: k = 1 patterns, then k = 2 slots
    DBuf<uint32_t> k1_off;    // word offset per table index (81 + k = 2 slots)
    DBuf<K2Pat> k2pat;
    DBuf<uint32_t> k2snap_off;  // complexes after two planes (lane-group cuts of 3..8 functions)
    DBuf<uint8_t> k2snap;
    bool k2_start = true;       // those cuts start from them
    DBuf<uint32_t> k2plan_off;
    DBuf<uint8_t> k2plan;
    DBuf<uint32_t> k2cnt;
    DBuf<uint8_t> ws_thread[3];  // thread-form cut state per bucket (LaneCut slots)
    // the path as a few captured graphs launched back to back; stream events
    // between them time each part on every replay
    std::vector<std::pair<std::string, hipGraph_t>> graphs;
    std::vector<std::pair<std::string, hipGraphExec_t>> parts;
    std::vector<hipEvent_t> ev;
    // the same path as one graph: replays without stage times
    hipGraph_t whole_graph = nullptr;
    hipGraphExec_t whole = nullptr;
    // side branches of the whole graph (the cut kernels beside each other and
    // beside the fill pass): streams to capture them on, fork and join events
    hipStream_t side = nullptr, side2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
    // pinned, device-mapped summary the closing kernel of a run writes (counters
    // and error flags), read by the host after the stream synchronises
    uint32_t* pub_h = nullptr;
    uint32_t* pub_d = nullptr;
    bool exec = false;  // graphs captured for the current capacities
    int grid[3] = {0, 0, 0};  // cut kernels' grids the graphs were captured with
    // the values the whole graph and the per-stage graphs read (a run with
    // other values of the same size re-records the graph and updates it in place)
    const double* bound_whole = nullptr;
    const double* bound_parts = nullptr;
    bool ctr_clear = false;  // counters and error flags are zero (left so by the last run)
    bool in_use = false;
    void drop_parts() {
        for (auto& p : parts) (void)hipGraphExecDestroy(p.second);
        for (auto& g : graphs) (void)hipGraphDestroy(g.second);
        parts.clear();
        graphs.clear();
        bound_parts = nullptr;
    }
    void drop_graphs() {
        drop_parts();
        if (whole) (void)hipGraphExecDestroy(whole);
        if (whole_graph) (void)hipGraphDestroy(whole_graph);
        whole = nullptr;
        whole_graph = nullptr;
        bound_whole = nullptr;
    }
    ~Workspace() {
        drop_graphs();
        for (auto e : ev) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (join2) (void)hipEventDestroy(join2);
        if (side2) (void)hipStreamDestroy(side2);
        if (side) (void)hipStreamDestroy(side);
        if (pub_h) (void)hipHostFree(pub_h);
    }
};

struct Plan {
    std::mutex mu;
    std::vector<std::shared_ptr<Workspace>> ws;
    int blocks_per_cu[3] = {0, 0, 0};  // autotuned grid of each bucket's cut kernel
    int cube_tile = 0;                  // autotuned cube tile of the count and fill passes (0: not yet)
    uint32_t hint_b[5] = {0, 0, 0, 0, 0};
};

// one plan per (mode, function count) of a grid
struct PlanCache {
    std::mutex mu;
    std::map<uint64_t, std::shared_ptr<Plan>> plans;
    std::shared_ptr<Plan> get(uint64_t key) {
        std::lock_guard<std::mutex> l(mu);
        auto& p = plans[key];
        if (!p) p = std::make_shared<Plan>();
        return p;
    }
};

std::shared_ptr<PlanCache> make_plan_cache() { return std::make_shared<PlanCache>(); }

namespace {

// Stars of the simplices of a regular grid, from the cube decomposition
// itself: every tet of a small grid contributes, for each of its corners,
// edges and faces whose lowest vertex is an interior anchor of either parity,
// its cube (offset from the anchor's cube, each coordinate -1 or 0) and its
// index in the cube.  The decomposition repeats with the parity of i+j+k, so
// the two anchors cover every simplex of every grid.  Entries are in tet order.
struct StarTables {
    std::vector<uint32_t> head;  // [2][kStarShapes]: byte offset << 8 | byte count
    std::vector<uint8_t> ent;    // pairs (packed cube offset, tet in cube)
};

const StarTables& star_tables() {
    static const StarTables T = [] {
        const int res = 4, N = res + 1;
        const uint32_t anchor[2] = {(uint32_t)((1 * N + 1) * N + 2), (uint32_t)((1 * N + 1) * N + 1)};  // parity 0, 1
        std::vector<std::vector<std::pair<uint32_t, uint8_t>>> lists(2 * kStarShapes);
        for (uint64_t t = 0; t < 5ull * res * res * res; ++t) {
            uint32_t v[4];
            dev::regular_tet(res, t, v);
            const uint64_t cube = t / 5;
            const int ci = (int)(cube / (res * res)), cj = (int)((cube / res) % res), ck = (int)(cube % res);
            for (int sub = 1; sub < 16; ++sub) {
                const int nvs = __builtin_popcount(sub);
                if (nvs == 4) continue;
                uint32_t w[3];
                int n = 0;
                for (int c = 0; c < 4; ++c)
                    if (sub >> c & 1) w[n++] = v[c];
                std::sort(w, w + n);
                const int p = w[0] == anchor[0] ? 0 : (w[0] == anchor[1] ? 1 : -1);
                if (p < 0) continue;
                uint32_t shape = 0;
                if (n == 2) shape = 1 + dev::lattice_dir(res, w[0], w[1]);
                if (n == 3) shape = 28 + dev::lattice_dir(res, w[0], w[1]) * 27 + dev::lattice_dir(res, w[0], w[2]);
                const int ai = (int)(w[0] / (N * N)), aj = (int)((w[0] / N) % N), ak = (int)(w[0] % N);
                const int di = ci - ai, dj = cj - aj, dk = ck - ak;
                if (di < -1 || di > 0 || dj < -1 || dj > 0 || dk < -1 || dk > 0)
                    throw std::logic_error("star table: cube offset out of range");
                const uint32_t order = (uint32_t)(((di + 1) * 4 + (dj + 1)) * 4 + (dk + 1)) * 8 + (uint32_t)(t % 5);
                const uint8_t packed = (uint8_t)((di + 1) | (dj + 1) << 2 | (dk + 1) << 4);
                auto& L = lists[p * kStarShapes + shape];
                bool seen = false;
                for (auto& e : L) seen = seen || (e.first == order);
                if (!seen) L.push_back({order, packed});
            }
        }
        StarTables S;
        S.head.assign(2 * kStarShapes, 0);
        for (size_t h = 0; h < lists.size(); ++h) {
            auto& L = lists[h];
            std::sort(L.begin(), L.end());
            if (2 * L.size() > 255) throw std::logic_error("star table: star too large");
            S.head[h] = (uint32_t)S.ent.size() << 8 | (uint32_t)(2 * L.size());
            for (auto& e : L) {
                S.ent.push_back(e.second);
                S.ent.push_back((uint8_t)(e.first & 7u));
            }
        }
        return S;
    }();
    return T;
}

unsigned cube_tiles(const Grid& g, int tile = kCubeTile) {
    const uint64_t cubes = g.view.T / 5;
    return (unsigned)std::max<uint64_t>(1, (cubes + tile - 1) / tile);
}

void alloc_ws(Workspace& W, const Grid& g, const Caps_& c, hipStream_t s) {
    W.cap = c;
    DevMem& m = W.mem;
    const uint64_t V = g.view.V, T = g.view.T;
    W.pos.alloc(m, V);
    W.neg.alloc(m, V);
    if (g.view.regular) {
        W.tile_sums.alloc(m, cube_tiles(g, 256) + 1);  // the smallest tile has the most tiles
        for (int q = 1; q < 3; ++q) W.early[q].alloc(m, c.b[q] + 1);
    } else {
        W.kcnt.alloc(m, T + 1);
        W.isact.alloc(m, T + 1);
    }
    W.csr_off.alloc(m, T + 1);
    W.rank.alloc(m, T + 1);
    W.act.alloc(m, c.active + 1);
    W.act_k.alloc(m, c.active + 1);
    W.act_off.alloc(m, c.active + 1);
    W.rec_loc.alloc(m, c.active + 1);
    W.cnt.alloc(m, 3 * (uint64_t)c.active + 3);
    W.fit.alloc(m, c.fit + 1);
    uint64_t words = 0;
    for (int q = 0; q < 3; ++q) {
        W.list[q].alloc(m, c.b[q] + 1);
        W.lts[q].alloc(m, c.b[q] + 1);
        W.rec_base[q] = (uint32_t)words;
        words += (uint64_t)c.b[q] * W.slot[q];
    }
    if (words >= (1ull << 31)) throw std::runtime_error("record workspace exceeds capacity");
    W.rec.alloc(m, words + 1);
    W.cv.alloc(m, c.active + 1);
    W.cf.alloc(m, c.active + 1);
    W.cfv.alloc(m, c.active + 1);
    W.okey.alloc(m, c.vo + 1);
    W.oa.alloc(m, c.vo + 1);
    W.oj.alloc(m, c.vo + 1);
    for (auto* b : {&W.first, &W.vid, &W.vrank, &W.first_occ, &W.hslot, &W.mo, &W.fail}) b->alloc(m, c.vo + 1);
    W.tile_first.alloc(m, (c.vo + kDedupTile - 1) / kDedupTile + 1);
    {
        const StarTables& ST = star_tables();
        W.star_head.alloc(m, ST.head.size());
        W.star_ent.alloc(m, ST.ent.size());
        RIN_HIP(hipMemcpyAsync(W.star_head.p, ST.head.data(), 4 * ST.head.size(), hipMemcpyHostToDevice, s));
        RIN_HIP(hipMemcpyAsync(W.star_ent.p, ST.ent.data(), ST.ent.size(), hipMemcpyHostToDevice, s));
    }
    W.hbits = 10;
    while (W.hbits < 31 && (1ull << W.hbits) < 2ull * c.vo) ++W.hbits;
    W.hkey.alloc(m, 1ull << W.hbits);
    W.hmx.alloc(m, 1ull << W.hbits);
    RIN_HIP(hipMemsetAsync(W.hkey.p, 0, 8ull << W.hbits, s));
    RIN_HIP(hipMemsetAsync(W.hmx.p, 0, 4ull << W.hbits, s));
    W.hash_dirty = false;
    W.fq_a.alloc(m, c.fo + 1);
    W.fq_vs.alloc(m, c.fo + 1);
    W.fq_l.alloc(m, c.fo + 1);
    W.fq_fn.alloc(m, c.fo + 1);
    W.fq_bnd.alloc(m, c.fo + 1);
    W.fve.alloc(m, c.fv + 1);
    for (auto* b : {&W.ffirst, &W.fmatch}) b->alloc(m, c.fo + 1);
    W.fvm.alloc(m, c.fv + 1);
    W.ftiles.alloc(m, 4 * ((c.fo + kFaceTile - 1) / kFaceTile) + 4);
    W.fsizes.alloc(m, 4);
    W.v_tet.alloc(m, c.vo + 1);
    W.v_j.alloc(m, c.vo + 1);
    W.v_ss.alloc(m, c.vo + 1);
    W.v_sv.alloc(m, 4 * (uint64_t)c.vo + 4);
    W.v_fn.alloc(m, 4 * (uint64_t)c.vo + 4);
    W.xyz.alloc(m, 3 * (uint64_t)c.vo + 3);
    W.f_vstart.alloc(m, c.fo + 1);
    W.f_tfstart.alloc(m, c.fo + 1);
    W.f_verts.alloc(m, c.fv + 1);
    W.f_tf_tet.alloc(m, c.fo + 1);
    W.f_tf_local.alloc(m, c.fo + 1);
    W.f_func.alloc(m, 2 * (uint64_t)c.fo + 2);
    W.ctr.alloc(m, kSpreadBase + SP_N * kSpread * kSpreadStride);
    W.errs.alloc(m, 64);
    const K1Table& K = k1_table();
    const K2Table& K2 = k2_table();
    W.k1.alloc(m, 81);
    W.k1_words.alloc(m, K.words.size() + K2.words.size() + 1);
    W.k1_off.alloc(m, 81 + K2.nslots);
    W.k2pat.alloc(m, 81);
    W.k2cnt.alloc(m, K2.nslots + 1);
    std::vector<uint32_t> off(81 + K2.nslots);
    for (int i = 0; i < 81; ++i) off[i] = K.e[i].rec_off;
    for (uint32_t i = 0; i < K2.nslots; ++i)
        off[81 + i] = K2.off[i] == K2_NONE ? K2_NONE : (uint32_t)K.words.size() + K2.off[i];
    RIN_HIP(hipMemcpyAsync(W.k1.p, K.e, sizeof(K.e), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k1_words.p, K.words.data(), 2 * K.words.size(), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k1_words.p + K.words.size(), K2.words.data(), 2 * K2.words.size(), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k1_off.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k2pat.p, K2.pat, sizeof(K2.pat), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k2cnt.p, K2.cnt.data(), 4 * K2.cnt.size(), hipMemcpyHostToDevice, s));
    // RIN_K2_START=0: the lane-group cuts insert every plane themselves
    const char* ks = std::getenv("RIN_K2_START");
    W.k2_start = !(ks && !std::strcmp(ks, "0"));
    W.k2snap_off.alloc(m, K2.snap_off.size() + 1);
    W.k2snap.alloc(m, K2.snap.size() + 1);
    RIN_HIP(hipMemcpyAsync(W.k2snap_off.p, K2.snap_off.data(), 4 * K2.snap_off.size(), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k2snap.p, K2.snap.data(), K2.snap.size(), hipMemcpyHostToDevice, s));
    const K2Plans& KP = k2_plans();
    W.k2plan_off.alloc(m, KP.off.size() + 1);
    W.k2plan.alloc(m, KP.bytes.size() + 1);
    RIN_HIP(hipMemcpyAsync(W.k2plan_off.p, KP.off.data(), 4 * KP.off.size(), hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(W.k2plan.p, KP.bytes.data(), KP.bytes.size(), hipMemcpyHostToDevice, s));
    // copies on the context's stream (a synchronous copy would go through the
    // legacy stream, which another thread's graph capture forbids); `off`
    // lives until they are done
    RIN_HIP(hipStreamSynchronize(s));
}

// stage 1 (also run eagerly to size the workspace): signs, filter, and
// either the per-tile sums (regular grids) or the per-tet CSR scans; zero:
// clear the counters first (the whole-path graph leaves them clear instead)
// early3, early5: list the k = 3..4 / 5..8 tets for launch_cut_early (regular grids)
void stage_count(Workspace& W, const Grid& g, const Values& val, hipStream_t s, bool zero = true,
                 bool early3 = false, bool early5 = false) {
    const uint64_t V = g.view.V, T = g.view.T;
    const uint32_t F = val.F;
    if (zero) hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, W.ctr.p, (int)C_N, W.errs.p);
    hipLaunchKernelGGL(k_signs_count, dim3(nblk(V, 1024)), dim3(1024), 0, s, (const double*)val.vals.p, V, F, W.pos.p,
                       W.neg.p, W.ctr.p);
    const uint64_t full = (F == 64) ? ~0ull : ((1ull << F) - 1);
    if (g.view.regular) {
        const int tile = W.cube_tile;
        auto count = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(cube_tiles(g, tile)), dim3(tile), 0, s, g.view, (const uint64_t*)W.pos.p,
                               (const uint64_t*)W.neg.p, full, W.tile_sums.p, early3 ? W.early[1].p : nullptr,
                               early5 ? W.early[2].p : nullptr, W.cap.b[1], W.cap.b[2], W.ctr.p);
        };
        if (tile == 256)
            count(k_cube_count<256>);
        else if (tile == 1024)
            count(k_cube_count<1024>);
        else
            count(k_cube_count<512>);
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kTileScanThreads), 0, s, W.tile_sums.p, cube_tiles(g, tile),
                           W.ctr.p);
        return;
    }
    hipLaunchKernelGGL(k_filter_count, dim3(nblk(T)), dim3(256), 0, s, g.view, (const uint64_t*)W.pos.p,
                       (const uint64_t*)W.neg.p, full, W.kcnt.p, W.isact.p, W.ctr.p);
    scan_u32(W.mem, W.kcnt.p, W.csr_off.p, Count(T), W.ctr.p + C_FIT, s);
    scan_u32(W.mem, W.isact.p, W.rank.p, Count(T), W.ctr.p + C_ACTIVE, s);
}

// the k = 2 table lookups' arguments (the bucket list's size and the abort
// flag are the caller's)
K2Args k2_args(Workspace& W, const Grid& g, const Values& val) {
    K2Args K = {};
    K.g = g.view;
    K.vals = val.vals.p;
    K.pos = W.pos.p;
    K.neg = W.neg.p;
    K.full = (val.F == 64) ? ~0ull : ((1ull << val.F) - 1);
    K.act = W.act.p;
    K.csr_off = W.csr_off.p;
    K.list = W.list[0].p;
    K.lts = W.lts[0].p;
    K.fit = W.fit.p;
    K.cap = W.cap.b[0];
    K.pat = W.k2pat.p;
    K.slot_off = W.k1_off.p + 81;
    K.slot_cnt = W.k2cnt.p;
    K.rec_loc = W.rec_loc.p;
    K.cnt = W.cnt.p;
    K.errs = W.errs.p;
    K.rec = W.rec.p;
    return K;
}

template <class C>
void launch_cut(Workspace& W, const Grid& g, const Values& val, int q, int grid, hipStream_t s) {
    CutArgs A;
    A.g = g.view;
    A.vals = val.vals.p;
    A.act = W.act.p;
    A.csr_off = W.csr_off.p;
    A.list = W.list[q].p;
    A.lts = W.lts[q].p;
    A.rec_loc = W.rec_loc.p;
    A.act_k = W.act_k.p;
    A.fit = W.fit.p;
    A.n = W.ctr.p + C_Q2 + q;
    A.abort = W.ctr.p + C_ABORT;
    A.queue = W.ctr.p + kQueueBase + q;
    A.cap = W.cap.b[q];
    A.slot = W.slot[q];
    A.F = val.F;
    A.rec = W.rec.p;
    A.cnt = W.cnt.p;
    A.errs = W.errs.p;
    A.ctr = W.ctr.p;
    A.stat_k = q == 1 ? SP_CUTK3 : (q == 2 ? SP_CUTK5 : -1);
    A.stat_w = q == 1 ? SP_CUTW3 : (q == 2 ? SP_CUTW5 : -1);
    A.k2s = W.k2_start ? K2Snap{W.k2pat.p, W.k2snap_off.p, W.k2snap.p} : K2Snap{nullptr, nullptr, nullptr};
    if (grid == -2) {  // k = 2 table
        K2Args K = k2_args(W, g, val);
        K.n = A.n;
        K.abort = A.abort;
        hipLaunchKernelGGL(k_k2_lookup, dim3(nblk(W.cap.b[q])), dim3(256), 0, s, K);
        return;
    }
    if (grid < 0) {  // wave form: one block per bucket entry
        // two tets per wavefront up to 4 active functions (their complexes
        // rarely fill 32 lanes); the whole wavefront for k = 5..8; -3: four
        // tets per wavefront (half the waves, each group slower)
        constexpr int G = C::MV > 64 ? 64 : 32;
        if constexpr (G == 32) {
            if (grid == -3) {
                hipLaunchKernelGGL((k_cut_wave<C, 16>), dim3((W.cap.b[q] + 3) / 4), dim3(64), 0, s, A);
                return;
            }
        }
        hipLaunchKernelGGL((k_cut_wave<C, G>), dim3((W.cap.b[q] + 64 / G - 1) / (64 / G)), dim3(64), 0, s, A);
        return;
    }
    fill_u32(W.ctr.p + kQueueBase + q, 0, 1, s);
    // the complexes of the thread form live in the workspace (grown outside
    // capture: the autotuner launches every candidate eagerly first)
    const size_t need = LaneCut<C>::stride * 64 * (size_t)grid;
    if (W.ws_thread[q].n < need) W.ws_thread[q].alloc(W.mem, need);
    A.ws = W.ws_thread[q].p;
    hipLaunchKernelGGL(k_cut_bucket<C>, dim3(grid), dim3(64), 0, s, A);
}

// the k = 3..4 (q = 1) or 5..8 (q = 2) cut from the count pass's list,
// beside the fill pass (regular grids, lane-group forms: grid -1, or -3 for
// four tets a wavefront)
void launch_cut_early(Workspace& W, const Grid& g, const Values& val, int q, int grid, hipStream_t s) {
    CutArgs A = {};
    A.g = g.view;
    A.vals = val.vals.p;
    A.cap = W.cap.b[q];
    A.slot = W.slot[q];
    A.F = val.F;
    A.rec = W.rec.p;
    A.cnt = W.cnt.p;
    A.errs = W.errs.p;
    A.ctr = W.ctr.p;
    A.stat_k = q == 1 ? SP_CUTK3 : SP_CUTK5;
    A.stat_w = q == 1 ? SP_CUTW3 : SP_CUTW5;
    A.early = W.early[q].p;
    A.n_early = W.ctr.p + (q == 1 ? C_E3 : C_E5);
    A.n_back = q == 1 ? W.ctr.p + C_E3B : nullptr;
    A.tiles = W.tile_sums.p;
    A.pos = W.pos.p;
    A.neg = W.neg.p;
    A.full = (val.F == 64) ? ~0ull : ((1ull << val.F) - 1);
    A.rec_base = W.rec_base[q];
    A.cap_active = W.cap.active;
    A.tf = q == 1 ? TF_B3 : TF_B5;
    A.k2s = W.k2_start ? K2Snap{W.k2pat.p, W.k2snap_off.p, W.k2snap.p} : K2Snap{nullptr, nullptr, nullptr};
    const uint32_t cap = W.cap.b[q];
    if (q == 2)
        hipLaunchKernelGGL((k_cut_wave<CBig, 64, true>), dim3(cap), dim3(64), 0, s, A);
    else if (grid == -3)
        hipLaunchKernelGGL((k_cut_wave<CMed, 16, true>), dim3((cap + 3) / 4), dim3(64), 0, s, A);
    else
        hipLaunchKernelGGL((k_cut_wave<CMed, 32, true>), dim3((cap + 1) / 2), dim3(64), 0, s, A);
}

void stage_cut(Workspace& W, const Grid& g, const Values& val, const int* grid, hipStream_t s) {
    if (W.cap.b[0]) launch_cut<CK2>(W, g, val, 0, grid[0], s);
    if (W.cap.b[1]) launch_cut<CMed>(W, g, val, 1, grid[1], s);
    if (W.cap.b[2]) launch_cut<CBig>(W, g, val, 2, grid[2], s);
}

void stage_fill(Workspace& W, const Grid& g, const Values& val, hipStream_t s) {
    const uint64_t T = g.view.T;
    const uint32_t F = val.F;
    FillArgs A;
    A.g = g.view;
    A.pos = W.pos.p;
    A.neg = W.neg.p;
    A.full = (F == 64) ? ~0ull : ((1ull << F) - 1);
    A.csr_off = W.csr_off.p;
    A.rank = W.rank.p;
    A.fit = W.fit.p;
    A.act = W.act.p;
    A.act_off = W.act_off.p;
    A.rec_loc = W.rec_loc.p;
    A.cnt = W.cnt.p;
    A.act_k = W.act_k.p;
    A.k1 = W.k1.p;
    for (int q = 0; q < 3; ++q) {
        A.list[q] = W.list[q].p;
        A.lts[q] = W.lts[q].p;
        A.cap[q] = W.cap.b[q];
        A.rec_base[q] = W.rec_base[q];
        A.slot[q] = W.slot[q];
    }
    A.ctr = W.ctr.p;
    A.errs = W.errs.p;
    A.cap_active = W.cap.active;
    A.cap_fit = W.cap.fit;
    A.csr_off_w = W.csr_off.p;
    A.rank_w = W.rank.p;
    if (g.view.regular) {
        const int tile = W.cube_tile;
        const TileSums* ts = W.tile_sums.p;
        if (tile == 256)
            hipLaunchKernelGGL(k_cube_fill<256>, dim3(cube_tiles(g, tile)), dim3(tile), 0, s, A, ts);
        else if (tile == 1024)
            hipLaunchKernelGGL(k_cube_fill<1024>, dim3(cube_tiles(g, tile)), dim3(tile), 0, s, A, ts);
        else
            hipLaunchKernelGGL(k_cube_fill<512>, dim3(cube_tiles(g, tile)), dim3(tile), 0, s, A, ts);
        return;
    }
    hipLaunchKernelGGL(k_fill_batched, dim3(nblk(T)), dim3(256), 0, s, A);
}

// everything after the cut
// the stages after the cut, each captured as its own graph so that the
// stream events between graphs time them separately
enum RestPart {
    REST_SCANS = 1,
    REST_EXTRACT_K1 = 2,
    REST_DEDUP = 4,
    REST_XYZ = 8,
    REST_FACES = 16,
    REST_EXTRACT_REC = 32,
    REST_ALL = 63
};

// grid of a grid-stride kernel whose element count lives on the device:
// bounded, so that capacity-sized launches do not pay for idle workgroups
unsigned cgrid(uint64_t cap) { return std::min(nblk(cap), 2048u); }

void stage_rest(Workspace& W, const Grid& g, const Values& val, hipStream_t s, int part) {
    DevMem& m = W.mem;
    const Caps_& c = W.cap;
    uint32_t* ctr = W.ctr.p;
    const uint32_t* abort = ctr + C_ABORT;
    const Count NA(ctr + C_ACTIVE, c.active, abort);
    if (part & REST_SCANS) {
        const uint32_t* in[3] = {W.cnt.p, W.cnt.p + 1, W.cnt.p + 2};
        uint32_t* out[3] = {W.cv.p, W.cf.p, W.cfv.p};
        uint32_t* tot[3] = {ctr + C_VO, ctr + C_FO, ctr + C_FV};
        scan3_u32(m, in, 3, out, NA, tot, s);
    }
    if (part & (REST_EXTRACT_K1 | REST_EXTRACT_REC)) {
        ExtractArgs E;
        E.g = g.view;
        E.act = W.act.p;
        E.act_off = W.act_off.p;
        E.csr_off = W.csr_off.p;
        E.rec_loc = W.rec_loc.p;
        E.fit = W.fit.p;
        E.rec = W.rec.p;
        E.k1 = W.k1.p;
        E.tab = W.k1_words.p;
        E.tab_off = W.k1_off.p;
        E.n_active = ctr + C_ACTIVE;
        E.cap_active = c.active;
        E.ctr = ctr;
        E.voff = W.cv.p;
        E.foff = W.cf.p;
        E.fvoff = W.cfv.p;
        E.tot = ctr;
        E.cap_vo = c.vo;
        E.cap_fo = c.fo;
        E.cap_fv = c.fv;
        E.okey = W.okey.p;
        E.oa = W.oa.p;
        E.oj = W.oj.p;
        E.fq_a = W.fq_a.p;
        E.fq_l = W.fq_l.p;
        E.fq_fn = W.fq_fn.p;
        E.fq_bnd = W.fq_bnd.p;
        E.fq_vs = W.fq_vs.p;
        E.fve = W.fve.p;
        E.errs = W.errs.p;
        for (int q = 0; q < 3; ++q) {
            E.list[q] = W.list[q].p;
            E.lts[q] = W.lts[q].p;
            E.list_cap[q] = c.b[q];
            E.slot_words[q] = W.slot[q];
            E.fo_table[q] = W.wave_cut[q] ? 1u : 0u;
        }
        E.plan_off = W.k2plan_off.p;
        E.plan = W.k2plan.p;
        E.n_plans = (uint32_t)k2_plans().off.size();
        if ((part & REST_EXTRACT_K1) && (part & REST_EXTRACT_REC)) {
            // up to 16384 workgroups a lane-per-tet role (one lane per active
            // tet at 101^3: 0.375 ms/step against 0.381 with 8192)
            const uint32_t ecap = 16384u;
            const uint32_t nb_rec = std::min<uint32_t>((c.b[0] + 63) / 64, ecap) + 1;
            const uint32_t nb_wave = std::max(1u, c.b[1] + c.b[2]);
            const uint32_t nb_k1 = std::min<uint32_t>((c.active + 63) / 64, ecap) + 1;
            hipLaunchKernelGGL(k_extract_all, dim3(nb_rec + nb_wave + nb_k1), dim3(64), 0, s, E, nb_rec, nb_wave,
                               nb_k1);
        } else {
            if (part & REST_EXTRACT_K1) hipLaunchKernelGGL(k_extract_k1, dim3(cgrid(c.active)), dim3(256), 0, s, E);
            if (part & REST_EXTRACT_REC) {
                hipLaunchKernelGGL(k_extract_records_wave, dim3(std::max(1u, c.b[1] + c.b[2])), dim3(64), 0, s, E);
                hipLaunchKernelGGL(k_extract_records, dim3(cgrid(c.b[0])), dim3(256), 0, s, E);
            }
        }
    }

    const Count NVO(ctr + C_VO, c.vo, abort), NFO(ctr + C_FO, c.fo, abort), NV(ctr + C_VERTS, c.vo, abort),
        NF(ctr + C_FACES, c.fo, abort);
    if (part & REST_DEDUP) {
        // regular grids: star tables, then the hash path for what they leave;
        // explicit meshes: the hash path for every occurrence
        const bool star = g.view.regular != 0;
        const Count NH = star ? Count(ctr + C_DFAIL, c.vo, abort) : NVO;
        const uint32_t* list = star ? W.fail.p : nullptr;
        const unsigned hgrid = star ? std::min(nblk(c.vo), 1024u) : nblk(c.vo);
        const unsigned tgrid = (unsigned)((c.vo + kDedupTile - 1) / kDedupTile);
        if (star) {
            StarArgs S;
            S.res = g.view.res;
            S.okey = W.okey.p;
            S.oa = W.oa.p;
            S.act = W.act.p;
            S.rank = W.rank.p;
            S.voff = W.cv.p;
            S.star_head = W.star_head.p;
            S.star_ent = W.star_ent.p;
            S.first = W.first.p;
            S.mo = W.mo.p;
            S.fail = W.fail.p;
            S.n_fail = ctr + C_DFAIL;
            S.cap_fail = c.vo;
            S.n_active = ctr + C_ACTIVE;
            S.tile_first = W.tile_first.p;
            hipLaunchKernelGGL(k_dedup_star, dim3(tgrid), dim3(256), 0, s, NVO, S);
        }
        if (star) {
            // the fallback and the scan of the star pass's tile counts in one
            // workgroup, then ranks per tile
            hipLaunchKernelGGL(k_dedup_fallback_scan, dim3(1), dim3(1024), 0, s, NH, list, (const uint64_t*)W.okey.p,
                               W.hkey.p, W.hmx.p, W.hbits, W.hslot.p, W.errs.p, W.first.p, W.mo.p, NVO,
                               W.tile_first.p, ctr + C_VERTS);
            hipLaunchKernelGGL(k_dedup_rank, dim3(tgrid), dim3(256), 0, s, NVO, (const uint32_t*)W.first.p,
                               (const uint32_t*)W.tile_first.p, W.vrank.p, W.first_occ.p, NH, list,
                               (const uint32_t*)W.hslot.p, W.hkey.p, W.hmx.p);
        } else {
            hipLaunchKernelGGL(k_hash_insert, dim3(hgrid), dim3(256), 0, s, NH, list, (const uint64_t*)W.okey.p,
                               W.hkey.p, W.hmx.p, W.hbits, W.hslot.p, W.errs.p);
            hipLaunchKernelGGL(k_hash_first, dim3(hgrid), dim3(256), 0, s, NH, list, (const uint32_t*)W.hslot.p,
                               (const uint32_t*)W.hmx.p, W.first.p, W.mo.p);
            scan_u32(m, W.first.p, W.vrank.p, NVO, ctr + C_VERTS, s);
            hipLaunchKernelGGL(k_dedup_assign, dim3(cgrid(c.vo)), dim3(256), 0, s, NVO, (const uint32_t*)W.mo.p,
                               (const uint32_t*)W.first.p, (const uint32_t*)W.vrank.p, W.vid.p, W.first_occ.p, NH,
                               list, (const uint32_t*)W.hslot.p, W.hkey.p, W.hmx.p, true);
        }
    }
    VertFieldArgs VF;
    VF.g = g.view;
    VF.vals = (const double*)val.vals.p;
    VF.act = W.act.p;
    VF.off = W.act_off.p;
    VF.per_rank = true;
    VF.fit = W.fit.p;
    VF.rec_loc = W.rec_loc.p;
    VF.rec = W.rec.p;
    VF.k1_words = W.k1_words.p;
    VF.k1_off = W.k1_off.p;
    VF.first_occ = W.first_occ.p;
    VF.oa = W.oa.p;
    VF.oj = W.oj.p;
    VF.nverts = NV;
    VF.v_tet = W.v_tet.p;
    VF.v_j = W.v_j.p;
    VF.v_ss = W.v_ss.p;
    VF.v_sv = W.v_sv.p;
    VF.v_fn = W.v_fn.p;
    VF.xyz = W.xyz.p;
    // coordinates and faces together: one launch for both first passes
    const bool xyz_with_faces = (part & REST_XYZ) && (part & REST_FACES);
    if ((part & REST_XYZ) && !xyz_with_faces)
        hipLaunchKernelGGL(k_vert_fields_ia, dim3(cgrid(c.vo)), dim3(256), 0, s, VF);
    if (part & REST_FACES) {
        // faces: a boundary face may repeat in the tet across its tet face
        FaceTileArgs FA;
        FA.fq_a = W.fq_a.p;
        FA.fq_vs = W.fq_vs.p;
        FA.fve = W.fve.p;
        FA.mo = W.mo.p;
        FA.vrank = W.vrank.p;
        FA.act = W.act.p;
        FA.rank = W.rank.p;
        FA.foff = W.cf.p;
        FA.nbr = g.nbr.p;
        FA.fq_bnd = W.fq_bnd.p;
        FA.fq_l = W.fq_l.p;
        FA.fq_fn = W.fq_fn.p;
        FA.nfv_total = ctr + C_FV;
        FA.nfo_total = ctr + C_FO;
        FA.nfv_cap = c.fv;
        FA.nfo_cap = c.fo;
        FA.nq = NFO;
        FA.n_active = NA;
        FA.fvm = W.fvm.p;
        FA.flags = W.ffirst.p;
        FA.partner = W.fmatch.p;
        FA.tiles = W.ftiles.p;
        FA.f_vstart = W.f_vstart.p;
        FA.f_verts = W.f_verts.p;
        FA.f_tfstart = W.f_tfstart.p;
        FA.f_tf_tet = W.f_tf_tet.p;
        FA.f_tf_local = W.f_tf_local.p;
        FA.f_func = W.f_func.p;
        FA.sizes = nullptr;
        FA.sizes_out = W.fsizes.p;
        const unsigned fgrid = (unsigned)((c.fo + kFaceTile - 1) / kFaceTile);
        // the coordinates role: up to 8192 workgroups (at 101^3 0.374 ms/step
        // against 0.381 with 2048)
        const unsigned xgrid = std::min(nblk(c.vo), 8192u);
        if (xyz_with_faces)
            hipLaunchKernelGGL(k_faces_and_xyz, dim3(fgrid + xgrid), dim3(256), 0, s, FA, VF, fgrid, xgrid);
        else
            hipLaunchKernelGGL(k_face_first_tiled, dim3(fgrid), dim3(256), 0, s, FA);
        // the emit launch closes the run (counters published and cleared); it
        // reads its sizes from fsizes, kept by the first pass
        FA.sizes = W.fsizes.p;
        hipLaunchKernelGGL(k_face_emit_close, dim3(fgrid + 1), dim3(256), 0, s, FA, ctr, g.view.T, W.csr_off.p,
                           W.errs.p, W.pub_d);
    }
}

// conservative per-tet output bounds by active count (maxima of the cut's
// vertex / face counts, see the capacities of the bucket kernels)
void caps_from_counts(const uint32_t* ctr, Caps_& c) {
    const uint64_t b1 = ctr[C_B1], b2 = ctr[C_B2], b3 = ctr[C_B3], b5 = ctr[C_B5];
    auto clamp = [](uint64_t x) -> uint32_t {
        if (x >= (1ull << 31)) throw std::runtime_error("workspace capacity exceeds 2^31 entries");
        return (uint32_t)x;
    };
    c.active = clamp(ctr[C_ACTIVE] + 64);
    c.fit = clamp(ctr[C_FIT] + 64);
    c.b[0] = clamp(b2 + 64);
    c.b[1] = clamp(b3 + 64);
    c.b[2] = clamp(b5 + 64);
    c.vo = clamp(4 * b1 + 16 * b2 + 40 * b3 + 132 * b5 + 64);
    c.fo = clamp(1 * b1 + 24 * b2 + 64 * b3 + 284 * b5 + 64);
    c.fv = clamp(4 * b1 + 24 * 6 * b2 + 64 * 8 * b3 + 284 * 10 * b5 + 64);
}

bool caps_ok(const Caps_& have, const Caps_& need) {
    return have.active >= need.active && have.fit >= need.fit && have.b[0] >= need.b[0] && have.b[1] >= need.b[1] &&
           have.b[2] >= need.b[2] && have.vo >= need.vo && have.fo >= need.fo && have.fv >= need.fv;
}

float time_ms(hipStream_t s, const std::function<void()>& f) {
    hipEvent_t a, b;
    RIN_HIP(hipEventCreate(&a));
    RIN_HIP(hipEventCreate(&b));
    RIN_HIP(hipEventRecord(a, s));
    f();
    RIN_HIP(hipEventRecord(b, s));
    RIN_HIP(hipEventSynchronize(b));
    float ms = 0;
    RIN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace

int debug_table_record_ia(int k, const double* rows, uint16_t* out, int cap) {
    auto sgn = [](double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); };
    if (k == 1) {
        const K1Table& K = k1_table();
        int pat = 0;
        for (int c = 0, w = 1; c < 4; ++c, w *= 3) pat += (sgn(rows[c]) + 1) * w;
        const K1Entry& E = K.e[pat];
        if (!E.valid) return -1;
        if (E.rec_len > cap) return -2;
        std::memcpy(out, K.words.data() + E.rec_off, 2 * (size_t)E.rec_len);
        return E.rec_len;
    }
    if (k != 2) return -3;
    // the slot arithmetic of k_k2_lookup
    const K2Table& K2 = k2_table();
    int pa = 0, sB[4];
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) {
        pa += (sgn(rows[c]) + 1) * w;
        sB[c] = sgn(rows[4 + c]);
    }
    const K2Pat& K = K2.pat[pa];
    if (!K.valid) return -1;
    uint32_t slot = K.base, w = 81;
    for (int c = 0, w3 = 1; c < 4; ++c, w3 *= 3) slot += (uint32_t)(sB[c] + 1) * w3;
    bool err = false;
    for (int q = 0; q + 4 < K.nv; ++q, w *= 3) {
        const int c1 = K.e[q][0], c2 = K.e[q][1];
        int sg;
        if (sB[c1] * sB[c2] < 0)
            sg = dev::orient_ia_edge(rows[c1], rows[c2], rows[4 + c1], rows[4 + c2], err);
        else
            sg = sB[c1] == sB[c2] ? sB[c1] : (sB[c1] != 0 ? sB[c1] : sB[c2]);
        slot += (uint32_t)(sg + 1) * w;
    }
    if (err || K2.off[slot] == K2_NONE) return -1;
    const uint16_t* r = K2.words.data() + K2.off[slot];
    const int len = r[6];
    if (len > cap) return -2;
    std::memcpy(out, r, 2 * (size_t)len);
    return len;
}

namespace {

std::string cut_label(const char* base, const Workspace& w, const Plan& P, int q) {
    return std::string(base) + (w.grid[q] == -2 ? " (sign-vector table)"
                                : w.grid[q] == -3 ? " (lane group of 16 per tet)"
                                : w.grid[q] < 0 ? " (wave per tet)"
                                                : " (thread per tet, " + std::to_string(P.blocks_per_cu[q]) +
                                                      " blocks/CU)");
}

// the whole path as one graph: the k = 5..8 cut (few tets, long serial
// insertions) and the k = 2 table lookups on a branch beside the k = 3..4 cut
// (the cut kernels share no outputs; on regular grids the k = 3..4 and 5..8
// cuts start beside the fill pass, from the count pass's lists); the three
// extraction passes write
// disjoint rows and run as one launch, and so do coordinates and the first
// face pass
void record_whole(Workspace& w, const Grid& g, const Values& val, hipStream_t s) {
    if (!w.side) {
        RIN_HIP(hipStreamCreateWithFlags(&w.side, hipStreamNonBlocking));
        RIN_HIP(hipEventCreateWithFlags(&w.fork, hipEventDisableTiming));
        RIN_HIP(hipEventCreateWithFlags(&w.join, hipEventDisableTiming));
        RIN_HIP(hipStreamCreateWithFlags(&w.side2, hipStreamNonBlocking));
        RIN_HIP(hipEventCreateWithFlags(&w.join2, hipEventDisableTiming));
    }
    hipGraph_t gr = nullptr;
    RIN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    // regular grids with lane-group cuts start them right after the counts,
    // beside the fill pass; the k = 2 lookups follow the fill pass
    const bool e3 = g.view.regular && w.cap.b[1] && w.grid[1] < 0;
    const bool e5 = g.view.regular && w.cap.b[2] && w.grid[2] < 0;
    stage_count(w, g, val, s, false, e3, e5);
    if (e3 || e5) {
        // the k = 3..4 cut, captured first on the capturing stream, starts
        // a few microseconds before the other branches: its wavefronts (all
        // resident at once when they find the CUs free) take their place
        // before the fill pass's workgroups fill the CUs; the k = 5..8 cut
        // (tens of wavefronts) and the fill pass follow on side streams
        RIN_HIP(hipEventRecord(w.fork, s));
        if (e3) launch_cut_early(w, g, val, 1, w.grid[1], s);
        if (e5) {
            RIN_HIP(hipStreamWaitEvent(w.side, w.fork, 0));
            launch_cut_early(w, g, val, 2, w.grid[2], w.side);
            RIN_HIP(hipEventRecord(w.join, w.side));
        }
        RIN_HIP(hipStreamWaitEvent(w.side2, w.fork, 0));
        stage_fill(w, g, val, w.side2);
        if (!e5 && w.cap.b[2]) launch_cut<CBig>(w, g, val, 2, w.grid[2], w.side2);
        if (!e3 && w.cap.b[1]) launch_cut<CMed>(w, g, val, 1, w.grid[1], w.side2);
        if (w.cap.b[0]) launch_cut<CK2>(w, g, val, 0, w.grid[0], w.side2);
        RIN_HIP(hipEventRecord(w.join2, w.side2));
        if (e5) RIN_HIP(hipStreamWaitEvent(s, w.join, 0));
        RIN_HIP(hipStreamWaitEvent(s, w.join2, 0));
    } else {
        stage_fill(w, g, val, s);
        RIN_HIP(hipEventRecord(w.fork, s));
        RIN_HIP(hipStreamWaitEvent(w.side, w.fork, 0));
        if (w.cap.b[2]) launch_cut<CBig>(w, g, val, 2, w.grid[2], w.side);
        if (w.cap.b[0]) launch_cut<CK2>(w, g, val, 0, w.grid[0], w.side);
        RIN_HIP(hipEventRecord(w.join, w.side));
        if (w.cap.b[1]) launch_cut<CMed>(w, g, val, 1, w.grid[1], s);
        RIN_HIP(hipStreamWaitEvent(s, w.join, 0));
    }
    stage_rest(w, g, val, s, REST_SCANS);
    stage_rest(w, g, val, s, REST_EXTRACT_K1 | REST_EXTRACT_REC);
    stage_rest(w, g, val, s, REST_DEDUP);
    stage_rest(w, g, val, s, REST_XYZ | REST_FACES);
    RIN_HIP(hipStreamEndCapture(s, &gr));
    bool updated = false;
    if (w.whole) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        updated = hipGraphExecUpdate(w.whole, gr, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(w.whole);
            w.whole = nullptr;
        }
    }
    if (!updated) RIN_HIP(hipGraphInstantiate(&w.whole, gr, nullptr, nullptr, 0));
    if (w.whole_graph) (void)hipGraphDestroy(w.whole_graph);
    w.whole_graph = gr;
    while (w.ev.size() < 2) {
        hipEvent_t e;
        RIN_HIP(hipEventCreate(&e));
        w.ev.push_back(e);
    }
    w.bound_whole = val.vals.p;
}

// the same path as one graph per stage, for per-stage device times
void record_parts(Workspace& w, const Grid& g, const Values& val, const Plan& P, hipStream_t s) {
    w.drop_parts();
    auto capture = [&](const std::string& label, const std::function<void()>& body) {
        hipGraph_t gr = nullptr;
        hipGraphExec_t ex = nullptr;
        RIN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        body();
        RIN_HIP(hipStreamEndCapture(s, &gr));
        RIN_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
        w.graphs.push_back({label, gr});
        w.parts.push_back({label, ex});
    };
    capture("signs + filter + CSR scans", [&] { stage_count(w, g, val, s); });
    capture("CSR fill + buckets + k=1 table", [&] { stage_fill(w, g, val, s); });
    if (w.cap.b[0]) capture(cut_label("cut k=2", w, P, 0), [&] { launch_cut<CK2>(w, g, val, 0, w.grid[0], s); });
    if (w.cap.b[1]) capture(cut_label("cut k=3..4", w, P, 1), [&] { launch_cut<CMed>(w, g, val, 1, w.grid[1], s); });
    if (w.cap.b[2]) capture(cut_label("cut k=5..8", w, P, 2), [&] { launch_cut<CBig>(w, g, val, 2, w.grid[2], s); });
    capture("extract: per-tet output scans", [&] { stage_rest(w, g, val, s, REST_SCANS); });
    capture("extract k=1 occurrences (k_extract_k1)", [&] { stage_rest(w, g, val, s, REST_EXTRACT_K1); });
    capture("extract record occurrences (k_extract_records)", [&] { stage_rest(w, g, val, s, REST_EXTRACT_REC); });
    capture("dedup verts (star tables)", [&] { stage_rest(w, g, val, s, REST_DEDUP); });
    capture("compute xyz (k_vert_fields_ia)", [&] { stage_rest(w, g, val, s, REST_XYZ); });
    capture("faces", [&] { stage_rest(w, g, val, s, REST_FACES); });
    while (w.ev.size() < w.parts.size() + 1) {
        hipEvent_t e;
        RIN_HIP(hipEventCreate(&e));
        w.ev.push_back(e);
    }
    w.bound_parts = val.vals.p;
}

}  // namespace

void run_batched(DevMem& ctx_mem, const Grid& g, const Values& val, uint32_t flags, RunOut& out, hipStream_t s, int n_cu) {
    if (flags & (RUN_MI | RUN_CUT_ONLY | RUN_REVERSE))
        throw std::runtime_error("batched form is not available for this mode (use the plain form)");
    // the batched form is built on the k = 1 and k = 2 tables
    if (flags & RUN_NO_LOOKUP) throw NeedsPlainForm("lookups off: the general cut runs for every tet");
    const bool staged = (flags & RUN_STAGE_TIMES) != 0;
    const uint64_t V = g.view.V, T = g.view.T;
    if (val.V != V) throw std::invalid_argument("values do not match the grid");
    ensure_nbr(ctx_mem, g, s);
    if (val.grid_id != g.id) throw std::invalid_argument("values were made on another grid");
    std::shared_ptr<Plan> planp = g.plans->get((uint64_t)val.F);
    Plan& P = *planp;
    std::lock_guard<std::mutex> lock(P.mu);
    std::shared_ptr<Workspace> W;
    for (auto& w : P.ws)
        if (!w->in_use) {
            W = w;
            break;
        }
    if (!W) {
        W = std::make_shared<Workspace>();
        P.ws.push_back(W);
    }
    Workspace& w = *W;

    if (!w.pub_h) {
        RIN_HIP(hipHostMalloc((void**)&w.pub_h, kPubWords * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
        RIN_HIP(hipHostGetDevicePointer((void**)&w.pub_d, w.pub_h, 0));
    }
    uint32_t hctr[32] = {};
    uint32_t herr = 0;
    // the run's closing kernel (the faces' tile scan) has written the summary;
    // after an eager pass that ends before it, `publish` launches the summary
    // kernel itself
    auto read_back = [&](bool publish) {
        if (publish) hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, w.ctr.p, w.errs.p, w.pub_d);
        RIN_HIP(hipGetLastError());
        RIN_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < C_N; ++i) hctr[i] = w.pub_h[i];
        herr = w.pub_h[32];
    };
    auto errors = [&] { return herr; };
    bool ran = false;
    for (int attempt = 0; attempt < 3 && !ran; ++attempt) {
        if (!w.exec) {
            w.ctr_clear = false;  // the eager passes below use the counters
            // size the workspace from an eager counting pass
            if (!w.ctr.p) {
                Caps_ tiny;
                tiny.active = tiny.fit = 64;
                for (int q = 0; q < 3; ++q) tiny.b[q] = 64;
                tiny.vo = tiny.fo = tiny.fv = 64;
                alloc_ws(w, g, tiny, s);
            }
            stage_count(w, g, val, s);
            read_back(true);
            if (hctr[C_BX]) throw NeedsPlainForm("a tet has more than 8 active functions");
            Caps_ need;
            caps_from_counts(hctr, need);
            if (!caps_ok(w.cap, need)) {
                // grow with head-room, keeping the k=1 table
                Caps_ grow = need;
                grow.active = need.active + need.active / 8;
                grow.fit = need.fit + need.fit / 8;
                for (int q = 0; q < 3; ++q) grow.b[q] = need.b[q] + need.b[q] / 8;
                grow.vo = need.vo + need.vo / 8;
                grow.fo = need.fo + need.fo / 8;
                grow.fv = need.fv + need.fv / 8;
                alloc_ws(w, g, grow, s);
                stage_count(w, g, val, s);
            }
            stage_fill(w, g, val, s);
            // a one-off run takes the usual choice without timing candidates
            if (P.blocks_per_cu[0] == 0 && (flags & RUN_FIXED_FORMS) && !std::getenv("RIN_CUT_FORM")) {
                P.blocks_per_cu[0] = -2;
                P.blocks_per_cu[1] = P.blocks_per_cu[2] = -1;
            }
            // autotune the cube tile of the count and fill passes once per plan
            // (regular grids): every tile size gives the same lists, so the
            // last pass timed leaves them as they must be
            if (P.cube_tile == 0) {
                P.cube_tile = kCubeTile;
                // RIN_CUBE_TILE=256|512|1024 takes that tile without timing
                const char* ct = std::getenv("RIN_CUBE_TILE");
                const int forced = ct ? std::atoi(ct) : 0;
                if (forced == 256 || forced == 512 || forced == 1024)
                    P.cube_tile = forced;
                else if (g.view.regular && !(flags & RUN_FIXED_FORMS)) {
                    float best = 1e30f;
                    for (int tile : {256, 512, 1024}) {
                        w.cube_tile = tile;
                        float ms = 0;
                        // the count pass as the whole path runs it with lane-group
                        // cuts: listing the k = 3..8 tets (the lists go unused here)
                        for (int rep = 0; rep < 2; ++rep)  // the second pass, caches warm
                            ms = time_ms(s, [&] {
                                stage_count(w, g, val, s, true, w.cap.b[1] != 0, w.cap.b[2] != 0);
                                stage_fill(w, g, val, s);
                            });
                        if (ms < best) {
                            best = ms;
                            P.cube_tile = tile;
                        }
                    }
                }
            }
            w.cube_tile = P.cube_tile;
            // autotune the bucket cut grids once per plan
            if (P.blocks_per_cu[0] == 0) {
                RIN_HIP(hipStreamSynchronize(s));
                // thread form at 1..16 blocks per CU, or the wave form (-1);
                // RIN_CUT_FORM=wave|thread restricts the choice
                // k = 2 also by table (-2); RIN_CUT_FORM=table forces that
                const int cand[8] = {1, 2, 4, 8, 16, -1, -3, -2};
                const char* form = std::getenv("RIN_CUT_FORM");
                // RIN_CUT_FORM=wave16: the wave form with four tets per
                // wavefront for k = 3..4
                const bool only_wave16 = form && !std::strcmp(form, "wave16");
                const bool only_wave = form && (!std::strcmp(form, "wave") || only_wave16);
                const bool only_thread = form && !std::strcmp(form, "thread");
                const bool only_table = form && !std::strcmp(form, "table");
                for (int q = 0; q < 3; ++q) {
                    float best = 1e30f;
                    int bestc = only_wave16 && q == 1 ? -3
                                : only_wave           ? -1
                                                      : (only_table && q == 0 ? -2 : (q == 2 ? 1 : 4));
                    for (int ci = 0; ci < 8; ++ci) {
                        const int c = cand[ci];
                        if (c == -2 && (q != 0 || only_wave || only_thread)) continue;
                        if (c == -3 && q != 1) continue;
                        if (only_wave16 && q == 1 && c != -3) continue;
                        if (only_table && q == 0 && c != -2) continue;
                        if ((only_wave && c >= 0) || (only_thread && c < 0)) continue;
                        // thread forms keep one complex per lane in the
                        // workspace: the larger complexes get smaller grids,
                        // and k = 5..8 (about 9 KB a lane) only the wave form
                        if (c > (q == 2 ? 0 : (q == 1 ? 4 : 8))) continue;
                        const int grid = c < 0 ? c : std::max(1, n_cu * c);
                        float ms = time_ms(s, [&] {
                            if (q == 0 && w.cap.b[0]) launch_cut<CK2>(w, g, val, 0, grid, s);
                            if (q == 1 && w.cap.b[1]) launch_cut<CMed>(w, g, val, 1, grid, s);
                            if (q == 2 && w.cap.b[2]) launch_cut<CBig>(w, g, val, 2, grid, s);
                        });
                        if (ms < best) {
                            best = ms;
                            bestc = cand[ci];
                        }
                    }
                    P.blocks_per_cu[q] = bestc;
                }
            }
            for (int q = 0; q < 3; ++q) {
                w.grid[q] = P.blocks_per_cu[q] < 0 ? P.blocks_per_cu[q] : std::max(1, n_cu * P.blocks_per_cu[q]);
                w.wave_cut[q] = w.grid[q] == -1 || w.grid[q] == -3;
            }
            RIN_HIP(hipStreamSynchronize(s));
            w.exec = true;
        }
        // (re-)record the graphs for these values: a new workspace capacity
        // drops them; new values of the same size re-record the whole graph
        // and update the instantiated one in place
        if (w.bound_whole != val.vals.p) record_whole(w, g, val, s);
        if (staged && w.bound_parts != val.vals.p) record_parts(w, g, val, P, s);
        if (w.hash_dirty) {
            RIN_HIP(hipMemsetAsync(w.hkey.p, 0, 8ull << w.hbits, s));
            RIN_HIP(hipMemsetAsync(w.hmx.p, 0, 4ull << w.hbits, s));
            w.hash_dirty = false;
        }
        if (!w.ctr_clear) hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, w.ctr.p, (int)C_N, w.errs.p);
        w.ctr_clear = false;
        RIN_HIP(hipEventRecord(w.ev[0], s));
        if (staged) {
            for (size_t i = 0; i < w.parts.size(); ++i) {
                RIN_HIP(hipGraphLaunch(w.parts[i].second, s));
                RIN_HIP(hipEventRecord(w.ev[i + 1], s));
            }
        } else {
            RIN_HIP(hipGraphLaunch(w.whole, s));
            RIN_HIP(hipEventRecord(w.ev[1], s));
        }
        read_back(false);
        w.ctr_clear = true;  // the run's closing kernel cleared them
        const uint32_t e = errors();
        if (e) w.hash_dirty = true;
        if (e & ERR_OVERFLOW) {
            w.drop_graphs();
            w.exec = false;
            continue;
        }
        if (e) {
            out.err = e;
            throw NeedsPlainForm("a complex exceeds the bucket capacities or the predicate range");
        }
        ran = true;
    }
    if (!ran) throw std::runtime_error("batched workspace did not converge");
    w.in_use = true;
    if (staged) {
        for (size_t i = 0; i < w.parts.size(); ++i) {
            float ms = 0;
            RIN_HIP(hipEventElapsedTime(&ms, w.ev[i], w.ev[i + 1]));
            out.times.ms.push_back({w.parts[i].first, ms});
        }
    } else {
        float ms = 0;
        RIN_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
        out.times.ms.push_back({"batched path (one graph)", ms});
    }
    // result views into the workspace
    out.mi = false;
    out.flags = flags;
    out.T = T;
    out.F = val.F;
    out.num_degenerate = hctr[C_DEGEN];
    out.n_active = out.num_intersecting = hctr[C_ACTIVE];
    out.n_fit = hctr[C_FIT];
    out.nb[0] = hctr[C_B1];
    out.nb[1] = hctr[C_B2];
    out.nb[2] = hctr[C_B3] + hctr[C_B5];
    out.n_verts = hctr[C_VERTS];
    out.n_faces = hctr[C_FACES];
    out.n_face_verts = hctr[C_FACEVERTS];
    out.n_tet_faces = hctr[C_TETFACES];
    out.n_vert_occ = hctr[C_VO];
    out.n_face_occ = hctr[C_FO];
    out.n_face_occ_verts = hctr[C_FV];
    out.n_cut34 = hctr[C_B3];
    out.n_cut58 = hctr[C_B5];
    out.n_rec_words = hctr[C_RECW];
    out.cut_k_sum[0] = hctr[C_CUTK3];
    out.cut_k_sum[1] = hctr[C_CUTK5];
    out.cut_words[0] = hctr[C_CUTW3];
    out.cut_words[1] = hctr[C_CUTW5];
    auto& vw = out.view;
    vw.fit = w.fit.p;
    vw.csr_off = w.csr_off.p;
    vw.act = w.act.p;
    vw.rank_of_tet = w.rank.p;
    vw.rec = w.rec.p;
    vw.rec_loc = w.rec_loc.p;
    vw.k1_rec = w.k1_words.p;
    vw.k1_rec_off = w.k1_off.p;
    vw.v_tet = w.v_tet.p;
    vw.v_j = w.v_j.p;
    vw.v_ss = w.v_ss.p;
    vw.v_sv = w.v_sv.p;
    vw.v_fn = w.v_fn.p;
    vw.xyz = w.xyz.p;
    vw.f_vstart = w.f_vstart.p;
    vw.f_verts = w.f_verts.p;
    vw.f_tfstart = w.f_tfstart.p;
    vw.f_tf_tet = w.f_tf_tet.p;
    vw.f_tf_local = w.f_tf_local.p;
    vw.f_func = w.f_func.p;
    std::weak_ptr<Plan> wp = planp;
    Workspace* raw = W.get();
    out.hold = std::shared_ptr<void>(nullptr, [W, wp, raw](void*) {
        if (auto p = wp.lock()) {
            std::lock_guard<std::mutex> l(p->mu);
            raw->in_use = false;
        }
    });
}

}  // namespace rin
