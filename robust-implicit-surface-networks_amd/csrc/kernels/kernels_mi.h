// Material-interface extraction on the device (extract_MI_mesh,
// src/extract_mesh.cpp:569-986, and compute_MI_vert_xyz, :1541-1637).
//
// Vertex numbering is the reference's first-creation order.  Per active tet,
// in tet order, the creation events are:
//   phase 0: every local vertex on an interface face (positive label > 3),
//            in local order: tet corners, tet-edge points (edge, two
//            materials), tet-face points (face, three materials) are shared
//            with other tets by key; interior points are always new;
//   phase 1: the corners of the tet-boundary faces this tet emits (below), in
//            face order and cycle order, created if not yet present.
// Each event becomes an occurrence with a 64-bit key; a stable sort by key and
// first-occurrence ranks over occurrence order give the ids (as for the
// implicit arrangement).
//
// A tet-boundary face lies on the tet face shared with the neighbouring tet.
// The reference pairs the two copies through a map keyed by (min, second-min,
// max) vertex id (:833-951): the first copy (lower tet) is stored, the second
// becomes an interface face when the material sets on the two sides share no
// material, with func_index (material of the first copy, material of the
// second).  The device finds the copy directly in the neighbour's record, so
// faces need no sort: face ids are a prefix sum in (tet, local face) order.
#pragma once
#include "kernels_ia.h"

namespace rin {
namespace {

constexpr uint32_t NO_NBR = 0xFFFFFFFFu;

__device__ inline uint64_t key_mi_edge(uint64_t e, uint32_t m1, uint32_t m2) {
    return ((((e << 6) | m1) << 6 | m2) << 2) | 1;
}
__device__ inline uint64_t key_mi_face(uint64_t fc, uint32_t m1, uint32_t m2, uint32_t m3) {
    return (((((fc << 6) | m1) << 6 | m2) << 6 | m3) << 2) | 2;
}

// ------------------------------------------------------------------ tet adjacency

// (global face id, 4 t + local face) of every tet face
__global__ void k_face_slots(GridView g, uint64_t* __restrict__ key, uint32_t* __restrict__ slot) {
    GRID_STRIDE(t, g.T) {
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        for (int c = 0; c < 4; ++c) {
            key[4 * t + c] = dev::face_id(g, t, v, c);
            slot[4 * t + c] = (uint32_t)(4 * t + c);
        }
    }
}
// equal keys are adjacent after the sort; a face is shared by at most two tets
__global__ void k_pair_faces(uint64_t n, const uint64_t* __restrict__ key, const uint32_t* __restrict__ slot,
                             uint32_t* __restrict__ nbr) {
    GRID_STRIDE(p, n) {
        if (p + 1 < n && key[p] == key[p + 1]) {
            nbr[slot[p]] = slot[p + 1];
            nbr[slot[p + 1]] = slot[p];
        }
    }
}

// ------------------------------------------------------------------ 2-material table
//
// With two active materials a and b (a < b, CSR order) the only
// classification the cut makes is material b against material a at the four
// tet corners (exact comparisons of the corner values), so the complex and
// its record depend on those four comparisons alone: one record per pattern
// sum_c s_c 3^c (s_c = 0: b < a, 1: tie, 2: b > a), built once by the cut
// itself (the reference's 2-material lookup table plays the same part,
// src/material_interface.cpp:320-328).

struct MI2Entry {
    uint16_t off, len;   // record words in the table's word array
    uint32_t cnt[3];     // interface vertices, faces, face corners
    uint32_t cut[3];     // vertices, faces, cells of the complex
};

// 3 materials a < b < c: the complex of a and b is one of the 81 patterns
// above; inserting c depends only on that complex and c's sign against each
// of its vertices (the only classifications the cut makes): at a corner an
// exact comparison, at a vertex where a and b tie on a tet edge one exact
// 3 x 3 predicate.  Every sign vector is cut once on the host.  Per pattern:
// the first slot of its sign vectors and the labels of its vertices.
// (MI3Pat and MI3_NONE: kernels_ia.h)

// the tables on the device: 2-material records by pattern; 3-material
// patterns (staged in LDS by the kernel), slots (word offset or MI3_NONE),
// packed counts per slot (interface vertices | faces << 8 | corners << 16,
// and vertices | faces << 8 | cells << 16 of the complex) and records
struct MITablesDev {
    const MI2Entry* e2 = nullptr;
    const uint16_t* w2 = nullptr;
    const MI3Pat* pat3 = nullptr;
    const uint32_t *off3 = nullptr, *cnt3 = nullptr, *cut3 = nullptr;
    const uint16_t* w3 = nullptr;
    const uint32_t* snap_off3 = nullptr;  // complexes after three materials (MI3Snap)
    const uint8_t* snap3 = nullptr;
};

inline MITablesDev mi_tables_view(const MITablesRaw& r) {
    MITablesDev d;
    d.e2 = (const MI2Entry*)r.e2;
    d.w2 = (const uint16_t*)r.w2;
    d.pat3 = (const MI3Pat*)r.pat3;
    d.off3 = (const uint32_t*)r.off3;
    d.cnt3 = (const uint32_t*)r.cnt3;
    d.cut3 = (const uint32_t*)r.cut3;
    d.w3 = (const uint16_t*)r.w3;
    d.snap_off3 = (const uint32_t*)r.snap_off3;
    d.snap3 = (const uint8_t*)r.snap3;
    return d;
}

// Active tets with two or three materials by table (list: their active
// ranks; the count at n_dev when given, else n).  A tet whose sign vector has
// no table entry, or whose record outgrows its slot, is marked for the
// general cut (CNT_RETRY).  rec_loc (optional): the record words hold the two
// tables' words at base2 and base3, and the tet's record location becomes the
// table record's instead of a copy into its slot.
__global__ void __launch_bounds__(256) k_mi_table(GridView g, const double* __restrict__ vals,
                                                  const uint32_t* __restrict__ act,
                                                  const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit,
                                                  const uint32_t* __restrict__ list, uint64_t n,
                                                  const uint32_t* __restrict__ n_dev, uint32_t flags,
                                                  const uint32_t* __restrict__ rec_off, uint16_t* __restrict__ rec,
                                                  uint32_t* __restrict__ cnt, uint32_t* __restrict__ errs,
                                                  MITablesDev T, uint32_t* __restrict__ rec_loc, uint32_t base2,
                                                  uint32_t base3) {
    __shared__ MI3Pat pats[81];
    for (int i = threadIdx.x; i < 81 * (int)sizeof(MI3Pat) / 4; i += blockDim.x)
        reinterpret_cast<uint32_t*>(pats)[i] = reinterpret_cast<const uint32_t*>(T.pat3)[i];
    __syncthreads();
    if (n_dev) n = *n_dev < n ? *n_dev : n;
    const uint64_t V = g.V;
    GRID_STRIDE(i, n) {
        const uint32_t a = list[i];
        const uint32_t t = act[a];
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        const int k = (int)(csr_off[t + 1] - s0);
        double P[3][4];
        for (int q = 0; q < k && q < 3; ++q) {
            const double* f = vals + (uint64_t)fit[s0 + q] * V;
            for (int c = 0; c < 4; ++c) P[q][c] = f[v[c]];
        }
        int pat = 0;
        for (int c = 0, w = 1; c < 4; ++c, w *= 3) pat += (P[1][c] > P[0][c] ? 2 : (P[1][c] < P[0][c] ? 0 : 1)) * w;
        const uint16_t* src = nullptr;
        uint32_t len = 0, c3[3], at_table = 0;
        if (k == 2) {
            const MI2Entry E = T.e2[pat];
            src = T.w2 + E.off;
            at_table = base2 + E.off;
            len = E.len;
            const uint32_t* cc = (flags & RUN_CUT_ONLY) ? E.cut : E.cnt;
            c3[0] = cc[0];
            c3[1] = cc[1];
            c3[2] = cc[2];
        } else {
            const MI3Pat& K = pats[pat];
            bool err = !K.valid || k != 3;
            uint32_t slot = K.base, w = 1;
            for (int q = 0; q < K.nv && !err; ++q, w *= 3) {
                const int lab[4] = {K.lab[q][0], K.lab[q][1], K.lab[q][2], K.lab[q][3]};
                slot += (uint32_t)(dev::orient_mi(P, lab, P[2], err) + 1) * w;
            }
            const uint32_t off = err ? MI3_NONE : T.off3[slot];
            if (off != MI3_NONE) {
                src = T.w3 + off;
                at_table = base3 + off;
                len = src[6];
                const uint32_t pk = (flags & RUN_CUT_ONLY) ? T.cut3[slot] : T.cnt3[slot];
                c3[0] = pk & 0xFF;
                c3[1] = (pk >> 8) & 0xFF;
                c3[2] = pk >> 16;
            }
        }
        const uint32_t at = rec_off[a];
        if (!src || (!rec_loc && len > rec_off[a + 1] - at)) {
            flag_error(errs, ERR_RETRY);
            cnt[3 * (uint64_t)a] = CNT_RETRY;
            cnt[3 * (uint64_t)a + 1] = cnt[3 * (uint64_t)a + 2] = 0;
            rec[at] = 0;
            continue;
        }
        if (rec_loc)
            rec_loc[a] = at_table;
        else
            for (uint32_t q = 0; q < len; ++q) rec[at + q] = src[q];
        for (int q = 0; q < 3; ++q) cnt[3 * (uint64_t)a + q] = c3[q];
    }
}

// ------------------------------------------------------------------ per-tet helpers

struct MITet {
    Rec R;
    uint32_t t, s0;
    uint32_t v[4];
    __device__ MITet(const GridView& g, const uint16_t* rec, uint32_t tet, uint32_t csr0) : R(rec), t(tet), s0(csr0) {
        dev::tet_corners(g, t, v);
    }
    __device__ const uint16_t* groups() const { return R.r + R.r[6] - R.L; }
    // corner c by selects (an indexed read would move v to scratch memory)
    __device__ uint32_t vc(int c) const { return c == 0 ? v[0] : (c == 1 ? v[1] : (c == 2 ? v[2] : v[3])); }
};

// global ids of the tet edge between corners ca < cb and of the tet face
// opposite corner c, from the corners in registers (as dev::edge_id and
// dev::face_id)
__device__ inline uint64_t mi_edge_id(const GridView& g, const MITet& M, int ca, int cb) {
    if (!g.regular) return g.tedge[6 * (uint64_t)M.t + dev::local_edge(ca, cb)];
    const uint32_t x = M.vc(ca), y = M.vc(cb);
    const uint32_t a = x < y ? x : y, b = x < y ? y : x;
    return (uint64_t)a * 27 + dev::lattice_dir(g.res, a, b);
}
__device__ inline uint64_t mi_face_id(const GridView& g, const MITet& M, int c) {
    if (!g.regular) return g.tface[4 * (uint64_t)M.t + c];
    uint32_t x = M.vc(c == 0 ? 1 : 0), y = M.vc(c <= 1 ? 2 : 1), z = M.vc(c <= 2 ? 3 : 2);
    uint32_t u;
    if (y < x) { u = x; x = y; y = u; }
    if (z < y) { u = y; y = z; z = u; }
    if (y < x) { u = x; x = y; y = u; }
    return ((uint64_t)x * 27 + dev::lattice_dir(g.res, x, y)) * 27 + dev::lattice_dir(g.res, x, z);
}

// key of local vertex j (labels sorted ascending: tet faces first); nb = number
// of tet-face labels.  Bit logic instead of indexed local arrays keeps the
// labels and corners in registers.
__device__ inline uint64_t mi_vkey(const GridView& g, const MITet& M, const uint8_t* __restrict__ fit, uint64_t a,
                                   int j, int& nb) {
    const uint16_t* L4 = M.R.vert(j);
    const uint32_t l0 = L4[0], l1 = L4[1], l2 = L4[2], l3 = L4[3];
    unsigned bm = 0;
    if (l0 <= 3) bm |= 1u << l0;
    if (l1 <= 3) bm |= 1u << l1;
    if (l2 <= 3) bm |= 1u << l2;
    if (l3 <= 3) bm |= 1u << l3;
    nb = __popc(bm);
    auto mat = [&](uint32_t l) { return (uint32_t)fit[M.s0 + l - 4]; };
    if (nb == 3) return key_corner(M.vc(__builtin_ctz(~bm & 15u)));
    if (nb == 2) {
        const unsigned m = ~bm & 15u;
        const int c0 = __builtin_ctz(m), c1 = __builtin_ctz(m & (m - 1));
        return key_mi_edge(mi_edge_id(g, M, c0, c1), mat(l2), mat(l3));
    }
    if (nb == 1) return key_mi_face(mi_face_id(g, M, (int)l0), mat(l1), mat(l2), mat(l3));
    return key_inner(a, j);
}

// Material set on the inner side of a tet-boundary face: its negative
// label's material, or every material of that label's coincidence group (in
// label order).  Enumerated, not stored: an indexed local array would live in
// scratch memory.
template <class Fn>
__device__ inline void mi_side_each(const MITet& M, const uint8_t* __restrict__ fit, int neg, Fn&& fn) {
    if (M.R.dup) {
        const uint16_t* grp = M.groups();
        int n = 0;
        for (int l = 4; l < M.R.L; ++l) n += grp[l] == grp[neg];
        if (n > 1) {
            for (int l = 4; l < M.R.L; ++l)
                if (grp[l] == grp[neg]) fn((uint8_t)fit[M.s0 + l - 4]);
            return;
        }
    }
    fn((uint8_t)fit[M.s0 + neg - 4]);
}

// (min, second-min, max) of a face's vertex keys
__device__ inline void mi_face_triple(const GridView& g, const MITet& M, const uint8_t* __restrict__ fit, uint64_t a,
                                      const uint16_t* F, uint64_t k3[3]) {
    const int n = F[3];
    uint64_t mn = ~0ull, sm = ~0ull, mx = 0;
    for (int q = 0; q < n; ++q) {
        int nb;
        const uint64_t k = mi_vkey(g, M, fit, a, F[4 + q], nb);
        if (k < mn) {
            sm = mn;
            mn = k;
        } else if (k < sm) {
            sm = k;
        }
        if (k > mx) mx = k;
    }
    k3[0] = mn;
    k3[1] = sm;
    k3[2] = mx;
}

struct MIArgs {
    GridView g;
    const uint32_t *act, *csr_off, *rank, *rec_loc, *nbr;
    const uint8_t* fit;
    const uint16_t* rec;
    Count n_active;
};

// whether tet-boundary face F of active tet a becomes an interface face
// here (the second copy); `other` is the first copy's material
__device__ inline bool mi_bface_emits(const MIArgs& A, const MITet& M, uint64_t a, const uint16_t* F,
                                      uint8_t& other) {
    const int l = F[0];
    const uint32_t nb = A.nbr[4 * (uint64_t)M.t + l];
    if (nb == NO_NBR) return false;
    const uint32_t t2 = nb >> 2, l2 = nb & 3;
    if (t2 >= M.t) return false;
    const uint32_t r2 = A.rank[t2];
    if (!(r2 < count_of(A.n_active) && A.act[r2] == t2)) return false;
    uint64_t k3[3];
    bool have_k3 = false;
    const MITet N(A.g, A.rec + A.rec_loc[r2], t2, A.csr_off[t2]);
    const uint16_t* G = N.R.faces();
    for (int f = 0; f < N.R.nf; ++f) {
        const int n = G[3];
        // a copy lies on the shared tet face and has as many corners
        if (G[0] == l2 && n == F[3]) {
            if (!have_k3) {
                mi_face_triple(A.g, M, A.fit, a, F, k3);
                have_k3 = true;
            }
            uint64_t q3[3];
            mi_face_triple(A.g, N, A.fit, r2, G, q3);
            if (q3[0] == k3[0] && q3[1] == k3[1] && q3[2] == k3[2]) {
                bool share = false, first = true;
                mi_side_each(N, A.fit, G[1], [&](uint8_t x) {
                    if (first) other = x;
                    first = false;
                    mi_side_each(M, A.fit, F[1], [&](uint8_t y) { share |= x == y; });
                });
                return !share;
            }
        }
        G += 4 + n;
    }
    return false;
}

// interface-vertex mask of a tet (vertices of faces with positive label > 3)
__device__ inline void mi_iface_mask(const Rec& R, uint64_t m[kMaskWords]) {
    for (int w = 0; w < kMaskWords; ++w) m[w] = 0;
    const uint16_t* F = R.faces();
    for (int f = 0; f < R.nf; ++f) {
        const int n = F[3];
        if (F[0] > 3)
            for (int q = 0; q < n; ++q) m[F[4 + q] >> 6] |= 1ull << (F[4 + q] & 63);
        F += 4 + n;
    }
}
__device__ inline uint32_t mask_rank(const uint64_t m[kMaskWords], int j) {
    uint32_t r = 0;
    for (int w = 0; w < (j >> 6); ++w) r += __popcll(m[w]);
    return r + __popcll(m[j >> 6] & ((1ull << (j & 63)) - 1));
}
__device__ inline bool is_corner(const Rec& R, int j) {
    const uint16_t* L4 = R.vert(j);
    return L4[2] <= 3;  // sorted labels: three tet faces
}
// the tet corner a corner vertex sits at: the one tet face not among its labels
__device__ inline int corner_of(const Rec& R, int j) {
    const uint16_t* L4 = R.vert(j);
    int c = 0;
    while (c == L4[0] || c == L4[1] || c == L4[2]) ++c;
    return c;
}

// Which tet-boundary faces of each active tet become interface faces here
// (the reference's two-map pairing): bit f for face f < 63; bit 63 marks a
// tet with more faces, whose faces from 63 on are tested where they are used.
// The faces' keys are computed once: per active tet its face count (k_mi_nf),
// a scan gives each tet's first face slot, k_mi_face_keys writes every face's
// tet face, corner count, record word offset and (boundary faces) its
// (min, second-min, max) vertex-key triple, and k_mi_bface compares a
// boundary face's triple with those of the neighbour's faces on the shared
// tet face.  kBfaceLanes lanes per tet (adjacent lanes of one wave) take
// faces f = lane, lane + kBfaceLanes, ...
constexpr int kBfaceLanes = 8;
constexpr int kBfaceBits = 63;

struct MIFaceKeys {
    const uint32_t* fb;  // first face slot of each active tet (n_active + 1 entries)
    uint16_t* fl;        // tet face (0..3, or 0xFF inside) | corners << 8
    uint16_t* fw;        // word offset of the face in its record
    uint64_t* ft;        // vertex-key triple (boundary faces)
};

__global__ void k_mi_nf(MIArgs A, uint32_t* __restrict__ nf) {
    GRID_STRIDE(a, A.n_active) {
        const uint16_t* r = A.rec + A.rec_loc[a];
        nf[a] = r[0] == 2 ? r[3] : 0u;
    }
}

__global__ void k_mi_face_keys(MIArgs A, MIFaceKeys K) {
    const uint64_t n = count_of(A.n_active) * kBfaceLanes;
    GRID_STRIDE(idx, n) {
        const uint64_t a = idx / kBfaceLanes;
        const int l0 = (int)(idx % kBfaceLanes);
        const uint32_t t = A.act[a];
        const MITet M(A.g, A.rec + A.rec_loc[a], t, A.csr_off[t]);
        if (M.R.r[0] != 2) continue;
        const int nf = M.R.nf;
        const uint32_t fb = K.fb[a];
        const uint16_t* F = M.R.faces();
        for (int f = 0; f < l0 && f < nf; ++f) F += 4 + F[3];
        for (int f = l0; f < nf; f += kBfaceLanes) {
            const uint32_t slot = fb + (uint32_t)f;
            const int l = F[0];
            K.fl[slot] = (uint16_t)((l <= 3 ? l : 0xFF) | F[3] << 8);
            K.fw[slot] = (uint16_t)(F - M.R.r);
            if (l <= 3) {
                uint64_t k3[3];
                mi_face_triple(A.g, M, A.fit, a, F, k3);
                K.ft[3 * (uint64_t)slot] = k3[0];
                K.ft[3 * (uint64_t)slot + 1] = k3[1];
                K.ft[3 * (uint64_t)slot + 2] = k3[2];
            }
            for (int q = 0; q < kBfaceLanes && f + q + 1 < nf; ++q) F += 4 + F[3];
        }
    }
}

__global__ void k_mi_bface(MIArgs A, MIFaceKeys K, uint64_t* __restrict__ emask) {
    const uint64_t n = count_of(A.n_active) * kBfaceLanes;
    GRID_STRIDE(idx, n) {
        const uint64_t a = idx / kBfaceLanes;
        const int l0 = (int)(idx % kBfaceLanes);
        const uint32_t t = A.act[a];
        const uint16_t* r = A.rec + A.rec_loc[a];
        uint64_t bits = 0;
        if (r[0] == 2) {
            const int nf = r[3], nt = nf < kBfaceBits ? nf : kBfaceBits;
            const uint32_t fb = K.fb[a];
            for (int f = l0; f < nt; f += kBfaceLanes) {
                const uint32_t slot = fb + (uint32_t)f;
                const uint16_t fl = K.fl[slot];
                const uint32_t l = fl & 0xFFu;
                if (l > 3) continue;
                const uint32_t nb = A.nbr[4 * (uint64_t)t + l];
                if (nb == NO_NBR) continue;
                const uint32_t t2 = nb >> 2, l2 = nb & 3;
                if (t2 >= t) continue;
                const uint32_t r2 = A.rank[t2];
                if (!(r2 < count_of(A.n_active) && A.act[r2] == t2)) continue;
                const uint64_t k0 = K.ft[3 * (uint64_t)slot], k1 = K.ft[3 * (uint64_t)slot + 1],
                               k2 = K.ft[3 * (uint64_t)slot + 2];
                // the copy: on the shared tet face, as many corners, same key triple
                const uint16_t want = (uint16_t)(l2 | (fl & 0xFF00u));
                const uint32_t g0 = K.fb[r2], g1 = K.fb[r2 + 1];
                for (uint32_t gs = g0; gs < g1; ++gs) {
                    if (K.fl[gs] != want) continue;
                    if (K.ft[3 * (uint64_t)gs] != k0 || K.ft[3 * (uint64_t)gs + 1] != k1 ||
                        K.ft[3 * (uint64_t)gs + 2] != k2)
                        continue;
                    const MITet M(A.g, r, t, A.csr_off[t]);
                    const MITet N(A.g, A.rec + A.rec_loc[r2], t2, A.csr_off[t2]);
                    const uint16_t* G = N.R.r + K.fw[gs];
                    const uint16_t* F = M.R.r + K.fw[slot];
                    bool share = false;
                    mi_side_each(N, A.fit, G[1], [&](uint8_t x) {
                        mi_side_each(M, A.fit, F[1], [&](uint8_t y) { share |= x == y; });
                    });
                    if (!share) bits |= 1ull << f;
                    break;
                }
            }
            if (nf > kBfaceBits) bits |= 1ull << kBfaceBits;
        }
        // the lanes of one tet are adjacent and leave the loop together
        for (int d = 1; d < kBfaceLanes; d <<= 1) bits |= __shfl_xor(bits, d, 64);
        if (l0 == 0) emask[a] = bits;
    }
}

// whether tet-boundary face f (record words F) emits here, from the mask
__device__ inline bool mi_emits(const MIArgs& A, const MITet& M, uint64_t a, const uint16_t* F, int f, uint64_t em,
                                uint8_t& other) {
    if (f < kBfaceBits) {
        if (!((em >> f) & 1)) return false;
        // rare: the scan again for the first copy's material
        return mi_bface_emits(A, M, a, F, other);
    }
    return mi_bface_emits(A, M, a, F, other);
}

// per active tet: vertex occurrences (phase 0 + phase 1), faces, face corners
__global__ void k_mi_count(MIArgs A, const uint64_t* __restrict__ emask, uint32_t* __restrict__ cnt) {
    GRID_STRIDE(a, A.n_active) {
        const uint32_t t = A.act[a];
        const MITet M(A.g, A.rec + A.rec_loc[a], t, A.csr_off[t]);
        uint32_t nv = 0, nf = 0, nfv = 0;
        if (M.R.r[0] == 2) {
            const uint64_t em = emask[a];
            uint64_t m[kMaskWords];
            mi_iface_mask(M.R, m);
            for (int w = 0; w < kMaskWords; ++w) nv += __popcll(m[w]);
            const uint16_t* F = M.R.faces();
            for (int f = 0; f < M.R.nf; ++f) {
                const int n = F[3];
                if (F[0] > 3) {
                    ++nf;
                    nfv += n;
                } else if (f < kBfaceBits ? ((em >> f) & 1) != 0 : [&] {
                               uint8_t other;
                               return mi_bface_emits(A, M, a, F, other);
                           }()) {
                    ++nf;
                    nfv += n;
                    for (int q = 0; q < n; ++q) nv += is_corner(M.R, F[4 + q]);
                }
                F += 4 + n;
            }
        }
        cnt[3 * a] = nv;
        cnt[3 * a + 1] = nf;
        cnt[3 * a + 2] = nfv;
    }
}

__global__ void __launch_bounds__(256) k_mi_extract(MIArgs A, const uint64_t* __restrict__ emask, const uint32_t* __restrict__ voff,
                             const uint32_t* __restrict__ foff, const uint32_t* __restrict__ fvoff,
                             uint64_t* __restrict__ okey,
                             uint32_t* __restrict__ oa, uint16_t* __restrict__ oj, uint32_t* __restrict__ fq_a,
                             uint16_t* __restrict__ fq_l, uint8_t* __restrict__ fq_m, uint32_t* __restrict__ fq_vs,
                             uint32_t* __restrict__ fve) {
    GRID_STRIDE(a, A.n_active) {
        const uint32_t t = A.act[a];
        const MITet M(A.g, A.rec + A.rec_loc[a], t, A.csr_off[t]);
        if (M.R.r[0] != 2) continue;
        uint64_t m[kMaskWords];
        mi_iface_mask(M.R, m);
        const uint32_t vb = voff[a];
        uint32_t o = vb;
        for (int j = 0; j < M.R.nv; ++j) {
            if (!((m[j >> 6] >> (j & 63)) & 1)) continue;
            int nb;
            okey[o] = mi_vkey(A.g, M, A.fit, a, j, nb);
            oa[o] = (uint32_t)a;
            oj[o] = (uint16_t)j;
            ++o;
        }
        uint32_t q = foff[a], e = fvoff[a];
        const uint64_t em = emask[a];
        const uint16_t* F = M.R.faces();
        for (int f = 0; f < M.R.nf; ++f) {
            const int n = F[3];
            uint8_t other = 0xFF;
            const bool iface = F[0] > 3;
            if (iface || mi_emits(A, M, a, F, f, em, other)) {
                fq_a[q] = (uint32_t)a;
                fq_l[q] = (uint16_t)f;
                fq_m[2 * q] = iface ? A.fit[M.s0 + F[0] - 4] : other;
                fq_m[2 * q + 1] = A.fit[M.s0 + F[1] - 4];
                fq_vs[q] = e;
                for (int i = 0; i < n; ++i) {
                    const int j = F[4 + i];
                    if (!iface && is_corner(M.R, j)) {
                        // phase 1: the corner is created here if it is new
                        okey[o] = key_corner(M.vc(corner_of(M.R, j)));
                        oa[o] = (uint32_t)a;
                        oj[o] = (uint16_t)j;
                        fve[e++] = o++;
                    } else {
                        fve[e++] = vb + mask_rank(m, j);
                    }
                }
                ++q;
            }
            F += 4 + n;
        }
    }
}

// vertex records and coordinates (src/extract_mesh.cpp:640-815, 1541-1637)
__global__ void k_vert_fields_mi(GridView g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                                 const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit,
                                 const uint32_t* __restrict__ rec_loc, const uint16_t* __restrict__ rec,
                                 const uint32_t* __restrict__ first_occ, const uint32_t* __restrict__ oa,
                                 const uint16_t* __restrict__ oj, Count nverts, uint32_t* __restrict__ v_tet,
                                 uint16_t* __restrict__ v_j, uint8_t* __restrict__ v_ss, uint32_t* __restrict__ v_sv,
                                 uint8_t* __restrict__ v_fn, double* __restrict__ xyz) {
    const uint64_t V = g.V;
    GRID_STRIDE(i, nverts) {
        const uint32_t o = first_occ[i];
        const uint32_t a = oa[o];
        const int j = oj[o];
        const uint32_t t = act[a];
        const Rec R(rec + rec_loc[a]);
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        const uint16_t* L4 = R.vert(j);
        int bnd[4], nb = 0;
        uint32_t fn[4] = {0xFF, 0xFF, 0xFF, 0xFF};
        int nm = 0;
        for (int q = 0; q < 4; ++q) {
            if (L4[q] > 3)
                fn[nm++] = fit[s0 + L4[q] - 4];
            else
                bnd[nb++] = L4[q];
        }
        uint32_t sv[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        int ss;
        double X0, X1, X2;
        auto P = [&](uint32_t vid, int ax) { return g.pts[3 * (uint64_t)vid + ax]; };
        auto fv = [&](uint32_t vid, uint32_t f) { return vals[(uint64_t)f * V + vid]; };
        if (nb == 3) {
            int c = 0;
            while (c == bnd[0] || c == bnd[1] || c == bnd[2]) ++c;
            ss = 1;
            sv[0] = v[c];
            for (int q = 0; q < 4; ++q) fn[q] = 0xFF;  // a tet corner carries no materials
            X0 = P(sv[0], 0);
            X1 = P(sv[0], 1);
            X2 = P(sv[0], 2);
        } else if (nb == 2) {
            uint32_t e[2];
            int ne = 0;
            for (int c = 0; c < 4; ++c)
                if (c != bnd[0] && c != bnd[1]) e[ne++] = v[c];
            sv[0] = e[0] < e[1] ? e[0] : e[1];
            sv[1] = e[0] < e[1] ? e[1] : e[0];
            ss = 2;
            const double f1 = fv(sv[0], fn[0]) - fv(sv[0], fn[1]);
            const double f2 = fv(sv[1], fn[0]) - fv(sv[1], fn[1]);
            const double b0 = f2 / (f2 - f1);
            const double b1 = 1 - b0;
            X0 = b0 * P(sv[0], 0) + b1 * P(sv[1], 0);
            X1 = b0 * P(sv[0], 1) + b1 * P(sv[1], 1);
            X2 = b0 * P(sv[0], 2) + b1 * P(sv[1], 2);
        } else if (nb == 1) {
            uint32_t e[3];
            int ne = 0;
            for (int c = 0; c < 4; ++c)
                if (c != bnd[0]) e[ne++] = v[c];
            for (int x = 1; x < 3; ++x)
                for (int y = x; y > 0 && e[y - 1] > e[y]; --y) {
                    uint32_t tmp = e[y];
                    e[y] = e[y - 1];
                    e[y - 1] = tmp;
                }
            for (int q = 0; q < 3; ++q) sv[q] = e[q];
            ss = 3;
            double p1[3], p2[3];
            for (int q = 0; q < 3; ++q) {
                p1[q] = fv(sv[q], fn[0]) - fv(sv[q], fn[1]);
                p2[q] = fv(sv[q], fn[1]) - fv(sv[q], fn[2]);
            }
            const double n1 = p1[2] * p2[1] - p1[1] * p2[2];
            const double n2 = p1[0] * p2[2] - p1[2] * p2[0];
            const double n3 = p1[1] * p2[0] - p1[0] * p2[1];
            const double d = n1 + n2 + n3;
            const double b0 = n1 / d, b1 = n2 / d, b2 = n3 / d;
            X0 = b0 * P(sv[0], 0) + b1 * P(sv[1], 0) + b2 * P(sv[2], 0);
            X1 = b0 * P(sv[0], 1) + b1 * P(sv[1], 1) + b2 * P(sv[2], 1);
            X2 = b0 * P(sv[0], 2) + b1 * P(sv[1], 2) + b2 * P(sv[2], 2);
        } else {
            for (int q = 0; q < 4; ++q) sv[q] = v[q];
            ss = 4;
            double a1[4], a2[4], a3[4];
            for (int q = 0; q < 4; ++q) {
                a1[q] = fv(sv[q], fn[0]) - fv(sv[q], fn[1]);
                a2[q] = fv(sv[q], fn[1]) - fv(sv[q], fn[2]);
                a3[q] = fv(sv[q], fn[2]) - fv(sv[q], fn[3]);
            }
            const double n1 = a1[3] * (a2[2] * a3[1] - a2[1] * a3[2]) + a1[2] * (a2[1] * a3[3] - a2[3] * a3[1]) +
                              a1[1] * (a2[3] * a3[2] - a2[2] * a3[3]);
            const double n2 = a1[3] * (a2[0] * a3[2] - a2[2] * a3[0]) + a1[2] * (a2[3] * a3[0] - a2[0] * a3[3]) +
                              a1[0] * (a2[2] * a3[3] - a2[3] * a3[2]);
            const double n3 = a1[3] * (a2[1] * a3[0] - a2[0] * a3[1]) + a1[1] * (a2[0] * a3[3] - a2[3] * a3[0]) +
                              a1[0] * (a2[3] * a3[1] - a2[1] * a3[3]);
            const double n4 = a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]) + a1[1] * (a2[2] * a3[0] - a2[0] * a3[2]) +
                              a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]);
            const double d = n1 + n2 + n3 + n4;
            const double b0 = n1 / d, b1 = n2 / d, b2 = n3 / d, b3 = n4 / d;
            X0 = b0 * P(sv[0], 0) + b1 * P(sv[1], 0) + b2 * P(sv[2], 0) + b3 * P(sv[3], 0);
            X1 = b0 * P(sv[0], 1) + b1 * P(sv[1], 1) + b2 * P(sv[2], 1) + b3 * P(sv[3], 1);
            X2 = b0 * P(sv[0], 2) + b1 * P(sv[1], 2) + b2 * P(sv[2], 2) + b3 * P(sv[3], 2);
        }
        v_tet[i] = t;
        v_j[i] = (uint16_t)j;
        v_ss[i] = (uint8_t)ss;
        for (int q = 0; q < 4; ++q) {
            v_sv[4 * i + q] = sv[q];
            v_fn[4 * i + q] = (uint8_t)fn[q];
        }
        xyz[3 * i] = X0;
        xyz[3 * i + 1] = X1;
        xyz[3 * i + 2] = X2;
    }
}

// faces are never merged: face q is occurrence q
__global__ void k_mi_face_out(Count nf, const uint32_t* __restrict__ act, const uint32_t* __restrict__ fq_a,
                              const uint16_t* __restrict__ fq_l, const uint8_t* __restrict__ fq_m,
                              const uint32_t* __restrict__ fq_vs, const uint32_t* __restrict__ fve,
                              const uint32_t* __restrict__ vid_of_occ, Count nfv_c, uint32_t* __restrict__ f_vstart,
                              uint32_t* __restrict__ f_verts, uint32_t* __restrict__ f_tfstart,
                              uint32_t* __restrict__ f_tf_tet, uint16_t* __restrict__ f_tf_local,
                              uint8_t* __restrict__ f_func) {
    const uint64_t n = count_of(nf), nfv = count_of(nfv_c);
    GRID_STRIDE(q, nf) {
        const uint32_t b = fq_vs[q], e = (q + 1 < n) ? fq_vs[q + 1] : (uint32_t)nfv;
        f_vstart[q] = b;
        for (uint32_t i = b; i < e; ++i) f_verts[i] = vid_of_occ[fve[i]];
        f_tfstart[q] = (uint32_t)q;
        f_tf_tet[q] = act[fq_a[q]];
        f_tf_local[q] = fq_l[q];
        f_func[2 * q] = fq_m[2 * q];
        f_func[2 * q + 1] = fq_m[2 * q + 1];
    }
}

}  // namespace
}  // namespace rin
