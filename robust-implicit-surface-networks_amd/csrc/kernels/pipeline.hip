// Plain form of the device hot path: one launch per stage.
//
//   values (F x V, function-major)
//   -> per-vertex sign masks (IA) / highest-material masks (MI)
//      src/implicit_arrangement.cpp:53-77, src/material_interface.cpp:53-92
//   -> per-tet active-function masks, CSR by scan + fill
//      src/implicit_arrangement.cpp:82-116, src/material_interface.cpp:94-152
//   -> per-tet cut, one thread per active tet, record into a slot
//      src/implicit_arrangement.cpp:244-306, src/material_interface.cpp:288-351
//   -> iso-vertex / iso-face occurrences in (tet, local) order
//   -> stable radix sort on 64-bit simplex keys, first-occurrence ranks
//      src/extract_mesh.cpp:10-265
//   -> vertex records + xyz, src/extract_mesh.cpp:1446-1538
#include <algorithm>
#include <cstring>
#include <map>
#include <stdexcept>

#include "cut_dev.h"
#include "pipeline.h"
#include "scan_sort.h"

namespace rin {

namespace {

using dev::Complex;
using CMed = dev::Caps<40, 64, 8, 8>;      // up to 4 active functions
using CBig = dev::Caps<132, 284, 10, 12>;  // up to 8 active functions

constexpr uint32_t ERR_CAPACITY = 1, ERR_RECORD = 4;

inline unsigned nblk(uint64_t n, unsigned t = 256) {
    uint64_t g = (n + t - 1) / t;
    if (g > 1u << 20) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__device__ inline void flag_error(uint32_t* errs, uint32_t bit) { errs[threadIdx.x & 63] |= bit; }

// ------------------------------------------------------------------ setup

__global__ void k_grid_pts(GridView g, double* pts) {
    const uint64_t N = (uint64_t)g.res + 1;
    GRID_STRIDE(v, g.V) {
        const uint64_t i = v / (N * N), j = (v / N) % N, k = v % N;
        const double d = double(N - 1);
        pts[3 * v + 0] = (double(i) / d) * (g.bmax[0] - g.bmin[0]) + g.bmin[0];
        pts[3 * v + 1] = (double(j) / d) * (g.bmax[1] - g.bmin[1]) + g.bmin[1];
        pts[3 * v + 2] = (double(k) / d) * (g.bmax[2] - g.bmin[2]) + g.bmin[2];
    }
}

__global__ void k_eval(const double* __restrict__ pts, uint64_t V, const FuncDesc* __restrict__ fd, uint32_t F,
                       double* __restrict__ vals) {
    GRID_STRIDE(v, V) {
        const double x0 = pts[3 * v], x1 = pts[3 * v + 1], x2 = pts[3 * v + 2];
        for (uint32_t f = 0; f < F; ++f) vals[(uint64_t)f * V + v] = dev::eval_func(fd[f], x0, x1, x2);
    }
}

__global__ void k_transpose_in(const double* __restrict__ rm, uint64_t V, uint32_t F, double* __restrict__ vals) {
    GRID_STRIDE(i, V * F) {
        const uint64_t v = i / F, f = i % F;
        vals[f * V + v] = rm[i];
    }
}

// ------------------------------------------------------------------ signs / filter

__global__ void k_signs_ia(const double* __restrict__ vals, uint64_t V, uint32_t F, uint64_t* __restrict__ pos,
                           uint64_t* __restrict__ neg, uint32_t* __restrict__ zeros) {
    GRID_STRIDE(v, V) {
        uint64_t p = 0, n = 0;
        uint32_t z = 0;
        for (uint32_t f = 0; f < F; ++f) {
            const double x = vals[(uint64_t)f * V + v];
            if (x > 0)
                p |= 1ull << f;
            else if (x < 0)
                n |= 1ull << f;
            else
                ++z;
        }
        pos[v] = p;
        neg[v] = n;
        zeros[v] = z;
    }
}

__global__ void k_filter_ia(GridView g, const uint64_t* __restrict__ pos, const uint64_t* __restrict__ neg,
                            uint64_t full, uint64_t* __restrict__ amask, uint32_t* __restrict__ kcnt,
                            uint32_t* __restrict__ isact) {
    GRID_STRIDE(t, g.T) {
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint64_t ap = pos[v[0]] & pos[v[1]] & pos[v[2]] & pos[v[3]];
        const uint64_t an = neg[v[0]] & neg[v[1]] & neg[v[2]] & neg[v[3]];
        const uint64_t m = full & ~(ap | an);
        amask[t] = m;
        kcnt[t] = (uint32_t)__popcll(m);
        isact[t] = m != 0;
    }
}

__global__ void k_highest_mi(const double* __restrict__ vals, uint64_t V, uint32_t F, uint64_t* __restrict__ hm,
                             uint32_t* __restrict__ degen) {
    GRID_STRIDE(v, V) {
        double mx = vals[v];
        uint32_t id = 0, cnt = 1;
        for (uint32_t j = 1; j < F; ++j) {
            const double x = vals[(uint64_t)j * V + v];
            if (x > mx) {
                mx = x;
                id = j;
                cnt = 1;
            } else if (x == mx) {
                ++cnt;
            }
        }
        uint64_t m = 1ull << id;
        if (cnt > 1) {
            m = 0;
            for (uint32_t j = 0; j < F; ++j)
                if (vals[(uint64_t)j * V + v] == mx) m |= 1ull << j;
        }
        hm[v] = m;
        degen[v] = cnt > 1;
    }
}

__global__ void k_filter_mi(GridView g, const double* __restrict__ vals, uint32_t F, const uint64_t* __restrict__ hm,
                            uint64_t* __restrict__ amask, uint32_t* __restrict__ kcnt, uint32_t* __restrict__ isact) {
    const uint64_t V = g.V;
    GRID_STRIDE(t, g.T) {
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        uint64_t u = hm[v[0]] | hm[v[1]] | hm[v[2]] | hm[v[3]];
        if (__popcll(u) < 2) {
            amask[t] = 0;
            kcnt[t] = 0;
            isact[t] = 0;
            continue;
        }
        double mn[4] = {1.79769313486231570815e308, 1.79769313486231570815e308, 1.79769313486231570815e308,
                        1.79769313486231570815e308};
        for (uint64_t m = u; m; m &= m - 1) {
            const uint32_t j = __ffsll((unsigned long long)m) - 1;
            for (int c = 0; c < 4; ++c) {
                const double x = vals[(uint64_t)j * V + v[c]];
                if (x < mn[c]) mn[c] = x;
            }
        }
        for (uint32_t j = 0; j < F; ++j) {
            int gc = 0;
            for (int c = 0; c < 4; ++c) gc += vals[(uint64_t)j * V + v[c]] > mn[c];
            if (gc > 1) u |= 1ull << j;
        }
        amask[t] = u;
        kcnt[t] = (uint32_t)__popcll(u);
        isact[t] = 1;
    }
}

__global__ void k_fill(uint64_t T, const uint64_t* __restrict__ amask, const uint32_t* __restrict__ csr_off,
                       const uint32_t* __restrict__ rank, uint8_t* __restrict__ fit, uint32_t* __restrict__ act,
                       uint8_t* __restrict__ act_k) {
    GRID_STRIDE(t, T) {
        uint64_t m = amask[t];
        if (!m) continue;
        uint32_t o = csr_off[t];
        const uint32_t r = rank[t];
        act[r] = (uint32_t)t;
        act_k[r] = (uint8_t)__popcll(m);
        for (; m; m &= m - 1) fit[o++] = (uint8_t)(__ffsll((unsigned long long)m) - 1);
    }
}

// record slot bound per active count
__device__ inline uint32_t rec_bound(int mi, int k) {
    const uint32_t ia[11] = {0, 128, 256, 448, 768, 1280, 1920, 2560, 3200, 4096, 5120};
    const uint32_t m[11] = {0, 64, 128, 256, 320, 400, 480, 560, 640, 720, 800};
    return (k < 0 || k > 10) ? 6000u : (mi ? m[k] : ia[k]);
}

__global__ void k_rec_bound(uint64_t n, const uint8_t* __restrict__ act_k, int mi, uint32_t* __restrict__ b) {
    GRID_STRIDE(a, n) b[a] = rec_bound(mi, act_k[a]);
}

// ------------------------------------------------------------------ per-tet cut

template <class C>
__device__ inline void ia_counts(const Complex<C>& X, uint32_t& niv, uint32_t& nif, uint32_t& nfv) {
    uint64_t iso[4] = {0, 0, 0, 0};
    nif = nfv = 0;
    for (int f = 0; f < X.nf; ++f) {
        bool is = X.fa[f] > 3;
        if (!is && X.dup) {
            const uint8_t rep = X.grp[X.fa[f]];
            for (int q = 4; q < C::MP; ++q) is |= (X.grp[q] == rep);
        }
        if (!is) continue;
        ++nif;
        nfv += X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) iso[X.fv[f][q] >> 6] |= 1ull << (X.fv[f][q] & 63);
    }
    niv = (uint32_t)(__popcll(iso[0]) + __popcll(iso[1]) + __popcll(iso[2]) + __popcll(iso[3]));
}

template <class C>
__device__ inline void mi_counts(const Complex<C>& X, uint32_t& niv, uint32_t& nif, uint32_t& nfv) {
    // MI faces: positive label > 3; boundary faces are resolved across tets
    uint64_t iso[4] = {0, 0, 0, 0};
    nif = nfv = 0;
    for (int f = 0; f < X.nf; ++f) {
        if (X.fa[f] <= 3) continue;
        ++nif;
        nfv += X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) iso[X.fv[f][q] >> 6] |= 1ull << (X.fv[f][q] & 63);
    }
    niv = (uint32_t)(__popcll(iso[0]) + __popcll(iso[1]) + __popcll(iso[2]) + __popcll(iso[3]));
}

template <class C, bool MI>
__global__ void __launch_bounds__(64) k_cut(GridView g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                                            const uint8_t* __restrict__ act_k, const uint32_t* __restrict__ csr_off,
                                            const uint8_t* __restrict__ fit, uint64_t n_active, int kmin, int kmax,
                                            uint32_t flags, const uint32_t* __restrict__ rec_off,
                                            uint16_t* __restrict__ rec, uint32_t* __restrict__ cnt,
                                            uint32_t* __restrict__ errs) {
    const uint64_t V = g.V;
    GRID_STRIDE(a, n_active) {
        const int k = act_k[a];
        if (k < kmin || k > kmax) continue;
        const uint32_t t = act[a];
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        double P[C::MP - 4][4];
        for (int q = 0; q < k; ++q) {
            const uint32_t f = fit[s0 + ((flags & RUN_REVERSE) ? k - 1 - q : q)];
            for (int c = 0; c < 4; ++c) P[q][c] = vals[(uint64_t)f * V + v[c]];
        }
        Complex<C> X;
        if (MI)
            dev::mi_cut<C>(P, k, X);
        else
            dev::ia_cut<C>(P, k, X);
        if (X.err) {
            flag_error(errs, ERR_CAPACITY);
            cnt[3 * a] = cnt[3 * a + 1] = cnt[3 * a + 2] = 0;
            rec[rec_off[a]] = 0;
            continue;
        }
        const int cap = (int)(rec_off[a + 1] - rec_off[a]);
        const int w = MI ? dev::write_record_mi<C>(X, k, rec + rec_off[a], cap)
                         : dev::write_record_ia<C>(X, k, rec + rec_off[a], cap);
        if (w < 0) flag_error(errs, ERR_RECORD);
        uint32_t a0, a1, a2;
        if (flags & RUN_CUT_ONLY) {
            a0 = X.nv;
            a1 = X.nf;
            a2 = X.nc;
        } else if (MI) {
            mi_counts<C>(X, a0, a1, a2);
        } else {
            ia_counts<C>(X, a0, a1, a2);
        }
        cnt[3 * a] = a0;
        cnt[3 * a + 1] = a1;
        cnt[3 * a + 2] = a2;
    }
}

__global__ void k_split3(uint64_t n, const uint32_t* __restrict__ cnt, uint32_t* a, uint32_t* b, uint32_t* c) {
    GRID_STRIDE(i, n) {
        a[i] = cnt[3 * i];
        b[i] = cnt[3 * i + 1];
        c[i] = cnt[3 * i + 2];
    }
}

// ------------------------------------------------------------------ extraction (IA)

// keys: (payload << 2) | type; type 0 corner, 1 tet edge, 2 tet face, 3 interior
__device__ inline uint64_t key_corner(uint32_t v) { return ((uint64_t)v << 2) | 0; }
__device__ inline uint64_t key_edge(uint64_t e, uint32_t f) { return ((e << 6 | f) << 2) | 1; }
__device__ inline uint64_t key_face(uint64_t fc, uint32_t f0, uint32_t f1) {
    return ((((fc << 6) | f0) << 6 | f1) << 2) | 2;
}
__device__ inline uint64_t key_inner(uint64_t a, uint32_t j) { return ((a << 8 | j) << 2) | 3; }

struct Rec {
    const uint16_t* r;
    int L, nv, nf;
    bool dup;
    __device__ Rec(const uint16_t* p) : r(p), L(p[1]), nv(p[2]), nf(p[3]), dup(p[5] & 1) {}
    __device__ const uint16_t* vert(int j) const { return r + 8 + 4 * j; }
    __device__ const uint16_t* faces() const { return r + 8 + 4 * nv; }
    __device__ const uint16_t* groups_ia() const { return r + r[6] - 2 * L; }
};

__device__ inline bool ia_face_is_iso(const Rec& R, const uint16_t* F) {
    if (F[0] > 3) return true;
    if (!R.dup) return false;
    const uint16_t* grp = R.groups_ia();
    const uint16_t rep = grp[F[0]];
    for (int q = 4; q < R.L; ++q)
        if (grp[q] == rep) return true;
    return false;
}

__global__ void k_extract_ia(GridView g, const uint32_t* __restrict__ act, const uint32_t* __restrict__ csr_off,
                             const uint8_t* __restrict__ fit, uint64_t n_active, const uint32_t* __restrict__ rec_off,
                             const uint16_t* __restrict__ rec, const uint32_t* __restrict__ voff,
                             const uint32_t* __restrict__ foff, const uint32_t* __restrict__ fvoff,
                             uint64_t* __restrict__ okey, uint32_t* __restrict__ oa, uint16_t* __restrict__ oj,
                             uint32_t* __restrict__ fq_a, uint16_t* __restrict__ fq_l, uint8_t* __restrict__ fq_fn,
                             uint8_t* __restrict__ fq_bnd, uint32_t* __restrict__ fq_vs, uint32_t* __restrict__ fve) {
    GRID_STRIDE(a, n_active) {
        const uint32_t t = act[a];
        const Rec R(rec + rec_off[a]);
        if (R.r[0] != 1) continue;
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        uint64_t iso[4] = {0, 0, 0, 0};
        const uint16_t* F = R.faces();
        for (int f = 0; f < R.nf; ++f) {
            const int n = F[3];
            if (ia_face_is_iso(R, F))
                for (int q = 0; q < n; ++q) iso[F[4 + q] >> 6] |= 1ull << (F[4 + q] & 63);
            F += 4 + n;
        }
        // vertex occurrences in local order
        uint32_t o = voff[a];
        for (int j = 0; j < R.nv; ++j) {
            if (!((iso[j >> 6] >> (j & 63)) & 1)) continue;
            const uint16_t* T3 = R.vert(j);
            int bnd[3], nb = 0;
            uint32_t fn[3];
            int ni = 0;
            for (int q = 0; q < 3; ++q) {
                if (T3[q] > 3)
                    fn[ni++] = fit[s0 + T3[q] - 4];
                else
                    bnd[nb++] = T3[q];
            }
            uint64_t key;
            if (nb == 3) {
                int c = 0;
                while (c == bnd[0] || c == bnd[1] || c == bnd[2]) ++c;
                key = key_corner(v[c]);
            } else if (nb == 2) {
                int l[2], nl = 0;
                for (int c = 0; c < 4; ++c)
                    if (c != bnd[0] && c != bnd[1]) l[nl++] = c;
                key = key_edge(dev::edge_id(g, t, v, l[0], l[1]), fn[0]);
            } else if (nb == 1) {
                key = key_face(dev::face_id(g, t, v, bnd[0]), fn[0], fn[1]);
            } else {
                key = key_inner(a, j);
            }
            okey[o] = key;
            oa[o] = (uint32_t)a;
            oj[o] = (uint16_t)j;
            ++o;
        }
        // face occurrences in local order; vertex slots point at occurrences
        uint32_t q = foff[a], e = fvoff[a];
        const uint32_t vb = voff[a];
        F = R.faces();
        for (int f = 0; f < R.nf; ++f) {
            const int n = F[3];
            if (ia_face_is_iso(R, F)) {
                int sp = F[0];
                if (sp <= 3) {
                    // a tet face coinciding with implicit planes: the smallest of them
                    const uint16_t* grp = R.groups_ia();
                    for (int p = 4; p < R.L; ++p)
                        if (grp[p] == grp[sp]) {
                            sp = p;
                            break;
                        }
                }
                fq_a[q] = (uint32_t)a;
                fq_l[q] = (uint16_t)f;
                fq_fn[q] = fit[s0 + sp - 4];
                fq_bnd[q] = F[2] == 0xFFFF;
                fq_vs[q] = e;
                for (int i = 0; i < n; ++i) {
                    const int j = F[4 + i];
                    uint32_t rank = 0;
                    for (int w = 0; w < (j >> 6); ++w) rank += __popcll(iso[w]);
                    rank += __popcll(iso[j >> 6] & ((1ull << (j & 63)) - 1));
                    fve[e++] = vb + rank;
                }
                ++q;
            }
            F += 4 + n;
        }
    }
}

// ------------------------------------------------------------------ dedup

__global__ void k_heads(uint64_t n, const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                        uint32_t* __restrict__ first, uint32_t* __restrict__ head) {
    GRID_STRIDE(p, n) {
        const uint32_t h = (p == 0 || skey[p] != skey[p - 1]) ? 1u : 0u;
        head[p] = h;
        first[sidx[p]] = h;
    }
}

// head_excl: exclusive scan of head over sorted positions; rank: exclusive
// scan of `first` over occurrence order
__global__ void k_assign(uint64_t n, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ head,
                         const uint32_t* __restrict__ head_excl, uint32_t* __restrict__ seg_first_occ) {
    GRID_STRIDE(p, n) {
        if (head[p]) seg_first_occ[head_excl[p]] = sidx[p];
    }
}
__global__ void k_assign2(uint64_t n, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ head,
                          const uint32_t* __restrict__ head_excl, const uint32_t* __restrict__ seg_first_occ,
                          const uint32_t* __restrict__ rank, uint32_t* __restrict__ id_of_occ) {
    GRID_STRIDE(p, n) {
        const uint32_t seg = head_excl[p] + head[p] - 1;
        id_of_occ[sidx[p]] = rank[seg_first_occ[seg]];
    }
}
__global__ void k_first_list(uint64_t n, const uint32_t* __restrict__ first, const uint32_t* __restrict__ rank,
                             uint32_t* __restrict__ first_occ) {
    GRID_STRIDE(o, n) {
        if (first[o]) first_occ[rank[o]] = (uint32_t)o;
    }
}

// vertex records and coordinates (src/extract_mesh.cpp:95-228, 1446-1538)
__global__ void k_vert_fields_ia(GridView g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                                 const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit,
                                 const uint32_t* __restrict__ rec_off, const uint16_t* __restrict__ rec,
                                 const uint32_t* __restrict__ first_occ, const uint32_t* __restrict__ oa,
                                 const uint16_t* __restrict__ oj, uint64_t nverts, uint32_t* __restrict__ v_tet,
                                 uint16_t* __restrict__ v_j, uint8_t* __restrict__ v_ss, uint32_t* __restrict__ v_sv,
                                 uint8_t* __restrict__ v_fn, double* __restrict__ xyz) {
    const uint64_t V = g.V;
    GRID_STRIDE(i, nverts) {
        const uint32_t o = first_occ[i];
        const uint32_t a = oa[o];
        const int j = oj[o];
        const uint32_t t = act[a];
        const Rec R(rec + rec_off[a]);
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        const uint16_t* T3 = R.vert(j);
        int bnd[3], nb = 0;
        uint32_t fn[3] = {0xFF, 0xFF, 0xFF};
        int ni = 0;
        for (int q = 0; q < 3; ++q) {
            if (T3[q] > 3)
                fn[ni++] = fit[s0 + T3[q] - 4];
            else
                bnd[nb++] = T3[q];
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
            const double f1 = fv(sv[0], fn[0]), f2 = fv(sv[1], fn[0]);
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
                p1[q] = fv(sv[q], fn[0]);
                p2[q] = fv(sv[q], fn[1]);
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
                a1[q] = fv(sv[q], fn[0]);
                a2[q] = fv(sv[q], fn[1]);
                a3[q] = fv(sv[q], fn[2]);
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
            v_fn[4 * i + q] = q < 3 ? (uint8_t)fn[q] : 0xFF;
        }
        xyz[3 * i] = X0;
        xyz[3 * i + 1] = X1;
        xyz[3 * i + 2] = X2;
    }
}

// ------------------------------------------------------------------ faces

// global vertex ids of every face occurrence; boundary faces get their
// (min, second-min, max) key (src/extract_mesh.h:68-92)
__global__ void k_face_keys(uint64_t nq, const uint32_t* __restrict__ fq_vs, const uint32_t* __restrict__ fq_vs_end,
                            uint32_t* __restrict__ fve, const uint32_t* __restrict__ vid_of_occ,
                            const uint8_t* __restrict__ fq_bnd, int vb, uint64_t* __restrict__ khi,
                            uint64_t* __restrict__ klo) {
    GRID_STRIDE(q, nq) {
        const uint32_t b = fq_vs[q], e = fq_vs_end[q];
        for (uint32_t i = b; i < e; ++i) fve[i] = vid_of_occ[fve[i]];
        if (!fq_bnd[q]) {
            // interior faces are never merged; they sort after every boundary key
            khi[q] = 1ull << (2 * vb);
            klo[q] = 0;
            continue;
        }
        uint32_t mn = fve[b], mx = fve[b], mp = b;
        for (uint32_t i = b + 1; i < e; ++i) {
            if (fve[i] < mn) {
                mn = fve[i];
                mp = i;
            } else if (fve[i] > mx) {
                mx = fve[i];
            }
        }
        uint32_t sm = mx + 1;
        for (uint32_t i = b; i < e; ++i)
            if (i != mp && fve[i] < sm) sm = fve[i];
        khi[q] = ((uint64_t)mn << vb) | sm;
        klo[q] = mx;
    }
}

__global__ void k_face_heads(uint64_t n, const uint32_t* __restrict__ perm, const uint64_t* __restrict__ khi,
                             const uint64_t* __restrict__ klo, const uint8_t* __restrict__ bnd,
                             uint32_t* __restrict__ first, uint32_t* __restrict__ head) {
    GRID_STRIDE(p, n) {
        const uint32_t q = perm[p];
        uint32_t h = 1;
        if (p > 0 && bnd[q]) {
            const uint32_t r = perm[p - 1];
            h = (bnd[r] && khi[r] == khi[q] && klo[r] == klo[q]) ? 0u : 1u;
        }
        head[p] = h;
        first[q] = h;
    }
}

// per unique face: occurrence count and vertex count
__global__ void k_face_counts(uint64_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ head,
                              const uint32_t* __restrict__ head_excl, const uint32_t* __restrict__ seg_first_occ,
                              const uint32_t* __restrict__ rank, uint32_t* __restrict__ tf_cnt) {
    GRID_STRIDE(p, n) {
        const uint32_t seg = head_excl[p] + head[p] - 1;
        atomicAdd(&tf_cnt[rank[seg_first_occ[seg]]], 1u);
    }
}

__global__ void k_face_out(uint64_t nfaces, const uint32_t* __restrict__ first_occ, const uint32_t* __restrict__ fq_vs,
                           const uint32_t* __restrict__ fq_vs_end, const uint32_t* __restrict__ fve,
                           const uint8_t* __restrict__ fq_fn, const uint32_t* __restrict__ fvstart,
                           uint32_t* __restrict__ fverts, uint8_t* __restrict__ ffunc) {
    GRID_STRIDE(i, nfaces) {
        const uint32_t q = first_occ[i];
        uint32_t o = fvstart[i];
        for (uint32_t e = fq_vs[q]; e < fq_vs_end[q]; ++e) fverts[o++] = fve[e];
        ffunc[2 * i] = fq_fn[q];
        ffunc[2 * i + 1] = 0xFF;
    }
}

__global__ void k_face_vcount(uint64_t nfaces, const uint32_t* __restrict__ first_occ, const uint32_t* __restrict__ fq_vs,
                              const uint32_t* __restrict__ fq_vs_end, uint32_t* __restrict__ cnt) {
    GRID_STRIDE(i, nfaces) {
        const uint32_t q = first_occ[i];
        cnt[i] = fq_vs_end[q] - fq_vs[q];
    }
}

// occurrences of each unique face in occurrence order: sorted positions of a
// segment are consecutive and (stable sort) already in occurrence order
__global__ void k_face_tf(uint64_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ head,
                          const uint32_t* __restrict__ head_excl, const uint32_t* __restrict__ seg_first_occ,
                          const uint32_t* __restrict__ seg_start, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ tfstart, const uint32_t* __restrict__ act,
                          const uint32_t* __restrict__ fq_a, const uint16_t* __restrict__ fq_l,
                          uint32_t* __restrict__ tf_tet, uint16_t* __restrict__ tf_local) {
    GRID_STRIDE(p, n) {
        const uint32_t seg = head_excl[p] + head[p] - 1;
        const uint32_t fid = rank[seg_first_occ[seg]];
        const uint32_t pos = tfstart[fid] + ((uint32_t)p - seg_start[seg]);
        const uint32_t q = perm[p];
        tf_tet[pos] = act[fq_a[q]];
        tf_local[pos] = fq_l[q];
    }
}
__global__ void k_seg_start(uint64_t n, const uint32_t* __restrict__ head, const uint32_t* __restrict__ head_excl,
                            uint32_t* __restrict__ seg_start) {
    GRID_STRIDE(p, n) {
        if (head[p]) seg_start[head_excl[p]] = (uint32_t)p;
    }
}

// ------------------------------------------------------------------ helpers

template <class T>
T read1(const T* d, hipStream_t s) {
    T h;
    RIN_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    return h;
}

struct Timer {
    hipStream_t s;
    std::vector<std::pair<std::string, hipEvent_t>> ev;
    explicit Timer(hipStream_t st) : s(st) { mark("start"); }
    void mark(const char* name) {
        hipEvent_t e;
        RIN_HIP(hipEventCreate(&e));
        RIN_HIP(hipEventRecord(e, s));
        ev.push_back({name, e});
    }
    void finish(StageTimes& out) {
        RIN_HIP(hipEventSynchronize(ev.back().second));
        for (size_t i = 1; i < ev.size(); ++i) {
            float ms = 0;
            RIN_HIP(hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second));
            out.ms.push_back({ev[i].first, ms});
        }
        for (auto& e : ev) (void)hipEventDestroy(e.second);
        ev.clear();
    }
};

inline int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

}  // namespace

// ------------------------------------------------------------------ API

void grid_regular(DevMem& mem, int res, const double bmin[3], const double bmax[3], Grid& g, hipStream_t s) {
    if (res < 1) throw std::invalid_argument("grid resolution must be >= 1");
    const uint64_t N = (uint64_t)res + 1;
    g.view = GridView{};
    g.view.regular = 1;
    g.view.res = res;
    for (int a = 0; a < 3; ++a) {
        g.view.bmin[a] = bmin[a];
        g.view.bmax[a] = bmax[a];
    }
    g.view.V = N * N * N;
    g.view.T = 5ull * res * res * res;
    if (g.view.V >= (1ull << 32) / 27) throw std::invalid_argument("grid too large for 32-bit vertex ids");
    g.view.n_edge_ids = g.view.V * 27;
    g.view.n_face_ids = g.view.V * 729;
    g.pts.alloc(mem, 3 * g.view.V);
    g.view.pts = g.pts.p;
    hipLaunchKernelGGL(k_grid_pts, dim3(nblk(g.view.V)), dim3(256), 0, s, g.view, g.pts.p);
    RIN_HIP(hipGetLastError());
}

void grid_explicit(DevMem& mem, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, Grid& g,
                   hipStream_t s) {
    if (V >= (1ull << 32)) throw std::invalid_argument("too many vertices");
    g.view = GridView{};
    g.view.regular = 0;
    g.view.V = V;
    g.view.T = T;
    std::vector<uint32_t> t32(4 * T), te(6 * T), tf(4 * T);
    for (uint64_t i = 0; i < 4 * T; ++i) {
        if (tets[i] >= V) throw std::invalid_argument("tet vertex index out of range");
        t32[i] = (uint32_t)tets[i];
    }
    // global edge and face ids, numbered in ascending order of their sorted
    // vertex tuples
    {
        std::vector<std::pair<uint64_t, uint64_t>> ek(6 * T);
        const int le[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        for (uint64_t t = 0; t < T; ++t)
            for (int e = 0; e < 6; ++e) {
                uint64_t a = t32[4 * t + le[e][0]], b = t32[4 * t + le[e][1]];
                if (a > b) std::swap(a, b);
                ek[6 * t + e] = {(a << 32) | b, 6 * t + e};
            }
        std::sort(ek.begin(), ek.end());
        uint32_t id = 0;
        for (size_t i = 0; i < ek.size(); ++i) {
            if (i && ek[i].first != ek[i - 1].first) ++id;
            te[ek[i].second] = id;
        }
        g.view.n_edge_ids = ek.empty() ? 0 : id + 1;
    }
    {
        struct FK {
            uint32_t a, b, c;
            uint64_t slot;
        };
        std::vector<FK> fk(4 * T);
        for (uint64_t t = 0; t < T; ++t)
            for (int c = 0; c < 4; ++c) {
                uint32_t f[3];
                int n = 0;
                for (int q = 0; q < 4; ++q)
                    if (q != c) f[n++] = t32[4 * t + q];
                std::sort(f, f + 3);
                fk[4 * t + c] = {f[0], f[1], f[2], 4 * t + c};
            }
        std::sort(fk.begin(), fk.end(), [](const FK& x, const FK& y) {
            if (x.a != y.a) return x.a < y.a;
            if (x.b != y.b) return x.b < y.b;
            if (x.c != y.c) return x.c < y.c;
            return x.slot < y.slot;
        });
        uint32_t id = 0;
        for (size_t i = 0; i < fk.size(); ++i) {
            if (i && (fk[i].a != fk[i - 1].a || fk[i].b != fk[i - 1].b || fk[i].c != fk[i - 1].c)) ++id;
            tf[fk[i].slot] = id;
        }
        g.view.n_face_ids = fk.empty() ? 0 : id + 1;
    }
    g.pts.alloc(mem, 3 * V);
    g.tets.alloc(mem, 4 * T);
    g.tedge.alloc(mem, 6 * T);
    g.tface.alloc(mem, 4 * T);
    RIN_HIP(hipMemcpyAsync(g.pts.p, pts, 24 * V, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(g.tets.p, t32.data(), 16 * T, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(g.tedge.p, te.data(), 24 * T, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(g.tface.p, tf.data(), 16 * T, hipMemcpyHostToDevice, s));
    RIN_HIP(hipStreamSynchronize(s));
    g.view.pts = g.pts.p;
    g.view.tets = g.tets.p;
    g.view.tedge = g.tedge.p;
    g.view.tface = g.tface.p;
}

void values_eval(DevMem& mem, const Grid& g, const FuncDesc* funcs, uint32_t F, Values& v, hipStream_t s) {
    if (F == 0 || F > kMaxFuncs) throw std::invalid_argument("function count must be 1..64");
    v.F = F;
    v.V = g.view.V;
    v.vals.alloc(mem, (size_t)F * v.V);
    DBuf<FuncDesc> fd(mem, F);
    RIN_HIP(hipMemcpyAsync(fd.p, funcs, sizeof(FuncDesc) * F, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_eval, dim3(nblk(v.V)), dim3(256), 0, s, (const double*)g.pts.p, v.V, (const FuncDesc*)fd.p, F,
                       v.vals.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipStreamSynchronize(s));
}

void values_upload(DevMem& mem, const Grid& g, const double* rowmajor, uint32_t F, Values& v, hipStream_t s) {
    if (F == 0 || F > kMaxFuncs) throw std::invalid_argument("function count must be 1..64");
    v.F = F;
    v.V = g.view.V;
    v.vals.alloc(mem, (size_t)F * v.V);
    DBuf<double> tmp(mem, (size_t)F * v.V);
    RIN_HIP(hipMemcpyAsync(tmp.p, rowmajor, 8 * (size_t)F * v.V, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_transpose_in, dim3(nblk(F * v.V)), dim3(256), 0, s, (const double*)tmp.p, v.V, F, v.vals.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipStreamSynchronize(s));
}

void run_pipeline(DevMem& mem, const Grid& g, const Values& val, uint32_t flags, RunOut& out, hipStream_t s) {
    const bool MI = flags & RUN_MI;
    const uint64_t V = g.view.V, T = g.view.T;
    const uint32_t F = val.F;
    if (val.V != V) throw std::invalid_argument("values do not match the grid");
    out.mi = MI;
    out.flags = flags;
    out.T = T;
    out.F = F;
    Timer tm(s);
    DBuf<uint32_t> errs(mem, 64);
    fill_u32(errs.p, 0, 64, s);

    // ---- signs / highest material, filter, CSR
    DBuf<uint64_t> amask(mem, T);
    DBuf<uint32_t> kcnt(mem, T + 1), isact(mem, T + 1);
    out.csr_off.alloc(mem, T + 1);
    out.rank_of_tet.alloc(mem, T + 1);
    DBuf<uint32_t> tot(mem, 2);
    DBuf<uint32_t> vflag(mem, V);
    if (!MI) {
        DBuf<uint64_t> pos(mem, V), neg(mem, V);
        hipLaunchKernelGGL(k_signs_ia, dim3(nblk(V)), dim3(256), 0, s, (const double*)val.vals.p, V, F, pos.p, neg.p,
                           vflag.p);
        tm.mark("func signs");
        const uint64_t full = (F == 64) ? ~0ull : ((1ull << F) - 1);
        hipLaunchKernelGGL(k_filter_ia, dim3(nblk(T)), dim3(256), 0, s, g.view, (const uint64_t*)pos.p,
                           (const uint64_t*)neg.p, full, amask.p, kcnt.p, isact.p);
    } else {
        DBuf<uint64_t> hm(mem, V);
        hipLaunchKernelGGL(k_highest_mi, dim3(nblk(V)), dim3(256), 0, s, (const double*)val.vals.p, V, F, hm.p,
                           vflag.p);
        tm.mark("highest func");
        hipLaunchKernelGGL(k_filter_mi, dim3(nblk(T)), dim3(256), 0, s, g.view, (const double*)val.vals.p, F,
                           (const uint64_t*)hm.p, amask.p, kcnt.p, isact.p);
    }
    RIN_HIP(hipGetLastError());
    {
        DBuf<uint32_t> vsum(mem, 1);
        scan_u32(mem, vflag.p, vflag.p, V, vsum.p, s);
        out.num_degenerate = read1(vsum.p, s);  // IA: zero entries; MI: tied vertices
    }
    scan_u32(mem, kcnt.p, out.csr_off.p, T, out.csr_off.p + T, s);
    scan_u32(mem, isact.p, out.rank_of_tet.p, T, tot.p, s);
    out.n_fit = read1(out.csr_off.p + T, s);
    out.n_active = read1(tot.p, s);
    out.num_intersecting = out.n_active;
    const uint64_t NA = out.n_active;
    out.fit.alloc(mem, out.n_fit + 1);
    out.act.alloc(mem, NA + 1);
    out.act_k.alloc(mem, NA + 1);
    hipLaunchKernelGGL(k_fill, dim3(nblk(T)), dim3(256), 0, s, T, (const uint64_t*)amask.p,
                       (const uint32_t*)out.csr_off.p, (const uint32_t*)out.rank_of_tet.p, out.fit.p, out.act.p,
                       out.act_k.p);
    RIN_HIP(hipGetLastError());
    tm.mark("filter");
    amask.reset();
    kcnt.reset();
    isact.reset();
    vflag.reset();

    // bucket counts (reference labels: IA 1/2/>=3, MI 2/3/>=4)
    {
        std::vector<uint8_t> hk(NA);
        if (NA) RIN_HIP(hipMemcpyAsync(hk.data(), out.act_k.p, NA, hipMemcpyDeviceToHost, s));
        RIN_HIP(hipStreamSynchronize(s));
        int kmax = 0;
        for (uint8_t k : hk) {
            const int lo = MI ? 2 : 1;
            out.nb[k <= lo ? 0 : (k == lo + 1 ? 1 : 2)]++;
            kmax = std::max<int>(kmax, k);
        }
        if (kmax > 8) {
            out.err |= ERR_CAPACITY;
            throw std::runtime_error("a tet has " + std::to_string(kmax) +
                                     " active functions; the device cut supports at most 8");
        }
    }

    // ---- per-tet cut
    out.rec_off.alloc(mem, NA + 1);
    {
        DBuf<uint32_t> b(mem, NA + 1);
        hipLaunchKernelGGL(k_rec_bound, dim3(nblk(NA)), dim3(256), 0, s, NA, (const uint8_t*)out.act_k.p, (int)MI,
                           b.p);
        scan_u32(mem, b.p, out.rec_off.p, NA, out.rec_off.p + NA, s);
    }
    const uint64_t nwords = read1(out.rec_off.p + NA, s);
    out.rec.alloc(mem, nwords + 1);
    DBuf<uint32_t> cnt(mem, 3 * NA + 3);
    if (NA) {
        if (!MI) {
            hipLaunchKernelGGL((k_cut<CMed, false>), dim3(nblk(NA, 64)), dim3(64), 0, s, g.view,
                               (const double*)val.vals.p, (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA, 1, 4, flags,
                               (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p);
            hipLaunchKernelGGL((k_cut<CBig, false>), dim3(nblk(NA, 64)), dim3(64), 0, s, g.view,
                               (const double*)val.vals.p, (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA, 5, 8, flags,
                               (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p);
        } else {
            hipLaunchKernelGGL((k_cut<CMed, true>), dim3(nblk(NA, 64)), dim3(64), 0, s, g.view,
                               (const double*)val.vals.p, (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA, 1, 4, flags,
                               (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p);
            hipLaunchKernelGGL((k_cut<CBig, true>), dim3(nblk(NA, 64)), dim3(64), 0, s, g.view,
                               (const double*)val.vals.p, (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA, 5, 8, flags,
                               (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p);
        }
        RIN_HIP(hipGetLastError());
    }
    tm.mark(MI ? "MI cut" : "simp_arr");
    {
        uint32_t he[64];
        RIN_HIP(hipMemcpyAsync(he, errs.p, sizeof(he), hipMemcpyDeviceToHost, s));
        RIN_HIP(hipStreamSynchronize(s));
        for (uint32_t e : he) out.err |= e;
        if (out.err) {
            tm.finish(out.times);
            throw std::runtime_error(std::string("device cut failed: ") +
                                     ((out.err & ERR_CAPACITY) ? "complex exceeds capacity or predicate range; "
                                                               : "") +
                                     ((out.err & ERR_RECORD) ? "record overflow" : ""));
        }
    }
    if (flags & RUN_CUT_ONLY) {
        out.cut_counts = std::move(cnt);
        tm.finish(out.times);
        return;
    }
    if (MI) {
        tm.finish(out.times);
        throw std::runtime_error("material-interface extraction is not available on the device path yet");
    }

    // ---- extraction
    DBuf<uint32_t> cv(mem, NA + 1), cf(mem, NA + 1), cfv(mem, NA + 1);
    hipLaunchKernelGGL(k_split3, dim3(nblk(NA)), dim3(256), 0, s, NA, (const uint32_t*)cnt.p, cv.p, cf.p, cfv.p);
    scan_u32(mem, cv.p, cv.p, NA, cv.p + NA, s);
    scan_u32(mem, cf.p, cf.p, NA, cf.p + NA, s);
    scan_u32(mem, cfv.p, cfv.p, NA, cfv.p + NA, s);
    const uint64_t NVO = read1(cv.p + NA, s), NFO = read1(cf.p + NA, s), NFV = read1(cfv.p + NA, s);
    DBuf<uint64_t> okey(mem, NVO + 1);
    DBuf<uint32_t> oa(mem, NVO + 1);
    DBuf<uint16_t> oj(mem, NVO + 1);
    DBuf<uint32_t> fq_a(mem, NFO + 1), fq_vs(mem, NFO + 1), fve(mem, NFV + 1);
    DBuf<uint16_t> fq_l(mem, NFO + 1);
    DBuf<uint8_t> fq_fn(mem, NFO + 1), fq_bnd(mem, NFO + 1);
    if (NA)
        hipLaunchKernelGGL(k_extract_ia, dim3(nblk(NA)), dim3(256), 0, s, g.view, (const uint32_t*)out.act.p,
                           (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA,
                           (const uint32_t*)out.rec_off.p, (const uint16_t*)out.rec.p, (const uint32_t*)cv.p,
                           (const uint32_t*)cf.p, (const uint32_t*)cfv.p, okey.p, oa.p, oj.p, fq_a.p, fq_l.p,
                           fq_fn.p, fq_bnd.p, fq_vs.p, fve.p);
    RIN_HIP(hipGetLastError());
    tm.mark("extract occurrences");

    // ---- vertex dedup: stable sort by key, first occurrence ranks
    const int kbits = 2 + std::max({bits_for(V), bits_for(g.view.n_edge_ids) + 6, bits_for(g.view.n_face_ids) + 12,
                                    bits_for(NA) + 8});
    if (kbits > 64) throw std::runtime_error("simplex keys exceed 64 bits for this mesh");
    DBuf<uint32_t> sidx(mem, NVO + 1), first(mem, NVO + 1), head(mem, NVO + 1), hexcl(mem, NVO + 1),
        seg_first(mem, NVO + 1), vid(mem, NVO + 1);
    DBuf<uint32_t> nvt(mem, 2);
    iota_u32(sidx.p, NVO, s);
    sort_pairs(mem, okey.p, sidx.p, NVO, kbits, s);
    hipLaunchKernelGGL(k_heads, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint64_t*)okey.p,
                       (const uint32_t*)sidx.p, first.p, head.p);
    scan_u32(mem, head.p, hexcl.p, NVO, nullptr, s);
    hipLaunchKernelGGL(k_assign, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, seg_first.p);
    DBuf<uint32_t> vrank(mem, NVO + 1);
    scan_u32(mem, first.p, vrank.p, NVO, nvt.p, s);
    hipLaunchKernelGGL(k_assign2, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, (const uint32_t*)seg_first.p,
                       (const uint32_t*)vrank.p, vid.p);
    out.n_verts = read1(nvt.p, s);
    const uint64_t NV = out.n_verts;
    DBuf<uint32_t> first_occ(mem, NV + 1);
    hipLaunchKernelGGL(k_first_list, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)first.p,
                       (const uint32_t*)vrank.p, first_occ.p);
    tm.mark("dedup verts");
    out.v_tet.alloc(mem, NV + 1);
    out.v_j.alloc(mem, NV + 1);
    out.v_ss.alloc(mem, NV + 1);
    out.v_sv.alloc(mem, 4 * NV + 4);
    out.v_fn.alloc(mem, 4 * NV + 4);
    out.xyz.alloc(mem, 3 * NV + 3);
    hipLaunchKernelGGL(k_vert_fields_ia, dim3(nblk(NV)), dim3(256), 0, s, g.view, (const double*)val.vals.p,
                       (const uint32_t*)out.act.p, (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p,
                       (const uint32_t*)out.rec_off.p, (const uint16_t*)out.rec.p, (const uint32_t*)first_occ.p,
                       (const uint32_t*)oa.p, (const uint16_t*)oj.p, NV, out.v_tet.p, out.v_j.p, out.v_ss.p,
                       out.v_sv.p, out.v_fn.p, out.xyz.p);
    RIN_HIP(hipGetLastError());
    tm.mark("compute xyz");

    // ---- faces: boundary faces dedup by key, interior faces always new
    // face occurrences are laid out back to back: each ends where the next starts
    DBuf<uint32_t> fq_end(mem, NFO + 1);
    if (NFO) {
        RIN_HIP(hipMemcpyAsync(fq_end.p, fq_vs.p + 1, sizeof(uint32_t) * (NFO - 1), hipMemcpyDeviceToDevice, s));
        fill_u32(fq_end.p + NFO - 1, (uint32_t)NFV, 1, s);
    }
    DBuf<uint64_t> khi(mem, NFO + 1), klo(mem, NFO + 1);
    const int vb = bits_for(NV + 1);
    if (2 * vb + 1 > 64) throw std::runtime_error("boundary face keys exceed 64 bits");
    hipLaunchKernelGGL(k_face_keys, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)fq_vs.p,
                       (const uint32_t*)fq_end.p, fve.p, (const uint32_t*)vid.p, (const uint8_t*)fq_bnd.p, vb, khi.p,
                       klo.p);
    DBuf<uint32_t> perm(mem, NFO + 1), ffirst(mem, NFO + 1), fhead(mem, NFO + 1), fhex(mem, NFO + 1),
        fsegfirst(mem, NFO + 1), frank(mem, NFO + 1), fsegstart(mem, NFO + 1);
    sort_index_by2(mem, khi.p, klo.p, 2 * vb + 1, vb, perm.p, NFO, s);
    hipLaunchKernelGGL(k_face_heads, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)perm.p,
                       (const uint64_t*)khi.p, (const uint64_t*)klo.p, (const uint8_t*)fq_bnd.p, ffirst.p, fhead.p);
    scan_u32(mem, fhead.p, fhex.p, NFO, nullptr, s);
    hipLaunchKernelGGL(k_assign, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)perm.p,
                       (const uint32_t*)fhead.p, (const uint32_t*)fhex.p, fsegfirst.p);
    hipLaunchKernelGGL(k_seg_start, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)fhead.p,
                       (const uint32_t*)fhex.p, fsegstart.p);
    DBuf<uint32_t> nft(mem, 2);
    scan_u32(mem, ffirst.p, frank.p, NFO, nft.p, s);
    out.n_faces = read1(nft.p, s);
    const uint64_t NF = out.n_faces;
    DBuf<uint32_t> ffirst_occ(mem, NF + 1);
    hipLaunchKernelGGL(k_first_list, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)ffirst.p,
                       (const uint32_t*)frank.p, ffirst_occ.p);
    out.f_tfstart.alloc(mem, NF + 1);
    fill_u32(out.f_tfstart.p, 0, NF + 1, s);
    hipLaunchKernelGGL(k_face_counts, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)perm.p,
                       (const uint32_t*)fhead.p, (const uint32_t*)fhex.p, (const uint32_t*)fsegfirst.p,
                       (const uint32_t*)frank.p, out.f_tfstart.p);
    scan_u32(mem, out.f_tfstart.p, out.f_tfstart.p, NF, out.f_tfstart.p + NF, s);
    out.n_tet_faces = NFO;
    out.f_tf_tet.alloc(mem, NFO + 1);
    out.f_tf_local.alloc(mem, NFO + 1);
    hipLaunchKernelGGL(k_face_tf, dim3(nblk(NFO)), dim3(256), 0, s, NFO, (const uint32_t*)perm.p,
                       (const uint32_t*)fhead.p, (const uint32_t*)fhex.p, (const uint32_t*)fsegfirst.p,
                       (const uint32_t*)fsegstart.p, (const uint32_t*)frank.p, (const uint32_t*)out.f_tfstart.p,
                       (const uint32_t*)out.act.p, (const uint32_t*)fq_a.p, (const uint16_t*)fq_l.p, out.f_tf_tet.p,
                       out.f_tf_local.p);
    out.f_vstart.alloc(mem, NF + 1);
    hipLaunchKernelGGL(k_face_vcount, dim3(nblk(NF)), dim3(256), 0, s, NF, (const uint32_t*)ffirst_occ.p,
                       (const uint32_t*)fq_vs.p, (const uint32_t*)fq_end.p, out.f_vstart.p);
    scan_u32(mem, out.f_vstart.p, out.f_vstart.p, NF, out.f_vstart.p + NF, s);
    out.n_face_verts = read1(out.f_vstart.p + NF, s);
    out.f_verts.alloc(mem, out.n_face_verts + 1);
    out.f_func.alloc(mem, 2 * NF + 2);
    hipLaunchKernelGGL(k_face_out, dim3(nblk(NF)), dim3(256), 0, s, NF, (const uint32_t*)ffirst_occ.p,
                       (const uint32_t*)fq_vs.p, (const uint32_t*)fq_end.p, (const uint32_t*)fve.p,
                       (const uint8_t*)fq_fn.p, (const uint32_t*)out.f_vstart.p, out.f_verts.p, out.f_func.p);
    RIN_HIP(hipGetLastError());
    tm.mark("faces");
    tm.finish(out.times);
    if (!(flags & RUN_KEEP_RECORDS)) {
        // records stay only when the host graph will ask for them
        out.rec.reset();
    }
}

// ------------------------------------------------------------------ diagnostics

namespace {
template <class C, bool MI>
__global__ void k_debug_cut(const double* rows, const int* ks, const int* starts, int ncase, int cap, uint16_t* words,
                            int* lens) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncase) return;
    const int k = ks[i];
    double P[C::MP - 4][4];
    for (int q = 0; q < k && q < C::MP - 4; ++q)
        for (int c = 0; c < 4; ++c) P[q][c] = rows[4 * (starts[i] + q) + c];
    Complex<C> X;
    if (k > C::MP - 4) {
        lens[i] = -2;
        return;
    }
    if (MI)
        dev::mi_cut<C>(P, k, X);
    else
        dev::ia_cut<C>(P, k, X);
    if (X.err) {
        lens[i] = -1;
        return;
    }
    lens[i] = MI ? dev::write_record_mi<C>(X, k, words + (size_t)i * cap, cap)
                 : dev::write_record_ia<C>(X, k, words + (size_t)i * cap, cap);
}
}  // namespace

void debug_cuts(DevMem& mem, int mi, const double* rows, const int* ks, int ncase, int cap, uint16_t* words, int* lens,
                hipStream_t s) {
    std::vector<int> starts(ncase);
    int total = 0;
    for (int i = 0; i < ncase; ++i) {
        starts[i] = total;
        total += ks[i];
    }
    DBuf<double> drows(mem, 4 * (size_t)total + 4);
    DBuf<int> dks(mem, ncase + 1), dst(mem, ncase + 1), dlens(mem, ncase + 1);
    DBuf<uint16_t> dw(mem, (size_t)ncase * cap + 1);
    RIN_HIP(hipMemcpyAsync(drows.p, rows, 32 * (size_t)total, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(dks.p, ks, 4 * (size_t)ncase, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(dst.p, starts.data(), 4 * (size_t)ncase, hipMemcpyHostToDevice, s));
    const unsigned nb = (unsigned)((ncase + 63) / 64);
    if (mi)
        hipLaunchKernelGGL((k_debug_cut<CBig, true>), dim3(nb), dim3(64), 0, s, (const double*)drows.p,
                           (const int*)dks.p, (const int*)dst.p, ncase, cap, dw.p, dlens.p);
    else
        hipLaunchKernelGGL((k_debug_cut<CBig, false>), dim3(nb), dim3(64), 0, s, (const double*)drows.p,
                           (const int*)dks.p, (const int*)dst.p, ncase, cap, dw.p, dlens.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(words, dw.p, 2 * (size_t)ncase * cap, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipMemcpyAsync(lens, dlens.p, 4 * (size_t)ncase, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

}  // namespace rin
