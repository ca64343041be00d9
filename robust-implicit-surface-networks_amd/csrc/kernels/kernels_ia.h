// Kernels of the device hot path shared by the plain and batched drivers.
// Included once per translation unit (anonymous namespace).
#pragma once
#include <type_traits>

#include "cut_dev.h"
#include "cut_wave_mi.h"
#include "pipeline.h"
#include "scan_sort.h"

namespace rin {
namespace {

using dev::Complex;
using CMed = dev::Caps<40, 64, 8, 8>;      // up to 4 active functions
using CBig = dev::Caps<132, 284, 10, 12>;  // up to 8 active functions
// more than 8 active functions (degenerate inputs: ties, coincident
// functions; many functions on coarse grids) and complexes beyond a bucket's
// capacities: up to 28 functions, 16-bit ids, the complex in device memory
using CHuge = dev::Caps<1023, 1500, 32, 32, uint16_t>;
constexpr int kMaskWords = (CHuge::MV + 63) / 64;  // vertex masks of any record
constexpr int kMaxHuge = CHuge::MP - 4;

constexpr uint32_t ERR_CAPACITY = 1, ERR_RECORD = 4;
// a bucket kernel's complex outgrew its capacities: the tet is cut again
// with the large-capacity kernel (plain form)
constexpr uint32_t ERR_RETRY = 16;
constexpr uint32_t CNT_RETRY = 0xFFFFFFFFu;

inline unsigned nblk(uint64_t n, unsigned t = 256) {
    uint64_t g = (n + t - 1) / t;
    if (g > 1u << 20) g = 1u << 20;
    return (unsigned)(g ? g : 1);
}

__device__ inline uint64_t count_of(uint64_t n) { return n; }
__device__ inline uint64_t count_of(const Count& n) { return n.get(); }
#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, i##_end = count_of(n); i < i##_end; \
         i += (uint64_t)gridDim.x * blockDim.x)

__device__ inline void flag_error(uint32_t* errs, uint32_t bit) { errs[threadIdx.x & 63] |= bit; }

// The chip's 8 XCDs each have their own L2, and a launch's workgroups go to
// them round-robin (workgroup b to XCD b mod 8).  Tiled kernels renumber
// their workgroups so that each XCD takes one contiguous range of the n live
// tiles: neighbouring tiles, which read the same lines (grid corners shared
// by neighbouring cubes, the star tets of a vertex, the occurrences of
// nearby tets), then meet in one L2 instead of being fetched by several.
// A bijection of [0, n) for b < n: XCD x takes n / 8 tiles, one more if
// x < n mod 8.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// One lane's thread-cut state (the complex and the insertion's working
// arrays) in a device-memory workspace: lane i of a launch owns bytes
// [i * stride, (i + 1) * stride).  Kept out of the private segment, so the
// runtime's per-queue scratch stays small and several queues can run cuts
// at once.
// lanes of a plain-form thread-cut launch (the grid strides beyond)
constexpr uint64_t kCutLanes = 16384;

template <class C>
struct LaneCut {
    static constexpr size_t stride =
        (sizeof(dev::Complex<C>) + sizeof(dev::InsertTemps<C>) + 255) & ~size_t(255);
    __device__ static dev::Complex<C>& X(uint8_t* ws, uint64_t lane) {
        return *reinterpret_cast<dev::Complex<C>*>(ws + lane * stride);
    }
    __device__ static dev::InsertTemps<C>& T(uint8_t* ws, uint64_t lane) {
        return *reinterpret_cast<dev::InsertTemps<C>*>(ws + lane * stride + sizeof(dev::Complex<C>));
    }
};

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

// function-major F x V -> row-major V x F (the reference's funcVals layout)
__global__ void k_transpose_out(const double* __restrict__ vals, uint64_t V, uint32_t F, double* __restrict__ rm) {
    GRID_STRIDE(i, V * F) {
        const uint64_t v = i / F, f = i % F;
        rm[i] = vals[f * V + v];
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

// degen: per vertex (null: not written); total (null: not counted): the
// number of tied vertices added per workgroup
__global__ void k_highest_mi(const double* __restrict__ vals, uint64_t V, uint32_t F, uint64_t* __restrict__ hm,
                             uint32_t* __restrict__ degen, uint32_t* __restrict__ total) {
    uint32_t ties = 0;
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
        if (degen) degen[v] = cnt > 1;
        ties += cnt > 1;
    }
    if (total) {
        __shared__ uint32_t red[4];
        for (int d = 32; d > 0; d >>= 1) ties += __shfl_xor(ties, d, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ties;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) t += red[w];
            if (t) atomicAdd(total, t);
        }
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
    const uint32_t ia[9] = {0, 128, 256, 448, 768, 1280, 1920, 2560, 3200};
    const uint32_t m[9] = {0, 64, 128, 256, 320, 400, 480, 560, 640};
    // beyond 8: everything a CHuge complex can hold
    constexpr uint32_t huge = 8 + 4 * CHuge::MV + CHuge::MF * (4 + CHuge::MFV) + 2 * CHuge::MP;
    return (k < 0 || k > 8) ? huge : (mi ? m[k] : ia[k]);
}

__global__ void k_rec_bound(Count n, const uint8_t* __restrict__ act_k, int mi, uint32_t* __restrict__ b) {
    GRID_STRIDE(a, n) b[a] = rec_bound(mi, act_k[a]);
}

// ------------------------------------------------------------------ per-tet cut

template <class C>
__device__ inline void ia_counts(const Complex<C>& X, uint32_t& niv, uint32_t& nif, uint32_t& nfv) {
    constexpr int W = (C::MV + 63) / 64;
    uint64_t iso[W];
    for (int w = 0; w < W; ++w) iso[w] = 0;
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
    niv = 0;
    for (int w = 0; w < W; ++w) niv += (uint32_t)__popcll(iso[w]);
}

template <class C>
__host__ __device__ inline void mi_counts(const Complex<C>& X, uint32_t& niv, uint32_t& nif, uint32_t& nfv) {
    // MI faces: positive label > 3; boundary faces are resolved across tets
    constexpr int W = (C::MV + 63) / 64;
    uint64_t iso[W];
    for (int w = 0; w < W; ++w) iso[w] = 0;
    nif = nfv = 0;
    for (int f = 0; f < X.nf; ++f) {
        if (X.fa[f] <= 3) continue;
        ++nif;
        nfv += X.fn[f];
        for (int q = 0; q < X.fn[f]; ++q) iso[X.fv[f][q] >> 6] |= 1ull << (X.fv[f][q] & 63);
    }
    niv = 0;
    for (int w = 0; w < W; ++w) niv += (uint32_t)__builtin_popcountll(iso[w]);
}

// One active tet's cut into its record slot.
template <class C, bool MI>
__device__ inline void cut_tet(const GridView& g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                               const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit, uint64_t a, int k,
                               uint32_t flags, const uint32_t* __restrict__ rec_off, uint16_t* __restrict__ rec,
                               uint32_t* __restrict__ cnt, uint32_t* __restrict__ errs, Complex<C>& X,
                               dev::InsertTemps<C>& Tm) {
    const uint64_t V = g.V;
    const uint32_t t = act[a];
    uint32_t v[4];
    dev::tet_corners(g, t, v);
    const uint32_t s0 = csr_off[t];
    double P[C::MP - 4][4];
    for (int q = 0; q < k; ++q) {
        const uint32_t f = fit[s0 + ((flags & RUN_REVERSE) ? k - 1 - q : q)];
        for (int c = 0; c < 4; ++c) P[q][c] = vals[(uint64_t)f * V + v[c]];
    }
    if (MI)
        dev::mi_cut<C>(P, k, X, Tm);
    else
        dev::ia_cut<C>(P, k, X, Tm);
    if (X.err) {
        // capacities or predicate range: cut again with the large kernel
        flag_error(errs, ERR_RETRY);
        cnt[3 * a] = CNT_RETRY;
        cnt[3 * a + 1] = cnt[3 * a + 2] = 0;
        rec[rec_off[a]] = 0;
        return;
    }
    const int cap = (int)(rec_off[a + 1] - rec_off[a]);
    const int w = MI ? dev::write_record_mi<C>(X, k, rec + rec_off[a], cap)
                     : dev::write_record_ia<C>(X, k, rec + rec_off[a], cap);
    if (w < 0) {
        flag_error(errs, ERR_RETRY);
        cnt[3 * a] = CNT_RETRY;
        cnt[3 * a + 1] = cnt[3 * a + 2] = 0;
        rec[rec_off[a]] = 0;
        return;
    }
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

template <class C, bool MI>
__global__ void __launch_bounds__(64) k_cut(GridView g, const double* __restrict__ vals, const uint32_t* __restrict__ act,
                                            const uint8_t* __restrict__ act_k, const uint32_t* __restrict__ csr_off,
                                            const uint8_t* __restrict__ fit, Count n_active, int kmin, int kmax,
                                            uint32_t flags, const uint32_t* __restrict__ rec_off,
                                            uint16_t* __restrict__ rec, uint32_t* __restrict__ cnt,
                                            uint32_t* __restrict__ errs, const uint32_t* __restrict__ list,
                                            uint8_t* __restrict__ ws) {
    const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Complex<C>& X = LaneCut<C>::X(ws, lane);
    dev::InsertTemps<C>& Tm = LaneCut<C>::T(ws, lane);
    GRID_STRIDE(i, n_active) {
        const uint64_t a = list ? list[i] : i;
        const int k = act_k[a];
        if (k < kmin || k > kmax) continue;
        cut_tet<C, MI>(g, vals, act, csr_off, fit, a, k, flags, rec_off, rec, cnt, errs, X, Tm);
    }
}

// The same with each lane's complex and working arrays in LDS (LANES lanes a
// block, blockDim.x == LANES): for short lists, where one lane's latency
// through its cut is the kernel's time and LDS answers several times faster
// than the device-memory workspace.
template <class C, bool MI, int LANES>
__global__ void __launch_bounds__(64) k_cut_lds(GridView g, const double* __restrict__ vals,
                                                const uint32_t* __restrict__ act, const uint8_t* __restrict__ act_k,
                                                const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit,
                                                Count n_active, int kmin, int kmax, uint32_t flags,
                                                const uint32_t* __restrict__ rec_off, uint16_t* __restrict__ rec,
                                                uint32_t* __restrict__ cnt, uint32_t* __restrict__ errs,
                                                const uint32_t* __restrict__ list) {
    __shared__ Complex<C> Xs[LANES];
    __shared__ dev::InsertTemps<C> Ts[LANES];
    const int l = threadIdx.x < LANES ? threadIdx.x : 0;
    if ((int)threadIdx.x >= LANES) return;
    Complex<C>& X = Xs[l];
    dev::InsertTemps<C>& Tm = Ts[l];
    const uint64_t n = count_of(n_active);
    for (uint64_t i = (uint64_t)blockIdx.x * LANES + l; i < n; i += (uint64_t)gridDim.x * LANES) {
        const uint64_t a = list ? list[i] : i;
        const int k = act_k[a];
        if (k < kmin || k > kmax) continue;
        cut_tet<C, MI>(g, vals, act, csr_off, fit, a, k, flags, rec_off, rec, cnt, errs, X, Tm);
    }
}

// Lane-group cuts over a list of active tets (64 / G tets a wavefront, the
// complex in LDS; cut_wave.h, cut_wave_mi.h), records at rec_off[a] in slots
// of rec_off[a + 1] - rec_off[a] words.  A tet whose complex outgrows the
// capacities is marked for the large-capacity retry (retry != 0) or sets the
// run's capacity error.  The grid strides over the list.
// k = 2 table, per first-plane sign pattern: the first slot of the pattern,
// the first plane's complex's vertex count and, for each of its vertices on a
// tet edge, that edge's corners.  A slot indexes the second plane's signs at
// those vertices (base-3 digits, corners first).
struct K2Pat {
    uint32_t base;  // first slot of this first-plane pattern
    uint8_t nv, valid, pad0, pad1;
    uint8_t e[4][2];  // corners of the tet edge under vertex 4 + q
};
static_assert(sizeof(K2Pat) == 16, "k2 pattern layout");
constexpr uint32_t K2_NONE = 0xFFFFFFFFu;
// snapshot layout of the complex after two planes: a 16-byte head, 3 bytes
// per vertex, 16 per face
constexpr int kK2SnapHead = 16, kK2SnapFace = 16;

// The lane-group cuts of 3..8 functions start from the complex after the
// tet's first two planes, the k = 2 snapshot of the tet's slot: two of their
// insertions become a table read.
struct K2Snap {
    const K2Pat* pat;
    const uint32_t* off;  // per slot: byte offset of the snapshot, or K2_NONE (null: start from the tet)
    const uint8_t* bytes;
};

// the slot of a tet whose first two rows are P[0], P[1], as k_k2_lookup
// computes it (corner signs and at most four exact edge predicates), or K2_NONE
__host__ __device__ __forceinline__ uint32_t k2_slot_of(const K2Pat* __restrict__ pats, const double (*P)[4]) {
    int pa = 0, sB[4];
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) {
        const double a = P[0][c], b = P[1][c];
        pa += (a > 0 ? 2 : (a < 0 ? 0 : 1)) * w;
        sB[c] = b > 0 ? 1 : (b < 0 ? -1 : 0);
    }
    const K2Pat K = pats[pa];
    if (!K.valid) return K2_NONE;
    uint32_t slot = K.base, w = 81;
    for (int c = 0, w3 = 1; c < 4; ++c, w3 *= 3) slot += (uint32_t)(sB[c] + 1) * w3;
    bool err = false;
    for (int q = 0; q + 4 < K.nv; ++q, w *= 3) {
        const int c1 = K.e[q][0], c2 = K.e[q][1];
        int sg;
        if (sB[c1] * sB[c2] < 0)
            sg = dev::orient_ia_edge(P[0][c1], P[0][c2], P[1][c1], P[1][c2], err);
        else
            sg = sB[c1] == sB[c2] ? sB[c1] : (sB[c1] != 0 ? sB[c1] : sB[c2]);
        slot += (uint32_t)(sg + 1) * w;
    }
    return err ? K2_NONE : slot;
}

// Loads the group's complex after the first two planes of its tet (rows in
// X.P, visible to the group) and returns true, or returns false (no snapshot:
// the cut starts from the tet).  Uniform over the group.
template <class C, int G>
__device__ __forceinline__ bool k2_start(dev::WaveIA<C>& X, const K2Snap& S, int k) {
    const int lane = dev::Grp<G>::lane();
    if (!S.off || k < 3 || 4 + k > C::MP) return false;
    const uint32_t slot = k2_slot_of(S.pat, X.P);
    if (slot == K2_NONE) return false;
    const uint32_t off = S.off[slot];
    if (off == K2_NONE) return false;
    const uint8_t* __restrict__ b = S.bytes + off;
    const int nv = b[0], nf = b[1];
    if (lane == 0) {
        X.nv = nv;
        X.nf = nf;
        X.nc = b[2];
        X.dup = b[3];
        X.err = 0;
    }
    if (lane < C::MP) {
        X.grp[lane] = lane < 6 ? b[4 + lane] : dev::NIL;
        X.ori[lane] = lane < 6 ? b[10 + lane] : (uint8_t)1;
    }
    const uint8_t* __restrict__ vb = b + kK2SnapHead;
    for (int v = lane; v < nv; v += G)
        for (int q = 0; q < 3; ++q) X.vl[v][q] = vb[3 * v + q];
    const uint8_t* __restrict__ fb = vb + 3 * nv;
    for (int f = lane; f < nf; f += G) {
        const uint8_t* __restrict__ r = fb + kK2SnapFace * f;
        X.fa[f] = r[0];
        X.fb[f] = r[1];
        X.fc[f] = r[2];
        X.fn[f] = r[3];
        for (int q = 0; q < C::MFV; ++q) {
            X.fv[f][q] = q < 6 ? r[4 + q] : (uint8_t)0;
            X.fe[f][q] = q < 6 ? r[10 + q] : (uint8_t)0;
        }
    }
    return true;
}

// 3 materials a < b < c (the material-interface table): per 2-material
// pattern, the first slot of its sign vectors and its vertices' labels
struct MI3Pat {
    uint32_t base;
    uint8_t nv, valid, pad0, pad1;
    uint8_t lab[8][4];
};
static_assert(sizeof(MI3Pat) == 40, "3-material pattern layout");
constexpr uint32_t MI3_NONE = 0xFFFFFFFFu;
// snapshot layout of the complex after three materials: a 20-byte head
// (nv, nf, nc, dup, grp[8], cells[8]), 4 bytes per vertex, 20 per face
// (fa, fb, fn, 0, fv[8], fe[8])
constexpr int kMI3SnapHead = 20, kMI3SnapFace = 20;

// The material-interface lane-group cut of 4..8 materials starts from the
// complex after the tet's first three, the 3-material table's snapshot of the
// tet's slot: three insertions become a table read.
struct MI3Snap {
    const MI3Pat* pat;
    const uint32_t* off;  // per slot: byte offset of the snapshot, or MI3_NONE (null: start from the tet)
    const uint8_t* bytes;
};

// the 3-material slot of a tet whose first three rows are P[0..2], as
// k_mi_table computes it, or MI3_NONE
__device__ __forceinline__ uint32_t mi3_slot_of(const MI3Pat* __restrict__ pats, const double (*P)[4]) {
    int pat = 0;
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) pat += (P[1][c] > P[0][c] ? 2 : (P[1][c] < P[0][c] ? 0 : 1)) * w;
    const MI3Pat& K = pats[pat];
    if (!K.valid) return MI3_NONE;
    bool err = false;
    uint32_t slot = K.base, w = 1;
    for (int q = 0; q < K.nv && !err; ++q, w *= 3) {
        const int lab[4] = {K.lab[q][0], K.lab[q][1], K.lab[q][2], K.lab[q][3]};
        slot += (uint32_t)(dev::orient_mi(P, lab, P[2], err) + 1) * w;
    }
    return err ? MI3_NONE : slot;
}

// Loads the group's complex after the first three materials of its tet (rows
// in X.P, visible to the group) and returns true, or returns false (the cut
// starts from the tet).  Uniform over the group.
template <class C, int G>
__device__ __forceinline__ bool mi3_start(dev::WaveMI<C>& X, const MI3Snap& S, int k) {
    const int lane = dev::Grp<G>::lane();
    if (!S.off || k < 4 || 4 + k > C::MP) return false;
    const uint32_t slot = mi3_slot_of(S.pat, X.P);
    if (slot == MI3_NONE) return false;
    const uint32_t off = S.off[slot];
    if (off == MI3_NONE) return false;
    const uint8_t* __restrict__ b = S.bytes + off;
    const int nv = b[0], nf = b[1];
    if (lane == 0) {
        X.nv = nv;
        X.nf = nf;
        X.nc = b[2];
        X.dup = b[3];
        X.err = 0;
    }
    if (lane < C::MP) {
        X.grp[lane] = lane < 8 ? b[4 + lane] : dev::NIL;
        X.cells[lane] = lane < 8 ? b[12 + lane] : dev::NIL;
    }
    const uint8_t* __restrict__ vb = b + kMI3SnapHead;
    for (int v = lane; v < nv; v += G)
        for (int q = 0; q < 4; ++q) X.vl[v][q] = vb[4 * v + q];
    const uint8_t* __restrict__ fb = vb + 4 * nv;
    for (int f = lane; f < nf; f += G) {
        const uint8_t* __restrict__ r = fb + kMI3SnapFace * f;
        X.fa[f] = r[0];
        X.fb[f] = r[1];
        X.fn[f] = r[2];
        for (int q = 0; q < C::MFV; ++q) {
            X.fv[f][q] = q < 8 ? r[4 + q] : (uint8_t)0;
            X.fe[f][q] = q < 8 ? r[12 + q] : (uint8_t)0;
        }
    }
    return true;
}

template <class C, int G, bool MI>
__global__ void __launch_bounds__(64) k_cut_group(GridView g, const double* __restrict__ vals,
                                                  const uint32_t* __restrict__ act, const uint8_t* __restrict__ act_k,
                                                  const uint32_t* __restrict__ csr_off,
                                                  const uint8_t* __restrict__ fit, const uint32_t* __restrict__ list,
                                                  Count nlist, const uint32_t* __restrict__ rec_off,
                                                  uint16_t* __restrict__ rec, uint32_t* __restrict__ cnt,
                                                  uint32_t* __restrict__ errs, int retry,
                                                  typename std::conditional<MI, MI3Snap, K2Snap>::type snap) {
    constexpr int NG = 64 / G;
    using Cx = typename std::conditional<MI, dev::WaveMI<C>, dev::WaveIA<C>>::type;
    __shared__ Cx Xs[NG];
    const int gid = threadIdx.x / G, lane = threadIdx.x & (G - 1);
    Cx& X = Xs[gid];
    const uint64_t n = count_of(nlist);
    const uint64_t V = g.V;
    for (uint64_t i = (uint64_t)blockIdx.x * NG + gid; i < n; i += (uint64_t)gridDim.x * NG) {
        const uint32_t a = list[i];
        const int k = act_k[a];
        const uint32_t t = act[a], s0 = csr_off[t];
        if (lane < 4 * k && 4 + k <= C::MP) {
            uint32_t v[4];
            dev::tet_corners(g, t, v);
            const uint32_t f = fit[s0 + (lane >> 2)];
            X.P[lane >> 2][lane & 3] = vals[(uint64_t)f * V + v[lane & 3]];
        }
        __syncthreads();
        if constexpr (MI) {
            const int start = mi3_start<C, G>(X, snap, k) ? 3 : 1;
            dev::wave_mi_cut<C, G>(X, k, start);
        } else {
            const int start = k2_start<C, G>(X, snap, k) ? 2 : 0;
            dev::wave_ia_cut<C, G>(X, k, start);
        }
        uint16_t* out = rec + rec_off[a];
        const int cap = (int)(rec_off[a + 1] - rec_off[a]);
        uint32_t a0 = 0, a1 = 0, a2 = 0;
        int len = -1;
        if (!X.err) {
            if constexpr (MI)
                len = dev::wave_mi_record<C, G>(X, k, out, cap, a0, a1, a2);
            else
                len = dev::wave_ia_record<C, G>(X, k, out, cap, a0, a1, a2);
        }
        if (lane == 0) {
            if (X.err || len < 0) {
                if (retry) {
                    atomicOr(errs, (uint32_t)ERR_RETRY);
                    a0 = CNT_RETRY;
                } else {
                    atomicOr(errs, (uint32_t)ERR_CAPACITY);
                    a0 = 0;
                }
                out[0] = 0;
                a1 = a2 = 0;
            }
            cnt[3 * a] = a0;
            cnt[3 * a + 1] = a1;
            cnt[3 * a + 2] = a2;
        }
        __syncthreads();  // the group's complex is reused by its next tet
    }
}

// active tets with more than 8 functions, and tets a bucket kernel could not
// hold, one lane each; ws holds one complex
// per lane of the grid
template <bool MI>
__global__ void __launch_bounds__(64) k_cut_huge(GridView g, const double* __restrict__ vals,
                                                 const uint32_t* __restrict__ act, const uint8_t* __restrict__ act_k,
                                                 const uint32_t* __restrict__ csr_off, const uint8_t* __restrict__ fit,
                                                 const uint32_t* __restrict__ list, Count n, uint32_t flags,
                                                 const uint32_t* __restrict__ rec_off, uint16_t* __restrict__ rec,
                                                 uint32_t* __restrict__ cnt, uint32_t* __restrict__ errs,
                                                 uint8_t* __restrict__ ws, const uint32_t* __restrict__ rec_loc,
                                                 uint32_t fixed_cap) {
    using C = CHuge;
    const uint64_t V = g.V;
    const uint64_t lane = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Complex<C>& X = LaneCut<C>::X(ws, lane);
    dev::InsertTemps<C>& Tm = LaneCut<C>::T(ws, lane);
    GRID_STRIDE(i, n) {
        const uint32_t a = list[i];
        const int k = act_k[a];
        const uint32_t t = act[a];
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        double P[C::MP - 4][4];
        const int kk = k < C::MP - 4 ? k : C::MP - 4;
        for (int q = 0; q < kk; ++q) {
            const uint32_t f = fit[s0 + ((flags & RUN_REVERSE) ? k - 1 - q : q)];
            for (int c = 0; c < 4; ++c) P[q][c] = vals[(uint64_t)f * V + v[c]];
        }
        if (MI)
            dev::mi_cut<C>(P, k, X, Tm);
        else
            dev::ia_cut<C>(P, k, X, Tm);
        // the record goes to rec_loc[a] (retries) or to the tet's own slot
        const uint32_t at = rec_loc ? rec_loc[a] : rec_off[a];
        if (X.err) {
            flag_error(errs, ERR_CAPACITY);
            cnt[3 * a] = cnt[3 * a + 1] = cnt[3 * a + 2] = 0;
            rec[at] = 0;
            continue;
        }
        const int cap = fixed_cap ? (int)fixed_cap : (int)(rec_off[a + 1] - rec_off[a]);
        const int w = MI ? dev::write_record_mi<C>(X, k, rec + at, cap) : dev::write_record_ia<C>(X, k, rec + at, cap);
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

__global__ void k_split3(Count n, const uint32_t* __restrict__ cnt, uint32_t* a, uint32_t* b, uint32_t* c) {
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
__device__ inline uint64_t key_inner(uint64_t a, uint32_t j) { return ((a << 10 | j) << 2) | 3; }

// Key of the iso vertex j of active tet a (tet t, corners v) given its plane
// triple (ids 0..3 tet faces, 4.. implicit planes in CSR order from s0):
// three tet faces meet at a corner, two at an edge, one at a face, none
// inside.  Bit logic instead of indexed local arrays keeps everything in
// registers.
__device__ inline uint64_t ia_vertex_key(const GridView& g, uint64_t t, const uint32_t v[4], uint64_t a, int j,
                                         int t0, int t1, int t2, const uint8_t* __restrict__ fit, uint32_t s0) {
    unsigned bm = 0;
    if (t0 < 4) bm |= 1u << t0;
    if (t1 < 4) bm |= 1u << t1;
    if (t2 < 4) bm |= 1u << t2;
    const int nb = __popc(bm);
    // implicit planes in triple order
    const int i0 = t0 >= 4 ? t0 : (t1 >= 4 ? t1 : t2);
    const int i1 = (t0 >= 4 && t1 >= 4) ? t1 : t2;
    if (nb == 3) return key_corner(v[__builtin_ctz(~bm & 15u)]);
    if (nb == 2) {
        const unsigned m = ~bm & 15u;
        const int l0 = __builtin_ctz(m), l1 = __builtin_ctz(m & (m - 1));
        return key_edge(dev::edge_id(g, t, v, l0, l1), fit[s0 + i0 - 4]);
    }
    if (nb == 1) return key_face(dev::face_id(g, t, v, __builtin_ctz(bm)), fit[s0 + i0 - 4], fit[s0 + i1 - 4]);
    return key_inner(a, j);
}

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
                             const uint8_t* __restrict__ fit, Count n_active, const uint32_t* __restrict__ rec_loc,
                             const uint16_t* __restrict__ rec, const uint32_t* __restrict__ voff,
                             const uint32_t* __restrict__ foff, const uint32_t* __restrict__ fvoff,
                             uint64_t* __restrict__ okey, uint32_t* __restrict__ oa, uint16_t* __restrict__ oj,
                             uint32_t* __restrict__ fq_a, uint16_t* __restrict__ fq_l, uint8_t* __restrict__ fq_fn,
                             uint8_t* __restrict__ fq_bnd, uint32_t* __restrict__ fq_vs, uint32_t* __restrict__ fve) {
    GRID_STRIDE(a, n_active) {
        const uint32_t t = act[a];
        const Rec R(rec + rec_loc[a]);
        if (R.r[0] != 1) continue;
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = csr_off[t];
        uint64_t iso[kMaskWords];
        for (int w = 0; w < kMaskWords; ++w) iso[w] = 0;
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
            const uint64_t key =
                ia_vertex_key(g, t, v, a, j, T3[0], T3[1], T3[2], fit, s0);
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
                fq_bnd[q] = F[2] == 0xFFFF ? (uint8_t)(1 + F[0]) : 0;
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

__global__ void k_heads(Count n, const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                        uint32_t* __restrict__ first, uint32_t* __restrict__ head) {
    GRID_STRIDE(p, n) {
        const uint32_t h = (p == 0 || skey[p] != skey[p - 1]) ? 1u : 0u;
        head[p] = h;
        first[sidx[p]] = h;
    }
}

// head_excl: exclusive scan of head over sorted positions; rank: exclusive
// scan of `first` over occurrence order
__global__ void k_assign(Count n, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ head,
                         const uint32_t* __restrict__ head_excl, uint32_t* __restrict__ seg_first_occ) {
    GRID_STRIDE(p, n) {
        if (head[p]) seg_first_occ[head_excl[p]] = sidx[p];
    }
}
__global__ void k_assign2(Count n, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ head,
                          const uint32_t* __restrict__ head_excl, const uint32_t* __restrict__ seg_first_occ,
                          const uint32_t* __restrict__ rank, uint32_t* __restrict__ id_of_occ) {
    GRID_STRIDE(p, n) {
        const uint32_t seg = head_excl[p] + head[p] - 1;
        id_of_occ[sidx[p]] = rank[seg_first_occ[seg]];
    }
}
__global__ void k_first_list(Count n, const uint32_t* __restrict__ first, const uint32_t* __restrict__ rank,
                             uint32_t* __restrict__ first_occ) {
    GRID_STRIDE(o, n) {
        if (first[o]) first_occ[rank[o]] = (uint32_t)o;
    }
}

// ------------------------------------------------------------------ iso-edges
// compute_mesh_edges (src/mesh_connectivity.cpp:10-56) on the device: face
// edge j of face f joins face vertices j and j + 1 (cyclically); its key is
// the sorted vertex pair; edge ids are the ranks of first occurrences in
// (face, j) order, and each edge lists its occurrences in that order.

// occurrence o (a face-vertex position) -> its face, and the edge key
__global__ void k_edge_keys(Count nf, const uint32_t* __restrict__ vstart, const uint32_t* __restrict__ verts,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ face_of) {
    GRID_STRIDE(f, nf) {
        const uint32_t b = vstart[f], e = vstart[f + 1];
        for (uint32_t o = b; o < e; ++o) {
            const uint64_t v1 = verts[o], v2 = verts[o + 1 == e ? b : o + 1];
            key[o] = v1 < v2 ? (v1 << 32 | v2) : (v2 << 32 | v1);
            face_of[o] = (uint32_t)f;
        }
    }
}

// per segment of equal keys (sorted positions): where it starts, and for its
// edge the endpoints and the occurrence count
__global__ void k_edge_segments(Count n, const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
                                const uint32_t* __restrict__ head, const uint32_t* __restrict__ head_excl,
                                const uint32_t* __restrict__ id_of_occ, uint32_t* __restrict__ seg_pos,
                                uint32_t* __restrict__ v12, uint32_t* __restrict__ count) {
    const uint64_t nn = count_of(n);
    GRID_STRIDE(p, n) {
        if (!head[p]) continue;
        const uint32_t g = head_excl[p];
        seg_pos[g] = (uint32_t)p;
        uint64_t q = p + 1;
        while (q < nn && !head[q]) ++q;
        const uint32_t e = id_of_occ[sidx[p]];
        v12[2 * (uint64_t)e] = (uint32_t)(skey[p] >> 32);
        v12[2 * (uint64_t)e + 1] = (uint32_t)skey[p];
        count[e] = (uint32_t)(q - p);
    }
}

// each occurrence into its edge's list, in occurrence order (the sort is stable)
__global__ void k_edge_fill(Count n, const uint32_t* __restrict__ sidx, const uint32_t* __restrict__ head,
                            const uint32_t* __restrict__ head_excl, const uint32_t* __restrict__ seg_pos,
                            const uint32_t* __restrict__ id_of_occ, const uint32_t* __restrict__ start,
                            uint32_t* __restrict__ occ) {
    GRID_STRIDE(p, n) {
        const uint32_t g = head_excl[p] + head[p] - 1;
        const uint32_t o = sidx[p];
        occ[start[id_of_occ[o]] + (uint32_t)(p - seg_pos[g])] = o;
    }
}

// ------------------------------------------------------------------ patches
// compute_patches (src/mesh_connectivity.cpp:58-93) on the device: a patch is
// a component of the faces joined across manifold edges (edges with exactly
// two occurrences), seeded at its smallest face and listed in breadth-first
// order (queue order, neighbours in edge order).  Components come from
// union-find (a root hooks under the smaller root, so each component's root
// is its smallest face); the breadth-first order from level-synchronous
// expansion where each newly reached face keeps the smallest (position of
// its parent in the frontier, edge index) that reaches it, exactly the queue
// order of one search per patch.

// the face across face corner position o's edge, or ~0u when the edge is
// not manifold
__global__ void k_face_adj(Count no, const uint32_t* __restrict__ e_of_occ, const uint32_t* __restrict__ e_start,
                           const uint32_t* __restrict__ e_occ, const uint32_t* __restrict__ face_of,
                           uint32_t* __restrict__ adj) {
    GRID_STRIDE(o, no) {
        const uint32_t e = e_of_occ[o];
        const uint32_t b = e_start[e];
        uint32_t other = ~0u;
        if (e_start[e + 1] - b == 2) {
            const uint32_t f0 = face_of[e_occ[b]], f1 = face_of[e_occ[b + 1]];
            other = f0 == face_of[o] ? f1 : f0;
        }
        adj[o] = other;
    }
}

__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        const uint32_t g = parent[p];
        if (g != p) atomicMin(&parent[x], g);  // path halving (parents only decrease)
        x = p;
        p = g;
    }
    return x;
}

// joins the components of a and b: the larger root hooks under the smaller,
// so every component's root is its smallest element
__device__ inline void uf_union(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        if (atomicCAS(&parent[hi], hi, lo) == hi) return;
    }
}

__global__ void k_uf_hook(Count no, const uint32_t* __restrict__ face_of, const uint32_t* __restrict__ adj,
                          uint32_t* parent) {
    GRID_STRIDE(o, no) {
        const uint32_t v = adj[o];
        if (v != ~0u) uf_union(parent, face_of[o], v);
    }
}

// roots (after hooking): each face's component root, and seed flags
__global__ void k_uf_roots(Count nf, uint32_t* parent, uint32_t* __restrict__ seed) {
    GRID_STRIDE(f, nf) {
        const uint32_t r = uf_find(parent, (uint32_t)f);
        parent[f] = r;
        seed[f] = r == f ? 1u : 0u;
    }
}

// Chains (compute_chains, src/mesh_connectivity.cpp:196-241): non-manifold
// edges (more than two face occurrences) join across the vertices that carry
// exactly two of them, so every component is a path or a cycle.  Its seed is
// its smallest edge, and its order the breadth-first one from the seed,
// neighbour across the first vertex before the one across the second.
__device__ inline bool edge_non_manifold(const uint32_t* __restrict__ e_start, uint32_t i) {
    return e_start[i + 1] - e_start[i] > 2;
}

// every non-manifold edge registers at its two vertices (the first two kept)
__global__ void k_chain_vert(Count ne, const uint32_t* __restrict__ v12, const uint32_t* __restrict__ e_start,
                             uint32_t* __restrict__ vdeg, uint32_t* __restrict__ vslot) {
    GRID_STRIDE(i, ne) {
        if (!edge_non_manifold(e_start, (uint32_t)i)) continue;
        for (int s = 0; s < 2; ++s) {
            const uint32_t v = v12[2 * i + s];
            const uint32_t k = atomicAdd(&vdeg[v], 1u);
            if (k < 2) vslot[2 * (uint64_t)v + k] = (uint32_t)i;
        }
    }
}

// each edge's neighbours across its first and second vertex (~0u: none), its
// union-find parent and a cleared visit flag
__global__ void k_chain_links(Count ne, const uint32_t* __restrict__ v12, const uint32_t* __restrict__ e_start,
                              const uint32_t* __restrict__ vdeg, const uint32_t* __restrict__ vslot,
                              uint32_t* __restrict__ nb, uint32_t* __restrict__ parent, uint32_t* __restrict__ seen) {
    GRID_STRIDE(i, ne) {
        uint32_t n[2] = {~0u, ~0u};
        if (edge_non_manifold(e_start, (uint32_t)i))
            for (int s = 0; s < 2; ++s) {
                const uint32_t v = v12[2 * i + s];
                if (vdeg[v] != 2) continue;  // a chain ends at a vertex of one or of three or more
                const uint32_t a = vslot[2 * (uint64_t)v], b = vslot[2 * (uint64_t)v + 1];
                n[s] = a == (uint32_t)i ? b : a;
            }
        nb[2 * i] = n[0];
        nb[2 * i + 1] = n[1];
        parent[i] = (uint32_t)i;
        seen[i] = 0;
    }
}

__global__ void k_chain_hook(Count nl, const uint32_t* __restrict__ nb, uint32_t* parent) {
    GRID_STRIDE(o, nl) {
        const uint32_t v = nb[o];
        if (v != ~0u) uf_union(parent, (uint32_t)(o >> 1), v);
    }
}

// roots, and seed flags: non-manifold edges that are their component's root
__global__ void k_chain_roots(Count ne, const uint32_t* __restrict__ e_start, uint32_t* parent,
                              uint32_t* __restrict__ seed) {
    GRID_STRIDE(i, ne) {
        const uint32_t r = uf_find(parent, (uint32_t)i);
        parent[i] = r;
        seed[i] = (r == (uint32_t)i && edge_non_manifold(e_start, (uint32_t)i)) ? 1u : 0u;
    }
}

// chain sizes, and each chain's seed
__global__ void k_chain_count(Count ne, const uint32_t* __restrict__ e_start, const uint32_t* __restrict__ root,
                              const uint32_t* __restrict__ seed, const uint32_t* __restrict__ srank,
                              uint32_t* __restrict__ size, uint32_t* __restrict__ seeds) {
    GRID_STRIDE(i, ne) {
        if (!edge_non_manifold(e_start, (uint32_t)i)) continue;
        atomicAdd(&size[srank[root[i]]], 1u);
        if (seed[i]) seeds[srank[i]] = (uint32_t)i;
    }
}

// one lane per chain: the breadth-first walk from the seed.  The queue never
// holds more than two edges (the seed adds at most two, every later edge at
// most one: one of its neighbours is the edge it was reached from); err
// counts chains whose walk does not fill their slice exactly.
__global__ void k_chain_walk(Count nc, const uint32_t* __restrict__ seeds, const uint32_t* __restrict__ start,
                             const uint32_t* __restrict__ nb, uint32_t* __restrict__ seen,
                             uint32_t* __restrict__ out, uint32_t* __restrict__ err) {
    GRID_STRIDE(c, nc) {
        const uint32_t s = seeds[c], end = start[c + 1];
        uint32_t o = start[c];
        bool bad = false;
        out[o++] = s;
        seen[s] = 1;
        uint32_t qa = s, qb = 0;
        int qn = 1;
        while (qn > 0 && !bad) {
            const uint32_t e = qa;
            qa = qb;
            --qn;
            for (int k = 0; k < 2; ++k) {
                const uint32_t x = nb[2 * (uint64_t)e + k];
                if (x == ~0u || seen[x]) continue;
                if (o >= end || qn == 2) {
                    bad = true;
                    break;
                }
                seen[x] = 1;
                out[o++] = x;
                if (qn == 0)
                    qa = x;
                else
                    qb = x;
                ++qn;
            }
        }
        if (bad || o != end) atomicAdd(err, 1u);
    }
}

// the first frontier (the seeds, in face order) and every face's patch id
__global__ void k_patch_seeds(Count nf, const uint32_t* __restrict__ root, const uint32_t* __restrict__ seed,
                              const uint32_t* __restrict__ seed_rank, uint32_t* __restrict__ pid,
                              uint32_t* __restrict__ level, uint32_t* __restrict__ frontier,
                              unsigned long long* __restrict__ claim) {
    GRID_STRIDE(f, nf) {
        pid[f] = seed_rank[root[f]];
        level[f] = seed[f] ? 0u : ~0u;
        claim[f] = ~0ull;
        if (seed[f]) frontier[seed_rank[f]] = (uint32_t)f;
    }
}

// one level: every frontier face (position i) offers its manifold
// neighbours not reached yet the key (i, corner j); the first offer to a face
// appends it to the next frontier
__global__ void k_patch_claim(const uint32_t* __restrict__ nfront, const uint32_t* __restrict__ frontier,
                              const uint32_t* __restrict__ vstart, const uint32_t* __restrict__ adj,
                              const uint32_t* __restrict__ level, unsigned long long* __restrict__ claim,
                              uint32_t* __restrict__ next, uint32_t* __restrict__ nnext) {
    const uint32_t n = *nfront;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = frontier[i];
        const uint32_t b = vstart[p], e = vstart[p + 1];
        for (uint32_t o = b; o < e; ++o) {
            const uint32_t c = adj[o];
            if (c == ~0u || level[c] != ~0u) continue;
            const unsigned long long key = (unsigned long long)i << 16 | (o - b);
            if (atomicMin(&claim[c], key) == ~0ull) next[atomicAdd(nnext, 1u)] = c;
        }
    }
}

// the next frontier: (claim key, face) pairs to sort
__global__ void k_patch_keys(const uint32_t* __restrict__ nnext, const uint32_t* __restrict__ next,
                             const unsigned long long* __restrict__ claim, uint64_t* __restrict__ key,
                             uint32_t* __restrict__ val) {
    const uint32_t n = *nnext;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = next[i];
        key[i] = claim[c];
        val[i] = c;
    }
}

// after the sort: the level of each reached face, and the frontier appended
// to the breadth-first list
__global__ void k_patch_level(const uint32_t* __restrict__ nnext, const uint32_t* __restrict__ sorted, uint32_t lv,
                              uint32_t* __restrict__ level, uint32_t* __restrict__ order, uint32_t base) {
    const uint32_t n = *nnext;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = sorted[i];
        level[c] = lv;
        order[base + i] = c;
    }
}

// patch sizes from the patch ids, and the faces by patch (the breadth-first
// list is level-major and patch-major within a level, so a stable placement
// by patch keeps each patch's queue order)
__global__ void k_patch_count(Count nf, const uint32_t* __restrict__ pid, uint32_t* __restrict__ cnt) {
    GRID_STRIDE(f, nf) atomicAdd(&cnt[pid[f]], 1u);
}
__global__ void k_patch_keys_pid(Count nf, const uint32_t* __restrict__ order, const uint32_t* __restrict__ pid,
                                 uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    GRID_STRIDE(i, nf) {
        const uint32_t f = order[i];
        key[i] = pid[f];
        val[i] = f;
    }
}

// First-occurrence dedup.  Vertex ids are the ranks of first occurrences in
// occurrence order, as a stable sort by key gives; every occurrence gets the
// index `mo` of its key's first occurrence.
//
// Regular grids resolve almost every key without atomics: an edge, face or
// corner key names a simplex of the grid, every tet of that simplex's star
// holds the key (its function is active there), and occurrences run in tet
// order, so the first occurrence lies in the lowest tet of the star.  A
// host-built table lists the star of each simplex shape (the lattice
// directions from its lowest vertex, by that vertex's parity) as cube offsets
// and tets, in tet order; the first in-grid entry is that lowest tet, whose
// occurrences are searched for the key.  Interior keys are unique.  A key
// whose lowest star tet does not hold it (only degenerate inputs) goes, with
// every other occurrence of it, to the hash path below.
//
// Hash path: the table holds key + 1 (0: empty) and ~(smallest occurrence
// index) (0: none) per slot, linear probing; it is zero between runs (the
// assign kernel clears the slots a run used, one per distinct key).
__device__ inline uint32_t hash_slot(uint64_t k, int bits) {
    return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

__device__ inline void hash_insert_one(uint32_t o, const uint64_t* __restrict__ okey, unsigned long long* hkey,
                                       uint32_t* hmx, int bits, uint32_t* hslot, uint32_t* errs) {
    const uint32_t mask = (uint32_t)((1ull << bits) - 1);
    const unsigned long long k = okey[o] + 1;
    uint32_t h = hash_slot(k, bits);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long prev = atomicCAS(&hkey[h], 0ull, k);
        if (prev == 0 || prev == k) {
            atomicMax(&hmx[h], ~o);
            hslot[o] = h;
            return;
        }
        h = (h + 1) & mask;
    }
    // full table (sized at twice the occurrence capacity: not reached)
    flag_error(errs, ERR_CAPACITY);
    hslot[o] = 0;
}

constexpr int kStarShapes = 1 + 27 + 27 * 27;  // corner, edge by direction, face by two directions

struct StarArgs {
    int res;
    const uint64_t* okey;
    const uint32_t *oa, *act, *rank, *voff;
    const uint32_t* star_head;  // [2][kStarShapes]: first entry << 8 | entry count
    const uint8_t* star_ent;    // (di+1) | (dj+1) << 2 | (dk+1) << 4 with di.. in {-1, 0}, then q
    uint32_t *first, *mo, *fail, *n_fail;
    uint32_t cap_fail;
    const uint32_t* n_active;
    uint32_t* tile_first;  // first occurrences per tile of kDedupTile occurrences
};

// Occurrences per lane of the star pass, handled side by side so that their
// chains of dependent loads (key, star tet, its rank and offsets, its keys)
// overlap; a workgroup takes one tile of consecutive occurrences, lane-major
// within each of the kDedupU rows.
constexpr int kDedupU = 2;  // measured against 1, 4 and 8 occurrences a lane
constexpr uint32_t kDedupTile = 256u * kDedupU;

// lowest tet of the star of key K's simplex inside the grid (~0: none)
__device__ inline uint64_t star_lowest_tet(const StarArgs& A, uint32_t R, uint32_t N, uint64_t K) {
    const uint32_t type = (uint32_t)(K & 3);
    uint32_t a, shape;
    if (type == 0) {
        a = (uint32_t)(K >> 2);
        shape = 0;
    } else if (type == 1) {
        const uint64_t e = K >> 8;
        a = (uint32_t)(e / 27);
        shape = 1 + (uint32_t)(e % 27);
    } else {
        const uint64_t fc = K >> 14;
        a = (uint32_t)(fc / 729);
        shape = 28 + (uint32_t)(fc % 729);
    }
    const uint32_t ai = a / (N * N), ar = a - ai * N * N, aj = ar / N, ak = ar - aj * N;
    const uint32_t head = A.star_head[((ai + aj + ak) & 1u) * kStarShapes + shape];
    for (uint32_t e = head >> 8, end = (head >> 8) + (head & 255u); e < end; e += 2) {
        const uint32_t d = A.star_ent[e];
        const int ci = (int)ai + (int)(d & 3u) - 1, cj = (int)aj + (int)((d >> 2) & 3u) - 1,
                  ck = (int)ak + (int)((d >> 4) & 3u) - 1;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= (int)R || cj >= (int)R || ck >= (int)R) continue;
        return (((uint64_t)ci * R + cj) * R + ck) * 5 + A.star_ent[e + 1];
    }
    return ~0ull;
}

__global__ void __launch_bounds__(256) k_dedup_star(Count n, StarArgs A) {
    const uint32_t R = (uint32_t)A.res, N = R + 1;
    const uint32_t NO = (uint32_t)count_of(n);
    const uint32_t nt = (NO + kDedupTile - 1) / kDedupTile;
    if (blockIdx.x >= nt) return;
    const uint32_t tile = xcd_tile(blockIdx.x, nt);
    const uint32_t base = tile * kDedupTile;
    const uint32_t NA = *A.n_active;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t K[kDedupU], ts[kDedupU];
    uint32_t o[kDedupU], oa[kDedupU], t[kDedupU], as[kDedupU], lo[kDedupU], hi[kDedupU], found[kDedupU];
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        o[u] = base + u * 256u + threadIdx.x;
        const bool in = o[u] < NO;
        K[u] = in ? A.okey[o[u]] : 3ull;  // out of range: an interior key, decided without loads
        oa[u] = in ? A.oa[o[u]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        const bool inner = (K[u] & 3) == 3;
        t[u] = inner ? 0u : A.act[oa[u]];
        ts[u] = inner ? ~0ull : star_lowest_tet(A, R, N, K[u]);
    }
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        // interior keys, and keys whose lowest star tet is this one, are first here
        found[u] = ((K[u] & 3) == 3 || ts[u] == t[u]) ? o[u] : ~0u;
        as[u] = (found[u] == ~0u && ts[u] != ~0ull) ? A.rank[ts[u]] : ~0u;
    }
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        const bool ok = as[u] < NA && A.act[as[u]] == (uint32_t)ts[u];
        lo[u] = ok ? A.voff[as[u]] : 0u;
        hi[u] = ok ? (as[u] + 1 < NA ? A.voff[as[u] + 1] : NO) : 0u;
        hi[u] = hi[u] < NO ? hi[u] : NO;
    }
    // the lowest tet's keys, eight at a time (a tet rarely has more)
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        for (uint32_t p0 = lo[u]; p0 < hi[u] && found[u] == ~0u; p0 += 8) {
            uint64_t kk[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) kk[j] = p0 + j < hi[u] ? A.okey[p0 + j] : ~K[u];
#pragma unroll
            for (int j = 7; j >= 0; --j)
                if (kk[j] == K[u]) found[u] = p0 + j;
        }
    }
    uint32_t nfirst = 0;
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        const bool in = o[u] < NO;
        const bool fail = in && found[u] == ~0u;
        if (in) {
            A.first[o[u]] = found[u] == o[u];
            if (!fail) A.mo[o[u]] = found[u];
        }
        nfirst += in && found[u] == o[u];
        // failures: one counter update per wave
        const uint64_t m = __ballot(fail);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t b0 = 0;
            if (lane == leader) b0 = atomicAdd(A.n_fail, (uint32_t)__popcll(m));
            b0 = __shfl(b0, leader, 64);
            if (fail) {
                const uint32_t slot = b0 + (uint32_t)__popcll(m & ((1ull << lane) - 1));
                if (slot < A.cap_fail) A.fail[slot] = o[u];
            }
        }
    }
    __shared__ uint32_t red[256 / 64];
    for (int d = 32; d > 0; d >>= 1) nfirst += __shfl_xor(nfirst, d, 64);
    if (lane == 0) red[wave] = nfirst;
    __syncthreads();
    if (threadIdx.x == 0) A.tile_first[tile] = red[0] + red[1] + red[2] + red[3];
}

// Exclusive prefixes of per-tile counts in place, one workgroup; the tiles
// are those of `tile` elements over n, field f of tile i at x[i * stride + f]
// (f < nfield <= 4); *total[f] = the field's sum.
struct TileTotals {
    uint32_t* p[4];
};

// one workgroup of 1024 lanes; the counts are read with device-scope loads,
// so that counts other workgroups of this kernel added atomically are seen
__device__ inline void scan_tiles_block(uint32_t nt, uint32_t* __restrict__ x, int stride, int nfield,
                                        const TileTotals& total) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int f = 0; f < nfield; ++f) {
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (uint32_t base = 0; base < nt; base += 4 * 1024) {
            const uint32_t i0 = base + 4 * threadIdx.x;
            uint32_t v[4];
            for (int q = 0; q < 4; ++q)
                v[q] = i0 + q < nt ? __hip_atomic_load(&x[(uint64_t)(i0 + q) * stride + f], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT)
                                   : 0u;
            const uint32_t sum = v[0] + v[1] + v[2] + v[3];
            uint32_t inc = sum;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(inc, d, 64);
                if (lane >= d) inc += y;
            }
            if (lane == 63) wsum[wave] = inc;
            __syncthreads();
            uint32_t before = carry;
            for (int w = 0; w < 16; ++w) before += w < wave ? wsum[w] : 0u;
            uint32_t run = before + inc - sum;
            for (int q = 0; q < 4; ++q) {
                if (i0 + q < nt) x[(uint64_t)(i0 + q) * stride + f] = run;
                run += v[q];
            }
            __syncthreads();  // every lane has read carry and wsum
            if (threadIdx.x == 1023) carry = run;
            __syncthreads();
        }
        if (threadIdx.x == 0) *total.p[f] = carry;
    }
}

// Star path, after the star pass, in one workgroup: the hash fallback for
// the occurrences the star tables left (only degenerate inputs leave any; a
// long list is taken in passes), their first occurrences joining the tile
// counts, then the scan of the tile counts (*total: the vertex count).
__global__ void __launch_bounds__(1024) k_dedup_fallback_scan(Count n_fail, const uint32_t* __restrict__ list,
                                                              const uint64_t* __restrict__ okey,
                                                              unsigned long long* __restrict__ hkey,
                                                              uint32_t* __restrict__ hmx, int bits,
                                                              uint32_t* __restrict__ hslot, uint32_t* __restrict__ errs,
                                                              uint32_t* __restrict__ first, uint32_t* __restrict__ mo,
                                                              Count n, uint32_t* __restrict__ tile_first,
                                                              uint32_t* __restrict__ total) {
    const uint32_t nf = (uint32_t)count_of(n_fail);
    for (uint32_t i = threadIdx.x; i < nf; i += blockDim.x) hash_insert_one(list[i], okey, hkey, hmx, bits, hslot, errs);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nf; i += blockDim.x) {
        const uint32_t o = list[i];
        const uint32_t m = ~__hip_atomic_load(&hmx[hslot[o]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        first[o] = m == o;
        mo[o] = m;
        if (m == o) atomicAdd(&tile_first[o / kDedupTile], 1u);
    }
    __syncthreads();
    const TileTotals tt = {{total, nullptr, nullptr, nullptr}};
    scan_tiles_block((uint32_t)((count_of(n) + kDedupTile - 1) / kDedupTile), tile_first, 1, 1, tt);
}

// Plain form: the active tets' cut lists by active-function count, built on
// the device in active-rank order.  List 0: MI two-material table, 1: up to
// 4 functions, 2: 5..8, 3: more.  Per chunk of kBucketChunk active tets,
// kBucketWords words: list sizes [0..4), the reference's bucket histogram
// [4..7) (IA: 1, 2, >= 3 functions; MI: 2, 3, >= 4), tets with 3..4 and
// 5..8 functions [7], [8], the largest count [9] (k_bucket_count).  One
// workgroup turns the list sizes into each chunk's list offsets and writes
// the totals (k_bucket_scan); each chunk then writes its entries
// (k_bucket_fill).  No atomics.
constexpr uint32_t kBucketChunk = 2048;
constexpr int kBucketWords = 10;

// table2: the material-interface tables take two- and three-material tets
// table2: 1 material interfaces (two and three materials by table), 2
// implicit arrangements (one and two functions by table), 0 no tables
__device__ inline int bucket_of_k(int k, int table2) {
    const bool tab = (table2 == 1 && (k == 2 || k == 3)) || (table2 == 2 && k <= 2);
    return k > 8 ? 3 : (k > 4 ? 2 : (tab ? 0 : 1));
}

__global__ void __launch_bounds__(256) k_bucket_count(Count nc, const uint8_t* __restrict__ act_k, int mi,
                                                      int table2, uint32_t* __restrict__ bc) {
    const uint64_t n = count_of(nc);
    const uint64_t base = (uint64_t)blockIdx.x * kBucketChunk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lo = mi ? 2 : 1;
    uint32_t w[kBucketWords] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < kBucketChunk; i += 256) {
        const uint64_t a = base + i;
        if (a >= n) break;
        const int k = act_k[a];
        const int b = bucket_of_k(k, table2);
        w[0] += b == 0;
        w[1] += b == 1;
        w[2] += b == 2;
        w[3] += b == 3;
        w[4] += k <= lo;
        w[5] += k == lo + 1;
        w[6] += k > lo + 1;
        w[7] += k >= 3 && k <= 4;
        w[8] += k >= 5 && k <= 8;
        w[9] = max(w[9], (uint32_t)k);
    }
    __shared__ uint32_t red[256 / 64][kBucketWords];
#pragma unroll
    for (int f = 0; f < kBucketWords; ++f) {
        uint32_t x = w[f];
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t y = __shfl_xor(x, d, 64);
            x = f == 9 ? max(x, y) : x + y;
        }
        if (lane == 0) red[wave][f] = x;
    }
    __syncthreads();
    if (threadIdx.x < kBucketWords) {
        const int f = threadIdx.x;
        uint32_t x = 0;
        for (int v = 0; v < 256 / 64; ++v) x = f == 9 ? max(x, red[v][f]) : x + red[v][f];
        bc[(uint64_t)blockIdx.x * kBucketWords + f] = x;
    }
}

// one workgroup of 1024 lanes; tot: the 10 totals
__global__ void __launch_bounds__(1024) k_bucket_scan(uint32_t nchunks, uint32_t* __restrict__ bc,
                                                      uint32_t* __restrict__ tot) {
    // statistics first (the scan overwrites the list sizes only)
    __shared__ uint32_t red[16][kBucketWords];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t w[kBucketWords - 4] = {0, 0, 0, 0, 0, 0};
    for (uint32_t c = threadIdx.x; c < nchunks; c += 1024)
        for (int f = 0; f < kBucketWords - 4; ++f) {
            const uint32_t x = bc[(uint64_t)c * kBucketWords + 4 + f];
            w[f] = f == 5 ? max(w[f], x) : w[f] + x;
        }
#pragma unroll
    for (int f = 0; f < kBucketWords - 4; ++f) {
        uint32_t x = w[f];
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t y = __shfl_xor(x, d, 64);
            x = f == 5 ? max(x, y) : x + y;
        }
        if (lane == 0) red[wave][f] = x;
    }
    __syncthreads();
    if (threadIdx.x < kBucketWords - 4) {
        const int f = threadIdx.x;
        uint32_t x = 0;
        for (int v = 0; v < 16; ++v) x = f == 5 ? max(x, red[v][f]) : x + red[v][f];
        tot[4 + f] = x;
    }
    __syncthreads();
    const TileTotals tt = {{tot, tot + 1, tot + 2, tot + 3}};
    scan_tiles_block(nchunks, bc, kBucketWords, 4, tt);
}

__global__ void __launch_bounds__(256) k_bucket_fill(Count nc, const uint8_t* __restrict__ act_k, int table2,
                                                     const uint32_t* __restrict__ bc, uint32_t* __restrict__ l0,
                                                     uint32_t* __restrict__ l1, uint32_t* __restrict__ l2,
                                                     uint32_t* __restrict__ l3) {
    const uint64_t n = count_of(nc);
    const uint64_t base = (uint64_t)blockIdx.x * kBucketChunk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    __shared__ uint32_t wc[256 / 64][4];
    uint32_t c0 = bc[(uint64_t)blockIdx.x * kBucketWords], c1 = bc[(uint64_t)blockIdx.x * kBucketWords + 1],
             c2 = bc[(uint64_t)blockIdx.x * kBucketWords + 2], c3 = bc[(uint64_t)blockIdx.x * kBucketWords + 3];
    for (uint32_t i0 = 0; i0 < kBucketChunk && base + i0 < n; i0 += 256) {
        const uint64_t a = base + i0 + threadIdx.x;
        const int b = a < n ? bucket_of_k(act_k[a], table2) : -1;
        const uint64_t m0 = __ballot(b == 0), m1 = __ballot(b == 1), m2 = __ballot(b == 2), m3 = __ballot(b == 3);
        if (lane == 0) {
            wc[wave][0] = (uint32_t)__popcll(m0);
            wc[wave][1] = (uint32_t)__popcll(m1);
            wc[wave][2] = (uint32_t)__popcll(m2);
            wc[wave][3] = (uint32_t)__popcll(m3);
        }
        __syncthreads();
        uint32_t p[4] = {c0, c1, c2, c3};
        uint32_t t[4] = {0, 0, 0, 0};
        for (int v = 0; v < 256 / 64; ++v)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                p[q] += v < wave ? wc[v][q] : 0u;
                t[q] += wc[v][q];
            }
        if (b == 0) l0[p[0] + (uint32_t)__popcll(m0 & lt)] = (uint32_t)a;
        if (b == 1) l1[p[1] + (uint32_t)__popcll(m1 & lt)] = (uint32_t)a;
        if (b == 2) l2[p[2] + (uint32_t)__popcll(m2 & lt)] = (uint32_t)a;
        if (b == 3) l3[p[3] + (uint32_t)__popcll(m3 & lt)] = (uint32_t)a;
        c0 += t[0];
        c1 += t[1];
        c2 += t[2];
        c3 += t[3];
        __syncthreads();  // before wc is reused
    }
}

// Star path: each first occurrence's vertex id (its rank among first
// occurrences, in occurrence order) from its tile's prefix, written where
// the id lookups read it (vrank at first occurrences: the id of occurrence o
// is vrank[mo[o]]), and first_occ.
// It also clears the hash slots the fallback used (one per distinct key:
// that of its first occurrence), leaving the table zero for the next run.
__global__ void __launch_bounds__(256) k_dedup_rank(Count n, const uint32_t* __restrict__ first,
                                                    const uint32_t* __restrict__ tile_first,
                                                    uint32_t* __restrict__ vrank, uint32_t* __restrict__ first_occ,
                                                    Count n_hash, const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ hslot,
                                                    unsigned long long* __restrict__ hkey, uint32_t* __restrict__ hmx) {
    GRID_STRIDE(i, n_hash) {
        const uint32_t o = list[i];
        if (first[o]) {
            const uint32_t h = hslot[o];
            hkey[h] = 0;
            hmx[h] = 0;
        }
    }
    const uint32_t NO = (uint32_t)count_of(n);
    const uint32_t nt = (NO + kDedupTile - 1) / kDedupTile;
    if (blockIdx.x >= nt) return;
    const uint32_t tile = xcd_tile(blockIdx.x, nt);
    const uint32_t base = tile * kDedupTile;
    __shared__ uint32_t ws[kDedupU][256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t f[kDedupU];
    uint64_t m[kDedupU];
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        const uint32_t o = base + u * 256u + threadIdx.x;
        f[u] = o < NO ? first[o] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        m[u] = __ballot(f[u] != 0);
        if (lane == 0) ws[u][wave] = (uint32_t)__popcll(m[u]);
    }
    __syncthreads();
    uint32_t carry = tile_first[tile];
#pragma unroll
    for (int u = 0; u < kDedupU; ++u) {
        uint32_t before = 0, tot = 0;
        for (int w = 0; w < 256 / 64; ++w) {
            before += w < wave ? ws[u][w] : 0u;
            tot += ws[u][w];
        }
        if (f[u]) {
            const uint32_t o = base + u * 256u + threadIdx.x;
            const uint32_t r = carry + before + (uint32_t)__popcll(m[u] & ((1ull << lane) - 1));
            vrank[o] = r;
            first_occ[r] = o;
        }
        carry += tot;
    }
}

// hash path over occurrences list[0..n) (all occurrences when list is null)
__global__ void k_hash_insert(Count n, const uint32_t* __restrict__ list, const uint64_t* __restrict__ okey,
                              unsigned long long* __restrict__ hkey, uint32_t* __restrict__ hmx, int bits,
                              uint32_t* __restrict__ hslot, uint32_t* __restrict__ errs) {
    GRID_STRIDE(i, n) hash_insert_one(list ? list[i] : (uint32_t)i, okey, hkey, hmx, bits, hslot, errs);
}

__global__ void k_hash_first(Count n, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hslot,
                             const uint32_t* __restrict__ hmx, uint32_t* __restrict__ first, uint32_t* __restrict__ mo) {
    GRID_STRIDE(i, n) {
        const uint32_t o = list ? list[i] : (uint32_t)i;
        const uint32_t m = ~hmx[hslot[o]];
        first[o] = m == o;
        mo[o] = m;
    }
}

// vrank: exclusive scan of `first`.  The hash slots the run used (those of
// list[0..n_hash), or of every occurrence when list is null) are cleared here.
__global__ void k_dedup_assign(Count n, const uint32_t* __restrict__ mo, const uint32_t* __restrict__ first,
                               const uint32_t* __restrict__ vrank, uint32_t* __restrict__ id_of_occ,
                               uint32_t* __restrict__ first_occ, Count n_hash, const uint32_t* __restrict__ list,
                               const uint32_t* __restrict__ hslot, unsigned long long* __restrict__ hkey,
                               uint32_t* __restrict__ hmx, bool set_first_occ) {
    GRID_STRIDE(o, n) {
        const uint32_t id = vrank[mo[o]];
        id_of_occ[o] = id;
        if (set_first_occ && first[o]) first_occ[id] = (uint32_t)o;
    }
    GRID_STRIDE(i, n_hash) {
        const uint32_t o = list ? list[i] : (uint32_t)i;
        if (first[o]) {
            const uint32_t h = hslot[o];
            hkey[h] = 0;
            hmx[h] = 0;
        }
    }
}

// vertex records and coordinates (src/extract_mesh.cpp:95-228, 1446-1538)
// off: CSR offsets by tet (per_rank false) or by active rank (per_rank true)
struct VertFieldArgs {
    GridView g;
    const double* vals;
    const uint32_t* act;
    const uint32_t* off;  // CSR offsets by tet (per_rank false) or by active rank (per_rank true)
    bool per_rank;
    const uint8_t* fit;
    const uint32_t* rec_loc;
    const uint16_t* rec;
    const uint16_t *k1_words;
    const uint32_t* k1_off;
    const uint32_t *first_occ, *oa;
    const uint16_t* oj;
    Count nverts;
    uint32_t* v_tet;
    uint16_t* v_j;
    uint8_t* v_ss;
    uint32_t* v_sv;
    uint8_t* v_fn;
    double* xyz;
};

// as a workgroup role: blk of nblk workgroups
__device__ inline void vert_fields_ia_role(const VertFieldArgs& P, uint32_t blk, uint32_t nblk) {
    const GridView& g = P.g;
    const double* __restrict__ vals = P.vals;
    const uint32_t* __restrict__ act = P.act;
    const uint32_t* __restrict__ off = P.off;
    const bool per_rank = P.per_rank;
    const uint8_t* __restrict__ fit = P.fit;
    const uint32_t* __restrict__ rec_loc = P.rec_loc;
    const uint16_t* __restrict__ rec = P.rec;
    const uint16_t* __restrict__ k1_words = P.k1_words;
    const uint32_t* __restrict__ k1_off = P.k1_off;
    const uint32_t* __restrict__ first_occ = P.first_occ;
    const uint32_t* __restrict__ oa = P.oa;
    const uint16_t* __restrict__ oj = P.oj;
    uint32_t* __restrict__ v_tet = P.v_tet;
    uint16_t* __restrict__ v_j = P.v_j;
    uint8_t* __restrict__ v_ss = P.v_ss;
    uint32_t* __restrict__ v_sv = P.v_sv;
    uint8_t* __restrict__ v_fn = P.v_fn;
    double* __restrict__ xyz = P.xyz;
    const uint64_t V = g.V;
    const uint64_t nv = P.nverts.get();
    for (uint64_t i = (uint64_t)blk * blockDim.x + threadIdx.x; i < nv; i += (uint64_t)nblk * blockDim.x) {
        const uint32_t o = first_occ[i];
        const uint32_t a = oa[o];
        const int j = oj[o];
        const uint32_t t = act[a];
        // a record in the workspace, or (bit 31 set) a table record: the k = 1
        // record of a sign pattern or the k = 2 record of a sign-vector slot
        const uint32_t loc = rec_loc[a];
        const Rec R((loc & 0x80000000u) ? k1_words + k1_off[loc & 0x7FFFFFFFu] : rec + loc);
        uint32_t v[4];
        dev::tet_corners(g, t, v);
        const uint32_t s0 = per_rank ? off[a] : off[t];
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

__global__ void k_vert_fields_ia(VertFieldArgs P) { vert_fields_ia_role(P, blockIdx.x, gridDim.x); }

// ------------------------------------------------------------------ faces
//
// Face occurrences come out of extraction in (tet, local face) order with
// fq_bnd = 0 for faces inside the tet and 1 + l for faces on tet face l.
// Interior faces are never merged.  A boundary face is merged by the
// reference through its (min, second-min, max) vertex-id key
// (src/extract_mesh.h:68-92, src/extract_mesh.cpp:240-253); two faces on
// different tet faces share at most two vertices, so the only possible copy
// lies in the tet across l, and is found there directly.

__global__ void k_face_map(Count nfv, uint32_t* __restrict__ fve, const uint32_t* __restrict__ vid_of_occ) {
    GRID_STRIDE(i, nfv) fve[i] = vid_of_occ[fve[i]];
}

__device__ inline void face_triple(const uint32_t* __restrict__ v, uint32_t n, uint32_t k3[3]) {
    uint32_t mn = v[0], mx = v[0], mp = 0;
    for (uint32_t i = 1; i < n; ++i) {
        if (v[i] < mn) {
            mn = v[i];
            mp = i;
        } else if (v[i] > mx) {
            mx = v[i];
        }
    }
    uint32_t sm = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i)
        if (i != mp && v[i] < sm) sm = v[i];
    k3[0] = mn;
    k3[1] = sm;
    k3[2] = mx;
}

struct FaceArgs {
    const uint32_t *fq_a, *fq_vs, *fve, *act, *rank, *foff, *nbr;
    const uint8_t* fq_bnd;
    const uint16_t* fq_l;
    const uint32_t* nfv_total;  // face-vertex entries (device)
    const uint32_t* nfo_total;  // face occurrences (device)
    uint64_t nfv_cap, nfo_cap;
    Count nq, n_active;
    uint32_t *first, *match, *tfc, *vcnt;
};

__device__ inline uint32_t clamp_count(const uint32_t* d, uint64_t cap) {
    const uint64_t n = *d;
    return (uint32_t)(n < cap ? n : cap);
}

// first-occurrence flag, the first occurrence of the same face (match), the
// face's occurrence count (on its first occurrence) and vertex count
__global__ void k_face_first(FaceArgs A) {
    const uint64_t NA = A.n_active.get(), NQ = A.nq.get();
    const uint32_t nfv = clamp_count(A.nfv_total, A.nfv_cap);
    GRID_STRIDE(q, A.nq) {
        const uint32_t b = A.fq_vs[q], e = (q + 1 < NQ) ? A.fq_vs[q + 1] : nfv;
        uint32_t first = 1, match = (uint32_t)q, dup = 0;
        if (A.fq_bnd[q]) {
            const uint32_t a = A.fq_a[q], t = A.act[a], l = A.fq_bnd[q] - 1u;
            const uint32_t nb = A.nbr[4 * (uint64_t)t + l];
            if (nb != 0xFFFFFFFFu) {
                const uint32_t t2 = nb >> 2, l2 = nb & 3u, r2 = A.rank[t2];
                if (r2 < NA && A.act[r2] == t2) {
                    uint32_t k3[3];
                    face_triple(A.fve + b, e - b, k3);
                    const uint32_t q0 = A.foff[r2];
                    const uint32_t q1 = (r2 + 1 < NA) ? A.foff[r2 + 1] : clamp_count(A.nfo_total, A.nfo_cap);
                    for (uint32_t p = q0; p < q1; ++p) {
                        if (A.fq_bnd[p] != 1 + l2) continue;
                        const uint32_t pb = A.fq_vs[p], pe = (p + 1 < NQ) ? A.fq_vs[p + 1] : nfv;
                        uint32_t p3[3];
                        face_triple(A.fve + pb, pe - pb, p3);
                        if (p3[0] == k3[0] && p3[1] == k3[1] && p3[2] == k3[2]) {
                            if (t2 < t) {
                                first = 0;
                                match = p;
                            } else {
                                dup = 1;
                            }
                            break;
                        }
                    }
                }
            }
        }
        A.first[q] = first;
        A.match[q] = match;
        A.tfc[q] = first ? 1 + dup : 0;
        A.vcnt[q] = first ? e - b : 0;
    }
}

struct FaceOutArgs {
    const uint32_t *fq_a, *fq_vs, *fve, *act, *first, *match, *frank, *tfpos, *vpos;
    const uint16_t* fq_l;
    const uint8_t* fq_fn;
    const uint32_t* nfv_total;
    uint64_t nfv_cap;
    Count nq;
    uint32_t *f_vstart, *f_verts, *f_tfstart, *f_tf_tet;
    uint16_t* f_tf_local;
    uint8_t* f_func;
};

// face records; tet-face occurrences in occurrence order (first, then the copy)
__global__ void k_face_emit(FaceOutArgs A) {
    const uint64_t NQ = A.nq.get();
    const uint32_t nfv = clamp_count(A.nfv_total, A.nfv_cap);
    GRID_STRIDE(q, A.nq) {
        const uint32_t t = A.act[A.fq_a[q]];
        if (A.first[q]) {
            const uint32_t fid = A.frank[q];
            const uint32_t b = A.fq_vs[q], e = (q + 1 < NQ) ? A.fq_vs[q + 1] : nfv;
            uint32_t o = A.vpos[q];
            A.f_vstart[fid] = o;
            for (uint32_t i = b; i < e; ++i) A.f_verts[o++] = A.fve[i];
            A.f_tfstart[fid] = A.tfpos[q];
            A.f_tf_tet[A.tfpos[q]] = t;
            A.f_tf_local[A.tfpos[q]] = A.fq_l[q];
            A.f_func[2 * fid] = A.fq_fn[q];
            A.f_func[2 * fid + 1] = 0xFF;
        } else {
            const uint32_t p = A.tfpos[A.match[q]] + 1;
            A.f_tf_tet[p] = t;
            A.f_tf_local[p] = A.fq_l[q];
        }
    }
}

// ---- faces in tiles (batched form)
//
// The same two passes over tiles of kFaceTile face occurrences (lane-major in
// each of kFaceU rows), with no global scans: the first pass maps corner
// occurrences to vertex ids (into fvm), decides each occurrence (flags: bit 0
// first, bit 1 merged with a later copy; partner: the other copy) and leaves
// per-tile counts of faces, tet-face entries and corners; one workgroup scans
// the tiles; the second pass ranks its tile itself and writes the faces.  A
// first occurrence writes its copy's tet-face entry too.
constexpr int kFaceU = 4;
constexpr uint32_t kFaceTile = 256u * kFaceU;

struct FaceTileArgs {
    const uint32_t *fq_a, *fq_vs, *fve, *act, *rank, *foff, *nbr;
    const uint32_t *mo, *vrank;  // vertex id of occurrence o: vrank[mo[o]]
    const uint8_t* fq_bnd;
    const uint16_t* fq_l;
    const uint8_t* fq_fn;
    const uint32_t* nfv_total;
    const uint32_t* nfo_total;
    uint64_t nfv_cap, nfo_cap;
    Count nq, n_active;
    uint32_t *fvm, *flags, *partner, *tiles;  // tiles: 4 words per tile (faces, tet-face entries, corners)
    const uint32_t* sizes;  // for the emit pass (or null): face occurrences and corner entries
    uint32_t* sizes_out = nullptr;  // the first pass keeps those two sizes here (or not, if null)
    uint32_t *f_vstart, *f_verts, *f_tfstart, *f_tf_tet;
    uint16_t* f_tf_local;
    uint8_t* f_func;
};

// (min, second-min, max) of a face's vertex ids
__device__ inline void face_triple_ids(const FaceTileArgs& A, uint32_t b, uint32_t e, uint32_t k3[3]) {
    uint32_t mn = 0xFFFFFFFFu, sm = 0xFFFFFFFFu, mx = 0;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t x = A.vrank[A.mo[A.fve[i]]];
        if (x < mn) {
            sm = mn;
            mn = x;
        } else if (x < sm) {
            sm = x;
        }
        if (x > mx) mx = x;
    }
    k3[0] = mn;
    k3[1] = sm;
    k3[2] = mx;
}

__device__ inline void face_first_role(const FaceTileArgs& A, uint32_t blk) {
    const uint32_t NQ = (uint32_t)A.nq.get();
    const uint32_t nfv = clamp_count(A.nfv_total, A.nfv_cap);
    if (A.sizes_out && blk == 0 && threadIdx.x == 0) {
        A.sizes_out[0] = NQ;
        A.sizes_out[1] = nfv;
    }
    const uint32_t nt = (NQ + kFaceTile - 1) / kFaceTile;
    if (blk >= nt) return;
    blk = xcd_tile(blk, nt);
    const uint32_t base = blk * kFaceTile;
    const uint32_t NA = (uint32_t)A.n_active.get();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t sf = 0, st = 0, sv = 0;
    for (int u = 0; u < kFaceU; ++u) {
        const uint32_t q = base + u * 256u + threadIdx.x;
        if (q >= NQ) break;
        const uint32_t b = A.fq_vs[q], e = (q + 1 < NQ) ? A.fq_vs[q + 1] : nfv;
        uint32_t mn = 0xFFFFFFFFu, sm = 0xFFFFFFFFu, mx = 0;
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t x = A.vrank[A.mo[A.fve[i]]];
            A.fvm[i] = x;
            if (x < mn) {
                sm = mn;
                mn = x;
            } else if (x < sm) {
                sm = x;
            }
            if (x > mx) mx = x;
        }
        uint32_t first = 1, partner = q, dup = 0;
        if (A.fq_bnd[q]) {
            const uint32_t a = A.fq_a[q], t = A.act[a], l = A.fq_bnd[q] - 1u;
            const uint32_t nb = A.nbr[4 * (uint64_t)t + l];
            if (nb != 0xFFFFFFFFu) {
                const uint32_t t2 = nb >> 2, l2 = nb & 3u, r2 = A.rank[t2];
                if (r2 < NA && A.act[r2] == t2) {
                    const uint32_t q0 = A.foff[r2];
                    const uint32_t q1 = (r2 + 1 < NA) ? A.foff[r2 + 1] : clamp_count(A.nfo_total, A.nfo_cap);
                    for (uint32_t p = q0; p < q1; ++p) {
                        if (A.fq_bnd[p] != 1 + l2) continue;
                        const uint32_t pb = A.fq_vs[p], pe = (p + 1 < NQ) ? A.fq_vs[p + 1] : nfv;
                        uint32_t p3[3];
                        face_triple_ids(A, pb, pe, p3);
                        if (p3[0] == mn && p3[1] == sm && p3[2] == mx) {
                            partner = p;
                            if (t2 < t)
                                first = 0;
                            else
                                dup = 1;
                            break;
                        }
                    }
                }
            }
        }
        A.flags[q] = first | dup << 1;
        A.partner[q] = partner;
        sf += first;
        st += first ? 1 + dup : 0;
        sv += first ? e - b : 0;
    }
    __shared__ uint32_t red[3][256 / 64];
    for (int d = 32; d > 0; d >>= 1) {
        sf += __shfl_xor(sf, d, 64);
        st += __shfl_xor(st, d, 64);
        sv += __shfl_xor(sv, d, 64);
    }
    if (lane == 0) {
        red[0][wave] = sf;
        red[1][wave] = st;
        red[2][wave] = sv;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int f = threadIdx.x;
        A.tiles[4 * (uint64_t)blk + f] = red[f][0] + red[f][1] + red[f][2] + red[f][3];
    }
}

__global__ void __launch_bounds__(256) k_face_first_tiled(FaceTileArgs A) { face_first_role(A, blockIdx.x); }

// the first face pass and the vertex coordinates in one launch (both read
// only the dedup results): face tiles first, then vertices
__global__ void __launch_bounds__(256) k_faces_and_xyz(FaceTileArgs A, VertFieldArgs P, uint32_t nb_face,
                                                       uint32_t nb_xyz) {
    if (blockIdx.x < nb_face)
        face_first_role(A, blockIdx.x);
    else
        vert_fields_ia_role(P, blockIdx.x - nb_face, nb_xyz);
}

// inclusive prefix over the lanes of a wave
__device__ inline uint32_t wave_inclusive(uint32_t x, int lane) {
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// A.tiles holds the raw per-tile counts (no scan): the workgroup adds up
// those of the tiles before its own
__device__ inline void face_emit_role(const FaceTileArgs& A, uint32_t blk) {
    const uint32_t NQ = A.sizes ? A.sizes[0] : (uint32_t)A.nq.get();
    const uint32_t nt = (NQ + kFaceTile - 1) / kFaceTile;
    if (blk >= nt) return;
    blk = xcd_tile(blk, nt);
    const uint32_t base = blk * kFaceTile;
    const uint32_t nfv = A.sizes ? A.sizes[1] : clamp_count(A.nfv_total, A.nfv_cap);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ uint32_t ws[kFaceU][3][256 / 64];
    uint32_t fl[kFaceU], vb[kFaceU], ve[kFaceU], inc[kFaceU][3];
#pragma unroll
    for (int u = 0; u < kFaceU; ++u) {
        const uint32_t q = base + u * 256u + threadIdx.x;
        const bool in = q < NQ;
        fl[u] = in ? A.flags[q] : 0u;
        vb[u] = in ? A.fq_vs[q] : 0u;
        ve[u] = in ? ((q + 1 < NQ) ? A.fq_vs[q + 1] : nfv) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kFaceU; ++u) {
        const uint32_t first = fl[u] & 1u;
        const uint32_t c[3] = {first, first ? 1u + ((fl[u] >> 1) & 1u) : 0u, first ? ve[u] - vb[u] : 0u};
        for (int f = 0; f < 3; ++f) {
            inc[u][f] = wave_inclusive(c[f], lane);
            if (lane == 63) ws[u][f][wave] = inc[u][f];
            inc[u][f] -= c[f];  // exclusive within the wave
        }
    }
    __syncthreads();
    __shared__ uint32_t pred[3][256 / 64];
    {
        uint32_t x[3] = {0, 0, 0};
        for (uint32_t b = threadIdx.x; b < blk; b += 256)
            for (int f = 0; f < 3; ++f) x[f] += A.tiles[4 * (uint64_t)b + f];
        for (int f = 0; f < 3; ++f) {
            for (int d = 32; d > 0; d >>= 1) x[f] += __shfl_xor(x[f], d, 64);
            if (lane == 0) pred[f][wave] = x[f];
        }
    }
    __syncthreads();
    uint32_t carry[3];
    for (int f = 0; f < 3; ++f) carry[f] = pred[f][0] + pred[f][1] + pred[f][2] + pred[f][3];
#pragma unroll
    for (int u = 0; u < kFaceU; ++u) {
        uint32_t pos[3];
        for (int f = 0; f < 3; ++f) {
            uint32_t before = 0, tot = 0;
            for (int w = 0; w < 256 / 64; ++w) {
                before += w < wave ? ws[u][f][w] : 0u;
                tot += ws[u][f][w];
            }
            pos[f] = carry[f] + before + inc[u][f];
            carry[f] += tot;
        }
        const uint32_t q = base + u * 256u + threadIdx.x;
        if (q >= NQ || !(fl[u] & 1u)) continue;
        const uint32_t fid = pos[0], tp = pos[1];
        uint32_t o = pos[2];
        A.f_vstart[fid] = o;
        for (uint32_t i = vb[u]; i < ve[u]; ++i) A.f_verts[o++] = A.fvm[i];
        A.f_tfstart[fid] = tp;
        A.f_tf_tet[tp] = A.act[A.fq_a[q]];
        A.f_tf_local[tp] = A.fq_l[q];
        A.f_func[2 * fid] = A.fq_fn[q];
        A.f_func[2 * fid + 1] = 0xFF;
        if (fl[u] & 2u) {
            // the copy in the tet across, after this one
            const uint32_t p = A.partner[q];
            A.f_tf_tet[tp + 1] = A.act[A.fq_a[p]];
            A.f_tf_local[tp + 1] = A.fq_l[p];
        }
    }
}

}  // namespace
}  // namespace rin
