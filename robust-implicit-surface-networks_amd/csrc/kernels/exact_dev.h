// Exact orientation predicates on the device.
//
// Every vertex/plane classification of the per-tet cut is the sign of a small
// determinant of input doubles (the reference's external simplicial_arrangement
// decides the same signs exactly; call sites src/implicit_arrangement.cpp:283,
// src/material_interface.cpp:327).  Each determinant is evaluated first in
// FP64 with a permanent-based error bound; when the bound does not separate the
// value from zero it is recomputed exactly with floating-point expansions
// (error-free sums and FMA products; J. R. Shewchuk, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997).
// Rows are first scaled by powers of two (exact) so that neither path
// overflows.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rin {
namespace dev {

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53

__host__ __device__ inline void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}
__host__ __device__ inline void fast_two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a;
    y = b - bv;
}
__host__ __device__ inline void two_prod(double a, double b, double& x, double& y) {
    x = a * b;
    y = fma(a, b, -x);
}

// h = e + b; h may alias e; returns length (>= 1)
__host__ __device__ inline int grow(const double* e, int n, double b, double* h) {
    double Q = b;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        double Qn, hh;
        two_sum(Q, e[i], Qn, hh);
        Q = Qn;
        if (hh != 0.0) h[k++] = hh;
    }
    if (Q != 0.0 || k == 0) h[k++] = Q;
    return k;
}

// h = e * b (h holds 2n entries, must not alias e)
__host__ __device__ inline int scale(const double* e, int n, double b, double* h) {
    double Q, hh;
    two_prod(e[0], b, Q, hh);
    int k = 0;
    if (hh != 0.0) h[k++] = hh;
    for (int i = 1; i < n; ++i) {
        double p1, p0, s;
        two_prod(e[i], b, p1, p0);
        two_sum(Q, p0, s, hh);
        if (hh != 0.0) h[k++] = hh;
        fast_two_sum(p1, s, Q, hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (Q != 0.0 || k == 0) h[k++] = Q;
    return k;
}

// in place: shortest non-overlapping form of e
__host__ __device__ inline int compress(double* e, int n) {
    int bottom = n - 1;
    double Q = e[bottom];
    for (int i = n - 2; i >= 0; --i) {
        double Qn, q;
        fast_two_sum(Q, e[i], Qn, q);
        if (q != 0.0) {
            e[bottom--] = Qn;
            Q = q;
        } else {
            Q = Qn;
        }
    }
    int top = 0;
    for (int i = bottom + 1; i < n; ++i) {
        double Qn, q;
        fast_two_sum(e[i], Q, Qn, q);
        if (q != 0.0) e[top++] = q;
        Q = Qn;
    }
    e[top] = Q;
    return top + 1;
}

__host__ __device__ inline int exp_sign(const double* e, int n) {
    double v = e[n - 1];
    return v > 0 ? 1 : (v < 0 ? -1 : 0);
}

// scale each row by a power of two so its largest entry lies in [1,2)
template <int N>
__host__ __device__ inline void normalise_rows(double (&M)[N][N]) {
    for (int r = 0; r < N; ++r) {
        double mx = 0.0;
        for (int c = 0; c < N; ++c) mx = fmax(mx, fabs(M[r][c]));
        if (mx == 0.0) continue;
        int e;
        frexp(mx, &e);
        for (int c = 0; c < N; ++c) M[r][c] = ldexp(M[r][c], 1 - e);
    }
}

// ---------------------------------------------------------------- filter
// determinant by cofactor recursion over column subsets (rows top to bottom)
template <int N>
__host__ __device__ inline void det_fp(const double (&M)[N][N], double& det, double& perm) {
    // value/permanent of the minor formed by the last s rows and column mask
    double val[1 << N], per[1 << N];
    for (int m = 0; m < (1 << N); ++m) {
        val[m] = 0.0;
        per[m] = 0.0;
    }
    for (int c = 0; c < N; ++c) {
        val[1 << c] = M[N - 1][c];
        per[1 << c] = fabs(M[N - 1][c]);
    }
    for (int s = 2; s <= N; ++s) {
        const int r = N - s;
        for (int m = 0; m < (1 << N); ++m) {
            if (__builtin_popcount(m) != s) continue;
            double acc = 0.0, pacc = 0.0;
            int k = 0;
            for (int c = 0; c < N; ++c) {
                if (!(m & (1 << c))) continue;
                double a = M[r][c];
                int sub = m & ~(1 << c);
                double t = a * val[sub];
                acc += (k & 1) ? -t : t;
                pacc += fabs(a) * per[sub];
                ++k;
            }
            val[m] = acc;
            per[m] = pacc;
        }
    }
    det = val[(1 << N) - 1];
    perm = per[(1 << N) - 1];
}

// exact determinant as an expansion, by cofactor recursion over the rows
// N-S..N-1 and the S columns in `cols`; returns the length or -1 when an
// expansion outgrows its capacity
constexpr int kCap = 32;
template <int S, int N>
__host__ __device__ inline int det_exp(const double (&M)[N][N], const int* cols, double* out) {
    if constexpr (S == 1) {
        out[0] = M[N - 1][cols[0]];
        return 1;
    } else {
        const int r = N - S;
        double sub[kCap];
        double prod[2 * kCap];
        int sc[S - 1];
        int alen = 1;
        out[0] = 0.0;
        for (int j = 0; j < S; ++j) {
            const double a = M[r][cols[j]];
            if (a == 0.0) continue;
            int n = 0;
            for (int i = 0; i < S; ++i)
                if (i != j) sc[n++] = cols[i];
            const int sl = det_exp<S - 1, N>(M, sc, sub);
            if (sl < 0) return -1;
            const int pl = scale(sub, sl, (j & 1) ? -a : a, prod);
            for (int i = 0; i < pl; ++i) {
                if (alen >= kCap) {
                    alen = compress(out, alen);
                    if (alen >= kCap) return -1;
                }
                alen = grow(out, alen, prod[i], out);
            }
            alen = compress(out, alen);
        }
        return alen;
    }
}

// The exact path is out of line and takes its matrix by value: the caller
// builds the copy only when the filter has failed, so on the common path the
// matrix stays in registers (a by-reference argument would give it an
// address, and every classification would store it to scratch memory).
template <int N>
struct ExactMat {
    double m[N][N];
};

// sign of the determinant, or 2 when an expansion outgrows its capacity
template <int N>
__host__ __device__ __noinline__ int det_exact(ExactMat<N> M) {
    int cols[N];
    for (int c = 0; c < N; ++c) cols[c] = c;
    double out[kCap];
    const int n = det_exp<N, N>(M.m, cols, out);
    if (n < 0) return 2;
    return exp_sign(out, n);
}

template <int N>
__host__ __device__ inline int det_sign(double (&M)[N][N], bool& err) {
    normalise_rows<N>(M);
    double d, p;
    det_fp<N>(M, d, p);
    if (p > 1e-200) {
        const double bound = 2.0 * N * N * kEps * p;
        if (d > bound) return 1;
        if (d < -bound) return -1;
    }
    // with every non-zero entry of the normalised rows at least 2^-100, every
    // partial product of the expansion has its lowest bit above 2^-760 and no
    // error-free transformation can underflow; smaller entries are reported
    ExactMat<N> E;
    for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
            if (M[r][c] != 0.0 && fabs(M[r][c]) < 0x1p-100) err = true;
            E.m[r][c] = M[r][c];
        }
    const int sgn = det_exact<N>(E);
    if (sgn == 2) {
        err = true;
        return 0;
    }
    return sgn;
}

// ---------------------------------------------------------------- IA
// Sign of implicit plane `q` (corner values) at the vertex where planes
// t0<t1<t2 meet.  Ids 0..3 are the tet faces (barycentric coordinate = 0).
// Reduced by the boundary planes: the free corners are those whose
// coordinate is not pinned, the vertex is the null vector of the implicit
// rows restricted to them, and sign = sign(det[rows; q]) * sign(det[rows; 1]).
// Sign of plane q at the point of tet edge (c1, c2) where plane f vanishes,
// from their values at the edge's corners (c1 < c2).
__host__ __device__ inline int orient_ia_edge(double f1, double f2, double q1, double q2, bool& err) {
    double A[2][2] = {{f1, f2}, {q1, q2}};
    double B[2][2] = {{f1, f2}, {1.0, 1.0}};
    int a = det_sign<2>(A, err), b = det_sign<2>(B, err);
    if (b == 0) err = true;
    return a * b;
}

// the row of M scaled by a power of two so its largest entry lies in [1,2)
// (normalise_rows for one row)
template <int N>
__host__ __device__ inline void normalise_row(double (&r)[N]) {
    double mx = 0.0;
    for (int c = 0; c < N; ++c) mx = fmax(mx, fabs(r[c]));
    if (mx == 0.0) return;
    int e;
    frexp(mx, &e);
    for (int c = 0; c < N; ++c) r[c] = ldexp(r[c], 1 - e);
}

// sign of det[top; R] for an N x N matrix whose rows 1..N-1 are R, given
// the values and permanents of R's (N-1)-minors (mv[c], mp[c]: column c
// left out): the filter of det_sign on that matrix, with the exact
// expansion when the bound does not decide.  The rows must be normalised.
template <int N>
__host__ __device__ inline int det_sign_top(const double (&top)[N], const double (&R)[N - 1][N], const double (&mv)[N],
                                            const double (&mp)[N], bool& err) {
    double d = 0.0, p = 0.0;
    for (int c = 0; c < N; ++c) {
        const double t = top[c] * mv[c];
        d += (c & 1) ? -t : t;
        p += fabs(top[c]) * mp[c];
    }
    if (p > 1e-200) {
        const double bound = 2.0 * N * N * kEps * p;
        if (d > bound) return 1;
        if (d < -bound) return -1;
    }
    ExactMat<N> E;
    for (int c = 0; c < N; ++c) {
        if (top[c] != 0.0 && fabs(top[c]) < 0x1p-100) err = true;
        E.m[0][c] = top[c];
    }
    for (int r = 0; r < N - 1; ++r)
        for (int c = 0; c < N; ++c) {
            if (R[r][c] != 0.0 && fabs(R[r][c]) < 0x1p-100) err = true;
            E.m[r + 1][c] = R[r][c];
        }
    const int sgn = det_exact<N>(E);
    if (sgn == 2) {
        err = true;
        return 0;
    }
    return sgn;
}

// the (N-1)-minors of the N-1 rows R (column c left out), by det_fp's
// recursion on those rows: values and permanents
template <int N>
__host__ __device__ inline void row_minors(const double (&R)[N - 1][N], double (&mv)[N], double (&mp)[N]) {
    constexpr int M = N - 1;
    double val[1 << N], per[1 << N];
    for (int m = 0; m < (1 << N); ++m) {
        val[m] = 0.0;
        per[m] = 0.0;
    }
    for (int c = 0; c < N; ++c) {
        val[1 << c] = R[M - 1][c];
        per[1 << c] = fabs(R[M - 1][c]);
    }
    for (int s = 2; s <= M; ++s) {
        const int r = M - s;
        for (int m = 0; m < (1 << N); ++m) {
            if (__builtin_popcount(m) != s) continue;
            double acc = 0.0, pacc = 0.0;
            int k = 0;
            for (int c = 0; c < N; ++c) {
                if (!(m & (1 << c))) continue;
                const double a = R[r][c];
                const int sub = m & ~(1 << c);
                const double t = a * val[sub];
                acc += (k & 1) ? -t : t;
                pacc += fabs(a) * per[sub];
                ++k;
            }
            val[m] = acc;
            per[m] = pacc;
        }
    }
    for (int c = 0; c < N; ++c) {
        mv[c] = val[((1 << N) - 1) & ~(1 << c)];
        mp[c] = per[((1 << N) - 1) & ~(1 << c)];
    }
}

// sign(det[q; R]) * sign(det[n; R]) for the query row q and the
// normalisation row n: both determinants share R's minors (the row order
// differs from [R; q] and [R; n] alike in both, so the product is the same
// sign).  n must be normalised already (its largest entry 1).
template <int N>
__host__ __device__ inline int orient_rows(double (&R)[N - 1][N], double (&q)[N], const double (&n)[N], bool& err) {
    for (int r = 0; r < N - 1; ++r) normalise_row<N>(R[r]);
    normalise_row<N>(q);
    double mv[N], mp[N];
    row_minors<N>(R, mv, mp);
    const int a = det_sign_top<N>(q, R, mv, mp, err);
    const int b = det_sign_top<N>(n, R, mv, mp, err);
    if (b == 0) err = true;
    return a * b;
}

// orient_ia with the plane rows' minors shared between the two determinants
__host__ __device__ inline int orient_ia(const double (*P)[4], int t0, int t1, int t2, const double* q, bool& err) {
    unsigned pm = 0;
    if (t0 < 4) pm |= 1u << t0;
    if (t1 < 4) pm |= 1u << t1;
    if (t2 < 4) pm |= 1u << t2;
    const int ni = (t0 >= 4) + (t1 >= 4) + (t2 >= 4);
    const int i0 = t0 >= 4 ? t0 - 4 : (t1 >= 4 ? t1 - 4 : t2 - 4);
    const int i1 = (t0 >= 4 && t1 >= 4) ? t1 - 4 : t2 - 4;
    const int i2 = t2 - 4;
    unsigned fm = ~pm & 15u;
    const int c0 = __builtin_ctz(fm | 16u);
    fm &= fm - 1;
    const int c1 = __builtin_ctz(fm | 16u);
    fm &= fm - 1;
    const int c2 = __builtin_ctz(fm | 16u);
    if (ni == 0) {
        const double v = q[c0];
        return v > 0 ? 1 : (v < 0 ? -1 : 0);
    }
    if (ni == 1) return orient_ia_edge(P[i0][c0], P[i0][c1], q[c0], q[c1], err);
    if (ni == 2) {
        double R[2][3] = {{P[i0][c0], P[i0][c1], P[i0][c2]}, {P[i1][c0], P[i1][c1], P[i1][c2]}};
        double qq[3] = {q[c0], q[c1], q[c2]};
        const double one[3] = {1.0, 1.0, 1.0};
        return orient_rows<3>(R, qq, one, err);
    }
    double R[3][4] = {{P[i0][0], P[i0][1], P[i0][2], P[i0][3]},
                      {P[i1][0], P[i1][1], P[i1][2], P[i1][3]},
                      {P[i2][0], P[i2][1], P[i2][2], P[i2][3]}};
    double qq[4] = {q[0], q[1], q[2], q[3]};
    const double one[4] = {1.0, 1.0, 1.0, 1.0};
    return orient_rows<4>(R, qq, one, err);
}

// ---------------------------------------------------------------- MI
// Sign of f_m - (common value of the vertex's materials) at the vertex given
// by four labels (0..3 tet faces, 4.. materials).  Unknowns: the free
// barycentric coordinates and the common value t; rows (f_a, -1) per
// material, then either the query (f_m, -1) or the normalisation (1..1, 0).
template <int S>
__host__ __device__ inline int orient_mi_n(const double (*P)[4], const int* mats, const int* fr, const double* q,
                                           bool& err) {
    constexpr int N = S + 1;
    double R[S][N], qq[N], nn[N];
    for (int r = 0; r < S; ++r) {
        for (int c = 0; c < S; ++c) R[r][c] = P[mats[r]][fr[c]];
        R[r][S] = -1.0;
    }
    for (int c = 0; c < S; ++c) {
        qq[c] = q[fr[c]];
        nn[c] = 1.0;
    }
    qq[S] = -1.0;
    nn[S] = 0.0;
    return orient_rows<N>(R, qq, nn, err);
}

__host__ __device__ inline int orient_mi(const double (*P)[4], const int* lab, const double* q, bool& err) {
    bool pinned[4] = {false, false, false, false};
    int mats[4], nm = 0;
    for (int i = 0; i < 4; ++i) {
        if (lab[i] < 4)
            pinned[lab[i]] = true;
        else
            mats[nm++] = lab[i] - 4;
    }
    int fr[4], nf = 0;
    for (int c = 0; c < 4; ++c)
        if (!pinned[c]) fr[nf++] = c;
    switch (nm) {
        case 1: {
            double a = q[fr[0]], b = P[mats[0]][fr[0]];
            return a > b ? 1 : (a < b ? -1 : 0);
        }
        case 2:
            return orient_mi_n<2>(P, mats, fr, q, err);
        case 3:
            return orient_mi_n<3>(P, mats, fr, q, err);
        default:
            return orient_mi_n<4>(P, mats, fr, q, err);
    }
}

}  // namespace dev
}  // namespace rin
