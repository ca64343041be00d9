// Device view of the background tet mesh and of the function values.
//
// Regular grids follow src/io.cpp:95-152 exactly: vertex (i,j,k) has index
// i*N*N + j*N + k, coordinate t*(max-min)+min with t = i/(N-1), and cube
// (i,j,k) holds five tets whose corner order depends on the parity of i+j+k.
// Their tets are never stored: a tet's corners, edge ids and face ids are
// computed from its index.  Explicit meshes carry corner, edge-id and face-id
// tables built once per mesh.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rin {

struct GridView {
    int regular;
    int res;
    double bmin[3], bmax[3];
    uint64_t V, T;
    const double* pts;       // V x 3
    const uint32_t* tets;    // explicit: T x 4
    const uint32_t* tedge;   // explicit: T x 6 global edge ids (local edges 01,02,03,12,13,23)
    const uint32_t* tface;   // explicit: T x 4 global face ids (face opposite local corner c)
    uint64_t n_edge_ids, n_face_ids;
};

namespace dev {

// corners (0..7: (i,j,k), (i+1,j,k), (i+1,j+1,k), (i,j+1,k), then the same
// at k+1) of the five tets of cube (i,j,k), for even and odd i+j+k
constexpr uint8_t kCubeTetEven[5][4] = {{4, 6, 1, 3}, {6, 3, 4, 7}, {1, 3, 0, 4}, {3, 1, 2, 6}, {4, 1, 6, 5}};
constexpr uint8_t kCubeTetOdd[5][4] = {{7, 0, 2, 5}, {2, 3, 0, 7}, {5, 7, 0, 4}, {7, 2, 6, 5}, {0, 1, 2, 5}};

__host__ __device__ inline void regular_tet(int res, uint64_t t, uint32_t v[4]) {
    // 32-bit index arithmetic: a regular grid has fewer than 2^32 / 27
    // vertices, so its cube and vertex indices fit
    const uint32_t R = (uint32_t)res, N = R + 1;
    const uint32_t t32 = (uint32_t)t, cube = t32 / 5u;
    const int q = (int)(t32 - cube * 5u);
    const uint32_t i = cube / (R * R), rem = cube - i * (R * R), j = rem / R, k = rem - j * R;
    const uint32_t b = i * N * N + j * N + k;
    const uint32_t c[8] = {b, b + N * N, b + N * N + N, b + N, b + 1, b + N * N + 1, b + N * N + N + 1, b + N + 1};
    const bool ev = ((i + j + k) & 1u) == 0;
    for (int a = 0; a < 4; ++a) v[a] = c[ev ? kCubeTetEven[q][a] : kCubeTetOdd[q][a]];
}

__host__ __device__ inline void tet_corners(const GridView& g, uint64_t t, uint32_t v[4]) {
    if (g.regular) {
        regular_tet(g.res, t, v);
    } else {
        for (int a = 0; a < 4; ++a) v[a] = g.tets[4 * t + a];
    }
}

// lattice direction code of b - a in {-1,0,1}^3, 0..26 (a and b corners of
// one cube).  The offsets come from the index difference alone: with N >= 3,
// b - a = di N^2 + dj N + dk splits at N^2 / 2 and N / 2.
__host__ __device__ inline uint32_t lattice_dir(int res, uint32_t a, uint32_t b) {
    const int64_t N = res + 1;
    if (N < 3) {
        const int64_t ai = a / (N * N), aj = (a / N) % N, ak = a % N;
        const int64_t bi = b / (N * N), bj = (b / N) % N, bk = b % N;
        return (uint32_t)((bi - ai + 1) * 9 + (bj - aj + 1) * 3 + (bk - ak + 1));
    }
    const int64_t NN = N * N;
    int64_t d = (int64_t)b - (int64_t)a;
    const int64_t di = 2 * d > NN ? 1 : (2 * d < -NN ? -1 : 0);
    d -= di * NN;
    const int64_t dj = 2 * d > N ? 1 : (2 * d < -N ? -1 : 0);
    d -= dj * N;
    return (uint32_t)((di + 1) * 9 + (dj + 1) * 3 + (d + 1));
}

__host__ __device__ inline int local_edge(int a, int b) {
    if (a > b) {
        int t = a;
        a = b;
        b = t;
    }
    // 01,02,03,12,13,23
    return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}

// global id of the tet edge between local corners la, lb
__host__ __device__ inline uint64_t edge_id(const GridView& g, uint64_t t, const uint32_t v[4], int la, int lb) {
    if (!g.regular) return g.tedge[6 * t + local_edge(la, lb)];
    uint32_t a = v[la], b = v[lb];
    if (a > b) {
        uint32_t x = a;
        a = b;
        b = x;
    }
    return (uint64_t)a * 27 + lattice_dir(g.res, a, b);
}

// global id of the tet face opposite local corner c
__host__ __device__ inline uint64_t face_id(const GridView& g, uint64_t t, const uint32_t v[4], int c) {
    if (!g.regular) return g.tface[4 * t + c];
    uint32_t f[3];
    int n = 0;
    for (int a = 0; a < 4; ++a)
        if (a != c) f[n++] = v[a];
    for (int i = 1; i < 3; ++i)
        for (int j = i; j > 0 && f[j - 1] > f[j]; --j) {
            uint32_t x = f[j];
            f[j] = f[j - 1];
            f[j - 1] = x;
        }
    return ((uint64_t)f[0] * 27 + lattice_dir(g.res, f[0], f[1])) * 27 + lattice_dir(g.res, f[0], f[2]);
}

}  // namespace dev
}  // namespace rin
