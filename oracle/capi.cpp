// ORACLE — test infrastructure only.  C entry points for the test suite
// (loaded with ctypes from tests/ and from bench.py's CPU-baseline leg).
#include <omp.h>

#include <cstring>
#include <string>

#include "exact.h"
#include "oracle.h"

using namespace oracle;

namespace {
thread_local std::string g_err;
struct Handle {
    Result r;
    std::vector<uint64_t> active;
    std::vector<uint64_t> rec_start;
    std::vector<uint16_t> words;
};
std::vector<Func> make_funcs(const int* types, const double* params, int F) {
    std::vector<Func> fs(F);
    for (int j = 0; j < F; ++j) {
        fs[j].type = types[j];
        std::memcpy(fs[j].p, params + 12 * j, sizeof(double) * 12);
    }
    return fs;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
// OpenMP threads of the per-tet cut and the evaluation (0: leave as is);
// returns the number in effect
int orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
uint64_t orc_exact_fallbacks() { return g_exact_fallbacks; }

int orc_grid(int res, const double* bmin, const double* bmax, double* pts, uint64_t* tets) {
    std::vector<std::array<double, 3>> P;
    std::vector<std::array<uint64_t, 4>> T;
    grid(res, bmin, bmax, P, T);
    std::memcpy(pts, P.data(), P.size() * sizeof(P[0]));
    std::memcpy(tets, T.data(), T.size() * sizeof(T[0]));
    return 0;
}

int orc_eval(const int* types, const double* params, int F, const double* pts, uint64_t V, double* vals) {
    try {
        std::vector<std::array<double, 3>> P(V);
        std::memcpy(P.data(), pts, V * sizeof(P[0]));
        std::vector<double> out;
        eval(make_funcs(types, params, F), P, out);
        std::memcpy(vals, out.data(), out.size() * sizeof(double));
        return 0;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void* orc_run(int mi, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, const double* vals, uint64_t F,
              int flags) {
    try {
        std::vector<std::array<double, 3>> P(V);
        std::memcpy(P.data(), pts, V * sizeof(P[0]));
        std::vector<std::array<uint64_t, 4>> Tt(T);
        std::memcpy(Tt.data(), tets, T * sizeof(Tt[0]));
        std::vector<double> vv(vals, vals + V * F);
        Handle* h = new Handle;
        if (mi)
            run_mi(P, Tt, vv, F, flags | F_KEEP_CUTS, h->r);
        else
            run_ia(P, Tt, vv, F, flags | F_KEEP_CUTS, h->r);
        for (uint64_t t = 0; t < T; ++t)
            if (h->r.cut_index[t] >= 0) h->active.push_back(t);
        h->rec_start.push_back(0);
        for (size_t a = 0; a < h->active.size(); ++a) {
            if (mi)
                serialize_mi(h->r.mi_cuts[a], h->words);
            else
                serialize_ia(h->r.ia_cuts[a], h->words);
            h->rec_start.push_back(h->words.size());
        }
        if (!(flags & F_KEEP_CUTS)) {
            h->r.ia_cuts.clear();
            h->r.mi_cuts.clear();
        }
        return h;
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_free(void* p) { delete (Handle*)p; }

// [0] degenerate entries [1] intersecting tets [2..4] bucket counts
// [5] CSR length [6] vertices [7] faces [8] face-vertex entries
// [9] tet-face entries [10] active tets [11] record words
void orc_sizes(void* p, uint64_t* o) {
    Handle* h = (Handle*)p;
    const Result& r = h->r;
    o[0] = r.num_degenerate;
    o[1] = r.num_intersecting;
    o[2] = r.nb[0];
    o[3] = r.nb[1];
    o[4] = r.nb[2];
    o[5] = r.func_in_tet.size();
    o[6] = r.verts.size();
    o[7] = r.faces.size();
    uint64_t fv = 0, tf = 0;
    for (const auto& f : r.faces) {
        fv += f.vs.size();
        tf += f.tet_faces.size();
    }
    o[8] = fv;
    o[9] = tf;
    o[10] = h->active.size();
    o[11] = h->words.size();
}

void orc_copy_csr(void* p, uint64_t* func_in_tet, uint64_t* start_index) {
    const Result& r = ((Handle*)p)->r;
    std::memcpy(func_in_tet, r.func_in_tet.data(), r.func_in_tet.size() * 8);
    std::memcpy(start_index, r.start_index.data(), r.start_index.size() * 8);
}

// per vertex 11 words: tet, tet_vert, simplex_size, sv[4], fn[4]
void orc_copy_verts(void* p, uint64_t* rec, double* xyz) {
    const Result& r = ((Handle*)p)->r;
    for (size_t i = 0; i < r.verts.size(); ++i) {
        const IsoVert& v = r.verts[i];
        uint64_t* o = rec + 11 * i;
        o[0] = v.tet;
        o[1] = v.tet_vert;
        o[2] = (uint64_t)v.simplex_size;
        for (int q = 0; q < 4; ++q) o[3 + q] = v.simplex_size > q ? v.sv[q] : UINT64_MAX;
        if (v.simplex_size == 4)
            for (int q = 0; q < 4; ++q) o[3 + q] = v.sv[q];
        for (int q = 0; q < 4; ++q) o[7 + q] = v.fn[q];
    }
    std::memcpy(xyz, r.pts.data(), r.pts.size() * 24);
}

void orc_copy_faces(void* p, uint64_t* vstart, uint64_t* vs, uint64_t* tfstart, uint64_t* tf, uint64_t* fn) {
    const Result& r = ((Handle*)p)->r;
    uint64_t a = 0, b = 0;
    vstart[0] = 0;
    tfstart[0] = 0;
    for (size_t i = 0; i < r.faces.size(); ++i) {
        const IsoFace& f = r.faces[i];
        for (uint64_t v : f.vs) vs[a++] = v;
        for (auto& x : f.tet_faces) {
            tf[2 * b] = x.first;
            tf[2 * b + 1] = x.second;
            ++b;
        }
        vstart[i + 1] = a;
        tfstart[i + 1] = b;
        fn[2 * i] = f.f0;
        fn[2 * i + 1] = f.f1;
    }
}

void orc_copy_records(void* p, uint64_t* active, uint64_t* rec_start, uint16_t* words) {
    Handle* h = (Handle*)p;
    std::memcpy(active, h->active.data(), h->active.size() * 8);
    std::memcpy(rec_start, h->rec_start.data(), h->rec_start.size() * 8);
    std::memcpy(words, h->words.data(), h->words.size() * 2);
}

int orc_timings(void* p, double* out, int cap) {
    const Result& r = ((Handle*)p)->r;
    int n = 0;
    for (auto& t : r.timings)
        if (n < cap) out[n++] = t.second;
    return n;
}
const char* orc_timing_label(void* p, int i) { return ((Handle*)p)->r.timings[i].first; }

// one cut, serialised; returns the number of words or -1
int orc_cut(int mi, const double* rows, int k, uint16_t* out, int cap) {
    try {
        std::vector<std::array<double, 4>> P(k);
        for (int j = 0; j < k; ++j)
            for (int c = 0; c < 4; ++c) P[j][c] = rows[4 * j + c];
        std::vector<uint16_t> w;
        if (mi)
            serialize_mi(mi_cut(P), w);
        else
            serialize_ia(ia_cut(P), w);
        if ((int)w.size() > cap) return -(int)w.size();
        std::memcpy(out, w.data(), w.size() * 2);
        return (int)w.size();
    } catch (std::exception& e) {
        g_err = e.what();
        return -1000000;
    }
}

}  // extern "C"
