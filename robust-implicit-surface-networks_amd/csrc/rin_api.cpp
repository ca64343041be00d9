// C ABI over the device pipeline (declarations and conventions in include/rin.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "devmem.h"
#include "kernels/pipeline.h"
#include "rin.h"

struct rin_ctx {
    int device = 0;
    int n_cu = 256;
    void* scratch = nullptr;  // the allocation-free path's array on this device
    hipStream_t stream = nullptr;
    rin::DevMem mem;
    std::mutex mu;
    std::string err;
};
struct rin_grid {
    rin::Grid g;
};
struct rin_values {
    rin::Values v;
};
struct rin_result {
    rin::RunOut r;
};

namespace {

int fail(rin_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <class F>
int guarded(rin_ctx* c, F&& f) {
    if (!c) return RIN_ERR_ARG;
    std::lock_guard<std::mutex> lock(c->mu);
    try {
        RIN_HIP(hipSetDevice(c->device));
        f();
        return RIN_OK;
    } catch (rin::HipError& e) {
        return fail(c, e.code == hipErrorOutOfMemory ? RIN_ERR_MEMORY : RIN_ERR_DEVICE, e.what());
    } catch (std::invalid_argument& e) {
        return fail(c, RIN_ERR_ARG, e.what());
    } catch (std::runtime_error& e) {
        const std::string w = e.what();
        if (w.find("active functions") != std::string::npos || w.find("capacity") != std::string::npos ||
            w.find("exceed") != std::string::npos)
            return fail(c, RIN_ERR_CAPACITY, w);
        if (w.find("not available") != std::string::npos) return fail(c, RIN_ERR_UNSUPPORTED, w);
        return fail(c, RIN_ERR_INTERNAL, w);
    } catch (std::exception& e) {
        return fail(c, RIN_ERR_INTERNAL, e.what());
    }
}

template <class T, class U>
void widen(const rin::DBuf<T>& d, size_t n, std::vector<U>& h, hipStream_t s) {
    std::vector<T> tmp(n);
    if (n) RIN_HIP(hipMemcpyAsync(tmp.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    h.assign(tmp.begin(), tmp.end());
}

template <class T>
std::vector<T> fetch(const T* d, size_t n, hipStream_t s) {
    std::vector<T> h(n);
    if (n) RIN_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace

extern "C" {

int rin_create(rin_ctx** out, int32_t device) {
    if (!out) return RIN_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RIN_ERR_DEVICE;
    if (device < 0 || device >= n) return RIN_ERR_ARG;
    auto* c = new rin_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RIN_ERR_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->n_cu = prop.multiProcessorCount;
    try {
        c->scratch = rin::noalloc_scratch_address();
    } catch (std::exception&) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return RIN_ERR_DEVICE;
    }
    *out = c;
    return RIN_OK;
}

int rin_device_memory(rin_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        size_t f = 0, t = 0;
        RIN_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
    });
}

int rin_destroy(rin_ctx* c) {
    if (!c) return RIN_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->mem.trim();
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RIN_OK;
}

const char* rin_last_error(const rin_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rin_grid_regular(rin_ctx* c, int32_t res, const double bmin[3], const double bmax[3], rin_grid** out) {
    if (!out || !bmin || !bmax) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto g = std::make_unique<rin_grid>();
        rin::grid_regular(c->mem, res, bmin, bmax, g->g, c->stream);
        RIN_HIP(hipStreamSynchronize(c->stream));
        *out = g.release();
    });
}

int rin_grid_explicit(rin_ctx* c, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, rin_grid** out) {
    if (!out || !pts || !tets) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto g = std::make_unique<rin_grid>();
        rin::grid_explicit(c->mem, pts, V, tets, T, g->g, c->stream);
        *out = g.release();
    });
}

int rin_grid_points(rin_ctx* c, const rin_grid* g, double* pts) {
    if (!g || !pts) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        RIN_HIP(hipMemcpyAsync(pts, g->g.pts.p, 24 * g->g.view.V, hipMemcpyDeviceToHost, c->stream));
        RIN_HIP(hipStreamSynchronize(c->stream));
    });
}

int rin_grid_release(rin_ctx* c, rin_grid* g) {
    return guarded(c, [&] {
        RIN_HIP(hipStreamSynchronize(c->stream));
        delete g;
    });
}

int rin_values_eval(rin_ctx* c, const rin_grid* g, const rin_func_desc* f, uint32_t F, rin_values** out) {
    if (!g || !f || !out) return fail(c, RIN_ERR_ARG, "null argument");
    static_assert(sizeof(rin_func_desc) == sizeof(rin::FuncDesc), "descriptor layout");
    return guarded(c, [&] {
        auto v = std::make_unique<rin_values>();
        rin::values_eval(c->mem, g->g, reinterpret_cast<const rin::FuncDesc*>(f), F, v->v, c->stream);
        *out = v.release();
    });
}

int rin_values_upload(rin_ctx* c, const rin_grid* g, const double* rm, uint32_t F, rin_values** out) {
    if (!g || !rm || !out) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto v = std::make_unique<rin_values>();
        rin::values_upload(c->mem, g->g, rm, F, v->v, c->stream);
        *out = v.release();
    });
}

int rin_eval_points(rin_ctx* c, const double* pts, uint64_t n, const rin_func_desc* f, uint32_t F, double* rm) {
    if ((!pts && n) || !f || (!rm && n)) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        rin::eval_points(c->mem, pts, n, reinterpret_cast<const rin::FuncDesc*>(f), F, rm, c->stream);
    });
}

int rin_eval_points_noalloc(rin_ctx* c, const double* pts, uint64_t n, const rin_func_desc* f, uint32_t F,
                            double* rm) {
    if ((!pts && n) || !f || (!rm && n)) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        rin::eval_points_noalloc(c->scratch, pts, n, reinterpret_cast<const rin::FuncDesc*>(f), F, rm, c->stream);
        RIN_HIP(hipStreamSynchronize(c->stream));
    });
}

int rin_values_download(rin_ctx* c, const rin_values* v, double* rm) {
    if (!v || !rm) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const uint64_t V = v->v.V, F = v->v.F;
        auto h = fetch(v->v.vals.p, V * F, c->stream);
        for (uint64_t f = 0; f < F; ++f)
            for (uint64_t i = 0; i < V; ++i) rm[i * F + f] = h[f * V + i];
    });
}

int rin_values_release(rin_ctx* c, rin_values* v) {
    return guarded(c, [&] {
        RIN_HIP(hipStreamSynchronize(c->stream));
        delete v;
    });
}

int rin_run(rin_ctx* c, const rin_grid* g, const rin_values* v, uint32_t flags, rin_result** out) {
    if (!g || !v || !out) return fail(c, RIN_ERR_ARG, "null argument");
    *out = nullptr;
    if (v->v.grid_id != g->g.id) return fail(c, RIN_ERR_ARG, "the values were made on another grid");
    return guarded(c, [&] {
        auto r = std::make_unique<rin_result>();
        if (flags & RIN_FLAG_BATCHED) {
            try {
                if (flags & RIN_FLAG_MI)
                    rin::run_batched_mi(c->mem, g->g, v->v, flags, r->r, c->stream);
                else
                    rin::run_batched(c->mem, g->g, v->v, flags, r->r, c->stream, c->n_cu);
            } catch (rin::NeedsPlainForm& e) {
                // the plain form handles every tet the buckets cannot
                RIN_HIP(hipStreamSynchronize(c->stream));
                r = std::make_unique<rin_result>();
                rin::run_pipeline(c->mem, g->g, v->v, flags & ~(uint32_t)RIN_FLAG_BATCHED, r->r, c->stream);
                r->r.times.ms.insert(r->r.times.ms.begin(), {std::string("batched form fell back: ") + e.what(), 0.f});
            }
        } else {
            rin::run_pipeline(c->mem, g->g, v->v, flags, r->r, c->stream);
        }
        RIN_HIP(hipStreamSynchronize(c->stream));
        *out = r.release();
    });
}

int rin_result_sizes(rin_ctx* c, const rin_result* res, rin_sizes* s) {
    if (!res || !s) return fail(c, RIN_ERR_ARG, "null argument");
    const auto& r = res->r;
    s->n_tets = r.T;
    s->n_funcs = r.F;
    s->num_degenerate = r.num_degenerate;
    s->num_intersecting = r.num_intersecting;
    for (int i = 0; i < 3; ++i) s->bucket[i] = r.nb[i];
    s->n_func_in_tet = r.n_fit;
    s->n_verts = r.n_verts;
    s->n_faces = r.n_faces;
    s->n_face_verts = r.n_face_verts;
    s->n_tet_faces = r.n_tet_faces;
    s->n_vert_occ = r.n_vert_occ;
    s->n_face_occ = r.n_face_occ;
    s->n_face_occ_verts = r.n_face_occ_verts;
    s->n_cut[0] = r.n_cut34;
    s->n_cut[1] = r.n_cut58;
    s->n_rec_words = r.n_rec_words;
    for (int i = 0; i < 2; ++i) {
        s->cut_k_sum[i] = r.cut_k_sum[i];
        s->cut_words[i] = r.cut_words[i];
    }
    return RIN_OK;
}

int rin_result_csr(rin_ctx* c, const rin_result* res, uint64_t* fit, uint64_t* start) {
    if (!res || !fit || !start) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const auto& r = res->r;
        auto a = fetch(r.view.fit, r.n_fit, c->stream);
        auto b = fetch(r.view.csr_off, r.T + 1, c->stream);
        for (uint64_t i = 0; i < r.n_fit; ++i) fit[i] = a[i];
        for (uint64_t i = 0; i <= r.T; ++i) start[i] = b[i];
    });
}

int rin_result_verts(rin_ctx* c, const rin_result* res, uint64_t* rec11, double* xyz) {
    if (!res || !rec11 || !xyz) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const auto& r = res->r;
        const uint64_t n = r.n_verts;
        auto tet = fetch(r.view.v_tet, n, c->stream);
        auto j = fetch(r.view.v_j, n, c->stream);
        auto ss = fetch(r.view.v_ss, n, c->stream);
        auto sv = fetch(r.view.v_sv, 4 * n, c->stream);
        auto fn = fetch(r.view.v_fn, 4 * n, c->stream);
        auto x = fetch(r.view.xyz, 3 * n, c->stream);
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t* o = rec11 + 11 * i;
            o[0] = tet[i];
            o[1] = j[i];
            o[2] = ss[i];
            for (int q = 0; q < 4; ++q) o[3 + q] = sv[4 * i + q] == 0xFFFFFFFFu ? UINT64_MAX : sv[4 * i + q];
            for (int q = 0; q < 4; ++q) o[7 + q] = fn[4 * i + q] == 0xFF ? UINT64_MAX : fn[4 * i + q];
        }
        std::memcpy(xyz, x.data(), 24 * n);
    });
}

int rin_result_faces(rin_ctx* c, const rin_result* res, uint64_t* vstart, uint64_t* verts, uint64_t* tfstart,
                     uint64_t* tf, uint64_t* func) {
    if (!res || !vstart || !verts || !tfstart || !tf || !func) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const auto& r = res->r;
        const uint64_t nf = r.n_faces;
        auto a = fetch(r.view.f_vstart, nf + 1, c->stream);
        auto b = fetch(r.view.f_verts, r.n_face_verts, c->stream);
        auto d = fetch(r.view.f_tfstart, nf + 1, c->stream);
        auto e = fetch(r.view.f_tf_tet, r.n_tet_faces, c->stream);
        auto f = fetch(r.view.f_tf_local, r.n_tet_faces, c->stream);
        auto g = fetch(r.view.f_func, 2 * nf, c->stream);
        for (uint64_t i = 0; i <= nf; ++i) {
            vstart[i] = a[i];
            tfstart[i] = d[i];
        }
        for (uint64_t i = 0; i < r.n_face_verts; ++i) verts[i] = b[i];
        for (uint64_t i = 0; i < r.n_tet_faces; ++i) {
            tf[2 * i] = e[i];
            tf[2 * i + 1] = f[i];
        }
        for (uint64_t i = 0; i < 2 * nf; ++i) func[i] = g[i] == 0xFF ? UINT64_MAX : g[i];
    });
}

int rin_result_edge_count(rin_ctx* c, const rin_result* res, uint64_t* n_edges) {
    if (!res || !n_edges) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;  // the edges are cached in the result
        rin::compute_edges(c->mem, r, c->stream);
        *n_edges = r.n_edges;
    });
}

int rin_result_edges(rin_ctx* c, const rin_result* res, uint64_t* v12, uint64_t* start, uint64_t* occ,
                     uint64_t* edge_of) {
    if (!res || !v12 || !start || !occ || !edge_of) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;
        rin::compute_edges(c->mem, r, c->stream);
        const uint64_t NE = r.n_edges, NO = r.n_face_verts, NF = r.n_faces;
        auto a = fetch(r.e_v12.p, 2 * NE, c->stream);
        auto b = fetch(r.e_start.p, NE + 1, c->stream);
        auto o = fetch(r.e_occ.p, NO, c->stream);
        auto e = fetch(r.e_of_occ.p, NO, c->stream);
        auto f = fetch(r.e_face_of.p, NO, c->stream);
        auto vs = fetch(r.view.f_vstart, NF + 1, c->stream);
        for (uint64_t i = 0; i < 2 * NE; ++i) v12[i] = a[i];
        for (uint64_t i = 0; i <= NE; ++i) start[i] = b[i];
        for (uint64_t i = 0; i < NO; ++i) {
            const uint32_t x = o[i], face = f[x];
            occ[2 * i] = face;
            occ[2 * i + 1] = x - vs[face];
            edge_of[i] = e[i];
        }
    });
}

int rin_result_patch_count(rin_ctx* c, const rin_result* res, uint64_t* n_patches) {
    if (!res || !n_patches) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;  // the patches are cached in the result
        rin::compute_patches(c->mem, r, c->stream);
        *n_patches = r.n_patches;
    });
}

int rin_result_patches(rin_ctx* c, const rin_result* res, uint64_t* start, uint64_t* faces) {
    if (!res || !start || (!faces && res->r.n_faces)) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;
        rin::compute_patches(c->mem, r, c->stream);
        auto a = fetch(r.p_start.p, r.n_patches + 1, c->stream);
        auto b = fetch(r.p_faces.p, r.n_faces, c->stream);
        for (uint64_t i = 0; i <= r.n_patches; ++i) start[i] = a[i];
        for (uint64_t i = 0; i < r.n_faces; ++i) faces[i] = b[i];
    });
}

int rin_result_chain_count(rin_ctx* c, const rin_result* res, uint64_t* n_chains) {
    if (!res || !n_chains) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;  // the chains are cached in the result
        rin::compute_chains(c->mem, r, c->stream);
        *n_chains = r.n_chains;
    });
}

int rin_result_chains(rin_ctx* c, const rin_result* res, uint64_t* start, uint64_t* edges) {
    if (!res || !start) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        auto& r = const_cast<rin_result*>(res)->r;
        rin::compute_chains(c->mem, r, c->stream);
        auto a = fetch(r.c_start.p, r.n_chains + 1, c->stream);
        const uint64_t n = a[r.n_chains];
        if (n && !edges) throw std::invalid_argument("null argument");
        auto b = fetch(r.c_edges.p, n, c->stream);
        for (uint64_t i = 0; i <= r.n_chains; ++i) start[i] = a[i];
        for (uint64_t i = 0; i < n; ++i) edges[i] = b[i];
    });
}

int rin_result_records(rin_ctx* c, const rin_result* res, const uint64_t* tets, uint64_t n, uint64_t* rec_start,
                       uint16_t* words, uint64_t cap) {
    if (!res || (!tets && n) || !rec_start) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const auto& r = res->r;
        if (!r.view.rec) throw std::invalid_argument("records were not kept (RIN_FLAG_KEEP_RECORDS)");
        rin::gather_records(c->mem, r, tets, n, rec_start, words, cap, c->stream);
    });
}

int rin_result_cut_counts(rin_ctx* c, const rin_result* res, uint64_t* counts) {
    if (!res || !counts) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] {
        const auto& r = res->r;
        if (!r.view.cut_counts) throw std::invalid_argument("run was not RIN_FLAG_CUT_ONLY");
        auto h = fetch(r.view.cut_counts, 3 * r.n_active, c->stream);
        for (uint64_t i = 0; i < 3 * r.n_active; ++i) counts[i] = h[i];
    });
}

int rin_result_stage_times(rin_ctx* c, const rin_result* res, double* ms, int32_t cap) {
    if (!res) return fail(c, RIN_ERR_ARG, "null argument");
    const auto& t = res->r.times.ms;
    int n = 0;
    for (auto& e : t)
        if (n < cap && ms) ms[n++] = e.second;
    return (int)t.size();
}

const char* rin_result_stage_label(rin_ctx* c, const rin_result* res, int32_t i) {
    (void)c;
    if (!res || i < 0 || i >= (int)res->r.times.ms.size()) return "";
    return res->r.times.ms[i].first.c_str();
}

int rin_result_release(rin_ctx* c, rin_result* res) {
    return guarded(c, [&] {
        RIN_HIP(hipStreamSynchronize(c->stream));
        delete res;
    });
}

int rin_debug_cuts(rin_ctx* c, int32_t mi, const double* rows, const int32_t* ks, int32_t n, int32_t cap,
                   uint16_t* words, int32_t* lens) {
    if (!rows || !ks || !words || !lens || n < 0 || cap <= 8) return fail(c, RIN_ERR_ARG, "bad argument");
    return guarded(c, [&] { rin::debug_cuts(c->mem, mi, rows, ks, n, cap, words, lens, c->stream); });
}

int rin_debug_table_cuts(int32_t mode, const double* rows, int32_t n, int32_t cap, uint16_t* words, int32_t* lens) {
    if (!rows || !words || !lens || n < 0 || cap <= 8 || mode < 0 || mode > 3) return RIN_ERR_ARG;
    try {
        const int nr = mode == 0 ? 1 : (mode == 3 ? 3 : 2);
        for (int32_t i = 0; i < n; ++i) {
            const double* r = rows + (size_t)4 * nr * i;
            uint16_t* w = words + (size_t)cap * i;
            lens[i] = mode == 3   ? rin::debug_table_record_mi3(r, w, cap)
                      : mode == 2 ? rin::debug_table_record_mi2(r, w, cap)
                                  : rin::debug_table_record_ia(nr, r, w, cap);
        }
        return RIN_OK;
    } catch (std::exception&) {
        return RIN_ERR_INTERNAL;
    }
}

int rin_debug_copy(rin_ctx* c, uint64_t bytes, int32_t reps, double* ms) {
    if (!ms) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] { *ms = rin::debug_copy(c->mem, bytes, reps, c->stream); });
}

int rin_debug_sqrt(rin_ctx* c, const double* in, uint64_t n, double* out) {
    if ((!in || !out) && n) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] { rin::debug_sqrt(c->mem, in, n, out, c->stream); });
}

int rin_debug_scan(rin_ctx* c, int32_t fields, int32_t stride, const uint32_t* in, uint64_t n, uint64_t count,
                   uint32_t* out, uint32_t* totals) {
    if ((!in && n) || (!out && n) || !totals) return fail(c, RIN_ERR_ARG, "null argument");
    return guarded(c, [&] { rin::debug_scan(c->mem, fields, stride, in, n, count, out, totals, c->stream); });
}

}  // extern "C"
