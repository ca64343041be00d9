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
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>

#include "cut_wave.h"
#include "kernels_ia.h"
#include "kernels_mi.h"

namespace rin {

namespace {
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

// vertex dedup shared by both modes: stable sort of the occurrence keys, then
// each key's id is the rank of its first occurrence in occurrence order
// the 2-material records, one per corner comparison pattern, from the cut
// itself (host build of the device algorithm)
struct MI2Table {
    MI2Entry e[81];
    std::vector<uint16_t> words;
};

const MI2Table& mi2_table() {
    static MI2Table T;
    static std::once_flag once;
    std::call_once(once, [] {
        static dev::Complex<CMed> X;
        for (int pat = 0; pat < 81; ++pat) {
            // material a is 0 at every corner; material b is -1, 0 or 1
            double P[2][4];
            for (int c = 0, q = pat; c < 4; ++c, q /= 3) {
                P[0][c] = 0.0;
                P[1][c] = (double)(q % 3 - 1);
            }
            dev::mi_cut<CMed>(P, 2, X);
            if (X.err) throw std::runtime_error("2-material table: cut failed");
            uint16_t rec[1024];
            const int len = dev::write_record_mi<CMed>(X, 2, rec, 1024);
            if (len < 0 || T.words.size() + len > 0xFFFF) throw std::runtime_error("2-material table: record");
            MI2Entry& E = T.e[pat];
            E.off = (uint16_t)T.words.size();
            E.len = (uint16_t)len;
            T.words.insert(T.words.end(), rec, rec + len);
            mi_counts<CMed>(X, E.cnt[0], E.cnt[1], E.cnt[2]);
            E.cut[0] = (uint32_t)X.nv;
            E.cut[1] = (uint32_t)X.nf;
            E.cut[2] = (uint32_t)X.nc;
        }
    });
    return T;
}

// the 3-material table: per 2-material pattern, every sign vector of the
// third material over that complex's vertices, cut once by the insertion
// itself; identical records are stored once
struct MI3Table {
    MI3Pat pat[81];
    uint32_t nslots = 0;
    std::vector<uint32_t> off, cnt, cut;
    std::vector<uint16_t> words;
    // the complex after the three materials, per slot (byte offset, or
    // MI3_NONE; equal complexes stored once): the state the lane-group cut of
    // 4..8 materials continues from (layout: kMI3SnapHead, kMI3SnapFace)
    std::vector<uint32_t> snap_off;
    std::vector<uint8_t> snap;
};

const MI3Table& mi3_table() {
    static MI3Table T;
    static std::once_flag once;
    std::call_once(once, [] {
        std::memset(T.pat, 0, sizeof(T.pat));
        static dev::Complex<CMed> X1, X2;
        auto two = [&](int pat) {
            double P[2][4];
            for (int c = 0, q = pat; c < 4; ++c, q /= 3) {
                P[0][c] = 0.0;
                P[1][c] = (double)(q % 3 - 1);
            }
            dev::mi_cut<CMed>(P, 2, X1);
            if (X1.err) throw std::runtime_error("3-material table: 2-material cut failed");
        };
        uint32_t base = 0;
        for (int pat = 0; pat < 81; ++pat) {
            two(pat);
            if (X1.nv > 8) throw std::runtime_error("3-material table: more than 8 vertices");
            MI3Pat& K = T.pat[pat];
            K.valid = 1;
            K.nv = (uint8_t)X1.nv;
            K.base = base;
            for (int v = 0; v < X1.nv; ++v)
                for (int q = 0; q < 4; ++q) K.lab[v][q] = X1.vl[v][q];
            uint32_t p3 = 1;
            for (int v = 0; v < X1.nv; ++v) p3 *= 3;
            base += p3;
        }
        T.nslots = base;
        T.off.assign(base, MI3_NONE);
        T.cnt.assign(base, 0);
        T.cut.assign(base, 0);
        T.snap_off.assign(base, MI3_NONE);
        std::map<std::vector<uint16_t>, uint32_t> seen;
        std::map<std::vector<uint8_t>, uint32_t> seen_snap;
        for (int pat = 0; pat < 81; ++pat) {
            const MI3Pat& K = T.pat[pat];
            two(pat);
            uint32_t p3 = 1;
            for (int v = 0; v < K.nv; ++v) p3 *= 3;
            for (uint32_t sv = 0; sv < p3; ++sv) {
                int8_t sg[CMed::MV];
                for (int v = 0, r = (int)sv; v < K.nv; ++v, r /= 3) sg[v] = (int8_t)(r % 3 - 1);
                X2 = X1;
                dev::mi_insert<CMed>(X2, 6, sg);
                if (X2.err) continue;  // no tet has this sign vector; the general cut would fail alike
                uint16_t rec[1024];
                const int len = dev::write_record_mi<CMed>(X2, 3, rec, 1024);
                if (len < 0) throw std::runtime_error("3-material table: record overflow");
                std::vector<uint16_t> key(rec, rec + len);
                auto it = seen.find(key);
                uint32_t off;
                if (it == seen.end()) {
                    off = (uint32_t)T.words.size();
                    T.words.insert(T.words.end(), rec, rec + len);
                    seen.emplace(std::move(key), off);
                } else {
                    off = it->second;
                }
                uint32_t a0, a1, a2;
                mi_counts<CMed>(X2, a0, a1, a2);
                if (a0 > 255 || a1 > 255 || X2.nv > 255 || X2.nf > 255) throw std::runtime_error("3-material table");
                T.off[K.base + sv] = off;
                T.cnt[K.base + sv] = a0 | a1 << 8 | a2 << 16;
                T.cut[K.base + sv] = (uint32_t)X2.nv | (uint32_t)X2.nf << 8 | (uint32_t)X2.nc << 16;
                static_assert(CMed::MP == 8 && CMed::MFV == 8, "3-material snapshot layout");
                std::vector<uint8_t> sn = {(uint8_t)X2.nv, (uint8_t)X2.nf, (uint8_t)X2.nc, (uint8_t)X2.dup};
                for (int p = 0; p < 8; ++p) sn.push_back(X2.grp[p]);
                for (int p = 0; p < 8; ++p) sn.push_back(p < X2.nc ? X2.cells[p] : (uint8_t)dev::NIL);
                for (int v = 0; v < X2.nv; ++v)
                    for (int q = 0; q < 4; ++q) sn.push_back(X2.vl[v][q]);
                for (int f = 0; f < X2.nf; ++f) {
                    sn.push_back(X2.fa[f]);
                    sn.push_back(X2.fb[f]);
                    sn.push_back(X2.fn[f]);
                    sn.push_back(0);
                    for (int q = 0; q < 8; ++q) sn.push_back(q < X2.fn[f] ? X2.fv[f][q] : (uint8_t)0);
                    for (int q = 0; q < 8; ++q) sn.push_back(q < X2.fn[f] ? X2.fe[f][q] : (uint8_t)0);
                }
                auto is = seen_snap.find(sn);
                if (is == seen_snap.end()) {
                    const uint32_t so = (uint32_t)T.snap.size();
                    T.snap.insert(T.snap.end(), sn.begin(), sn.end());
                    is = seen_snap.emplace(std::move(sn), so).first;
                }
                T.snap_off[K.base + sv] = is->second;
            }
        }
    });
    return T;
}

}  // namespace

// the material-interface tables on each device, uploaded once per process
// and device (they stay until the process ends)
MITablesRaw mi_tables_device(hipStream_t s) {
    static std::mutex mu;
    static std::map<int, MITablesRaw> per_device;
    int dev = 0;
    RIN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(dev);
    if (it != per_device.end()) return it->second;
    const MI2Table& T2 = mi2_table();
    const MI3Table& T3 = mi3_table();
    auto up = [&](const void* src, size_t bytes) {
        void* p = nullptr;
        RIN_HIP(hipMalloc(&p, bytes ? bytes : 4));
        if (bytes) RIN_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s));
        return p;
    };
    MITablesRaw D;
    D.e2 = up(T2.e, sizeof(T2.e));
    D.w2 = up(T2.words.data(), 2 * T2.words.size());
    D.pat3 = up(T3.pat, sizeof(T3.pat));
    D.off3 = up(T3.off.data(), 4 * T3.off.size());
    D.cnt3 = up(T3.cnt.data(), 4 * T3.cnt.size());
    D.cut3 = up(T3.cut.data(), 4 * T3.cut.size());
    D.w3 = up(T3.words.data(), 2 * T3.words.size());
    D.snap_off3 = up(T3.snap_off.data(), 4 * T3.snap_off.size());
    D.snap3 = up(T3.snap.data(), T3.snap.size());
    D.n_w2 = T2.words.size();
    D.n_w3 = T3.words.size();
    RIN_HIP(hipStreamSynchronize(s));
    per_device[dev] = D;
    return D;
}

namespace {

MITablesDev mi_tables_dev(hipStream_t s) { return mi_tables_view(mi_tables_device(s)); }

struct Dedup {
    DBuf<uint32_t> vid;        // id of every occurrence
    DBuf<uint32_t> first_occ;  // first occurrence of every id
    uint64_t n = 0;
};
void dedup_vertices(DevMem& mem, DBuf<uint64_t>& okey, uint64_t NVO, int kbits, Dedup& d, hipStream_t s) {
    if (kbits > 64) throw std::runtime_error("simplex keys exceed 64 bits for this mesh");
    DBuf<uint32_t> sidx(mem, NVO + 1), first(mem, NVO + 1), head(mem, NVO + 1), hexcl(mem, NVO + 1),
        seg_first(mem, NVO + 1), vrank(mem, NVO + 1), nvt(mem, 2);
    d.vid.alloc(mem, NVO + 1);
    iota_u32(sidx.p, NVO, s);
    sort_pairs(mem, okey.p, sidx.p, NVO, kbits, s);
    hipLaunchKernelGGL(k_heads, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint64_t*)okey.p,
                       (const uint32_t*)sidx.p, first.p, head.p);
    scan_u32(mem, head.p, hexcl.p, NVO, nullptr, s);
    hipLaunchKernelGGL(k_assign, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, seg_first.p);
    scan_u32(mem, first.p, vrank.p, NVO, nvt.p, s);
    hipLaunchKernelGGL(k_assign2, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, (const uint32_t*)seg_first.p,
                       (const uint32_t*)vrank.p, d.vid.p);
    d.n = read1(nvt.p, s);
    d.first_occ.alloc(mem, d.n + 1);
    hipLaunchKernelGGL(k_first_list, dim3(nblk(NVO)), dim3(256), 0, s, NVO, (const uint32_t*)first.p,
                       (const uint32_t*)vrank.p, d.first_occ.p);
    RIN_HIP(hipGetLastError());
}

}  // namespace

// tet adjacency across faces, once per grid
void ensure_nbr(DevMem& mem, const Grid& g, hipStream_t s) {
    if (g.nbr.p) return;
    const uint64_t n = 4 * g.view.T;
    if (n >= (1ull << 32)) throw std::runtime_error("too many tets for the adjacency table");
    DBuf<uint64_t> key(mem, n + 1);
    DBuf<uint32_t> slot(mem, n + 1);
    g.nbr.alloc(mem, n + 1);
    fill_u32(g.nbr.p, NO_NBR, n + 1, s);
    hipLaunchKernelGGL(k_face_slots, dim3(nblk(g.view.T)), dim3(256), 0, s, g.view, key.p, slot.p);
    sort_pairs(mem, key.p, slot.p, Count(n), bits_for(g.view.n_face_ids), s);
    hipLaunchKernelGGL(k_pair_faces, dim3(nblk(n)), dim3(256), 0, s, n, (const uint64_t*)key.p,
                       (const uint32_t*)slot.p, g.nbr.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipStreamSynchronize(s));
}

namespace {

// material-interface extraction and coordinates (src/extract_mesh.cpp:569-986,
// 1541-1637); faces are never merged, so they need no sort
void extract_mi(DevMem& mem, const Grid& g, const Values& val, RunOut& out, Timer& tm, hipStream_t s) {
    const uint64_t NA = out.n_active;
    ensure_nbr(mem, g, s);
    MIArgs A;
    A.g = g.view;
    A.act = out.act.p;
    A.csr_off = out.csr_off.p;
    A.rank = out.rank_of_tet.p;
    A.rec_loc = out.rec_loc.p;
    A.nbr = g.nbr.p;
    A.fit = out.fit.p;
    A.rec = out.rec.p;
    A.n_active = Count(NA);
    DBuf<uint32_t> cnt(mem, 3 * NA + 3), cv(mem, NA + 1), cf(mem, NA + 1), cfv(mem, NA + 1);
    DBuf<uint64_t> emask(mem, NA + 1);
    if (NA) {
        // every face's keys once, then the boundary-face pairing from them
        DBuf<uint32_t> fb(mem, NA + 1);
        hipLaunchKernelGGL(k_mi_nf, dim3(nblk(NA)), dim3(256), 0, s, A, fb.p);
        scan_u32(mem, fb.p, fb.p, NA, fb.p + NA, s);
        const uint64_t NFT = read1(fb.p + NA, s);
        DBuf<uint16_t> fl(mem, NFT + 1), fw(mem, NFT + 1);
        DBuf<uint64_t> ft(mem, 3 * NFT + 3);
        MIFaceKeys K;
        K.fb = fb.p;
        K.fl = fl.p;
        K.fw = fw.p;
        K.ft = ft.p;
        hipLaunchKernelGGL(k_mi_face_keys, dim3(nblk(NA * kBfaceLanes)), dim3(256), 0, s, A, K);
        hipLaunchKernelGGL(k_mi_bface, dim3(nblk(NA * kBfaceLanes)), dim3(256), 0, s, A, K, emask.p);
        RIN_HIP(hipGetLastError());
        RIN_HIP(hipStreamSynchronize(s));  // the face-key buffers return to the pool
        hipLaunchKernelGGL(k_mi_count, dim3(nblk(NA)), dim3(256), 0, s, A, (const uint64_t*)emask.p, cnt.p);
    }
    hipLaunchKernelGGL(k_split3, dim3(nblk(NA)), dim3(256), 0, s, Count(NA), (const uint32_t*)cnt.p, cv.p, cf.p,
                       cfv.p);
    scan_u32(mem, cv.p, cv.p, NA, cv.p + NA, s);
    scan_u32(mem, cf.p, cf.p, NA, cf.p + NA, s);
    scan_u32(mem, cfv.p, cfv.p, NA, cfv.p + NA, s);
    const uint64_t NVO = read1(cv.p + NA, s), NF = read1(cf.p + NA, s), NFV = read1(cfv.p + NA, s);
    out.n_vert_occ = NVO;
    out.n_face_occ = NF;
    out.n_face_occ_verts = NFV;
    DBuf<uint64_t> okey(mem, NVO + 1);
    DBuf<uint32_t> oa(mem, NVO + 1), fq_a(mem, NF + 1), fq_vs(mem, NF + 1), fve(mem, NFV + 1);
    DBuf<uint16_t> oj(mem, NVO + 1), fq_l(mem, NF + 1);
    DBuf<uint8_t> fq_m(mem, 2 * NF + 2);
    if (NA)
        hipLaunchKernelGGL(k_mi_extract, dim3(nblk(NA)), dim3(256), 0, s, A, (const uint64_t*)emask.p,
                           (const uint32_t*)cv.p,
                           (const uint32_t*)cf.p, (const uint32_t*)cfv.p, okey.p, oa.p, oj.p, fq_a.p, fq_l.p, fq_m.p,
                           fq_vs.p, fve.p);
    RIN_HIP(hipGetLastError());
    tm.mark("extract occurrences");
    const int kbits = 2 + std::max({bits_for(g.view.V), bits_for(g.view.n_edge_ids) + 12,
                                    bits_for(g.view.n_face_ids) + 18, bits_for(NA) + 10});
    Dedup d;
    dedup_vertices(mem, okey, NVO, kbits, d, s);
    tm.mark("dedup verts");
    const uint64_t NV = d.n;
    out.n_verts = NV;
    out.v_tet.alloc(mem, NV + 1);
    out.v_j.alloc(mem, NV + 1);
    out.v_ss.alloc(mem, NV + 1);
    out.v_sv.alloc(mem, 4 * NV + 4);
    out.v_fn.alloc(mem, 4 * NV + 4);
    out.xyz.alloc(mem, 3 * NV + 3);
    hipLaunchKernelGGL(k_vert_fields_mi, dim3(nblk(NV)), dim3(256), 0, s, g.view, (const double*)val.vals.p,
                       (const uint32_t*)out.act.p, (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p,
                       (const uint32_t*)out.rec_loc.p, (const uint16_t*)out.rec.p, (const uint32_t*)d.first_occ.p,
                       (const uint32_t*)oa.p, (const uint16_t*)oj.p, Count(NV), out.v_tet.p, out.v_j.p, out.v_ss.p,
                       out.v_sv.p, out.v_fn.p, out.xyz.p);
    RIN_HIP(hipGetLastError());
    tm.mark("compute xyz");
    out.n_faces = NF;
    out.n_face_verts = NFV;
    out.n_tet_faces = NF;
    out.f_vstart.alloc(mem, NF + 1);
    out.f_verts.alloc(mem, NFV + 1);
    out.f_tfstart.alloc(mem, NF + 1);
    out.f_tf_tet.alloc(mem, NF + 1);
    out.f_tf_local.alloc(mem, NF + 1);
    out.f_func.alloc(mem, 2 * NF + 2);
    hipLaunchKernelGGL(k_mi_face_out, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), (const uint32_t*)out.act.p,
                       (const uint32_t*)fq_a.p, (const uint16_t*)fq_l.p, (const uint8_t*)fq_m.p,
                       (const uint32_t*)fq_vs.p, (const uint32_t*)fve.p, (const uint32_t*)d.vid.p, Count(NFV),
                       out.f_vstart.p, out.f_verts.p, out.f_tfstart.p, out.f_tf_tet.p, out.f_tf_local.p,
                       out.f_func.p);
    fill_u32(out.f_vstart.p + NF, (uint32_t)NFV, 1, s);
    fill_u32(out.f_tfstart.p + NF, (uint32_t)NF, 1, s);
    RIN_HIP(hipGetLastError());
    tm.mark("faces");
}

}  // namespace

// ------------------------------------------------------------------ API

namespace {
std::atomic<uint64_t> g_grid_ids{0};
}

void grid_regular(DevMem& mem, int res, const double bmin[3], const double bmax[3], Grid& g, hipStream_t s) {
    if (res < 1) throw std::invalid_argument("grid resolution must be >= 1");
    g.id = ++g_grid_ids;
    g.plans = make_plan_cache();
    g.mi_plans = make_mi_plan_cache();
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
    g.id = ++g_grid_ids;
    g.plans = make_plan_cache();
    g.mi_plans = make_mi_plan_cache();
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
    v.grid_id = g.id;
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

void eval_points(DevMem& mem, const double* pts, uint64_t n, const FuncDesc* funcs, uint32_t F, double* rowmajor,
                 hipStream_t s) {
    if (F == 0) throw std::invalid_argument("no functions");
    if (n == 0) return;
    DBuf<double> dp(mem, 3 * n), dv(mem, (size_t)F * n), drm(mem, (size_t)F * n);
    DBuf<FuncDesc> fd(mem, F);
    RIN_HIP(hipMemcpyAsync(dp.p, pts, 24 * n, hipMemcpyHostToDevice, s));
    RIN_HIP(hipMemcpyAsync(fd.p, funcs, sizeof(FuncDesc) * F, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_eval, dim3(nblk(n)), dim3(256), 0, s, (const double*)dp.p, n, (const FuncDesc*)fd.p, F, dv.p);
    hipLaunchKernelGGL(k_transpose_out, dim3(nblk(F * n)), dim3(256), 0, s, (const double*)dv.p, n, F, drm.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(rowmajor, drm.p, 8 * (size_t)F * n, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

void values_upload(DevMem& mem, const Grid& g, const double* rowmajor, uint32_t F, Values& v, hipStream_t s) {
    if (F == 0 || F > kMaxFuncs) throw std::invalid_argument("function count must be 1..64");
    v.grid_id = g.id;
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
                           vflag.p, (uint32_t*)nullptr);
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

    // bucket counts (reference labels: IA 1/2/>=3, MI 2/3/>=4) and the cut
    // lists, built on the device
    // MI with lookups on: two-material tets take their record from the table
    // IA with lookups on: one- and two-function tets likewise (their records
    // from the k = 1 and k = 2 tables)
    // (the IA table kernel writes output counts, so the -R check's complex
    // counts take the general cut)
    const bool use_tables = !(flags & (RUN_REVERSE | RUN_NO_LOOKUP)) && (MI || !(flags & RUN_CUT_ONLY));
    const int table_mode = use_tables ? (MI ? 1 : 2) : 0;
    DBuf<uint32_t> blists(mem, 4 * (NA + 1));
    uint32_t* blist[4];
    for (int q = 0; q < 4; ++q) blist[q] = blists.p + (uint64_t)q * (NA + 1);
    uint32_t bcnt[4] = {0, 0, 0, 0};  // list sizes: two-material table, up to 4, 5..8, more
    if (NA) {
        const uint32_t nch = (uint32_t)((NA + kBucketChunk - 1) / kBucketChunk);
        DBuf<uint32_t> bc(mem, (uint64_t)kBucketWords * nch + kBucketWords), bo(mem, 16);
        hipLaunchKernelGGL(k_bucket_count, dim3(nch), dim3(256), 0, s, Count(NA), (const uint8_t*)out.act_k.p, (int)MI,
                           table_mode, bc.p);
        hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, nch, bc.p, bo.p);
        hipLaunchKernelGGL(k_bucket_fill, dim3(nch), dim3(256), 0, s, Count(NA), (const uint8_t*)out.act_k.p, table_mode,
                           (const uint32_t*)bc.p, blist[0], blist[1], blist[2], blist[3]);
        RIN_HIP(hipGetLastError());
        uint32_t hb[16];
        RIN_HIP(hipMemcpyAsync(hb, bo.p, sizeof(hb), hipMemcpyDeviceToHost, s));
        RIN_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < 4; ++q) bcnt[q] = hb[q];
        for (int q = 0; q < 3; ++q) out.nb[q] = hb[4 + q];
        out.n_cut34 = hb[7];
        out.n_cut58 = hb[8];
        if ((int)hb[9] > kMaxHuge) {
            out.err |= ERR_CAPACITY;
            throw std::runtime_error("a tet has " + std::to_string(hb[9]) +
                                     " active functions; the device cut supports at most " +
                                     std::to_string(kMaxHuge));
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
    if (bcnt[0] && !MI) {
        table_cut_plain(g.view, (const double*)val.vals.p, (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                        (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, (const uint32_t*)blist[0],
                        (uint64_t)bcnt[0], (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p, s);
    } else if (bcnt[0]) {
        // two- and three-material tets by table
        const MITablesDev TT = mi_tables_dev(s);
        hipLaunchKernelGGL(k_mi_table, dim3(nblk(bcnt[0])), dim3(256), 0, s, g.view, (const double*)val.vals.p,
                           (const uint32_t*)out.act.p, (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p,
                           (const uint32_t*)blist[0], (uint64_t)bcnt[0], (const uint32_t*)nullptr, flags,
                           (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p, TT, (uint32_t*)nullptr, 0u, 0u);
        RIN_HIP(hipGetLastError());
    }
    // the cuts beyond the tables: lane groups with the complex in LDS (the
    // -R check's counts and reversed order take the thread cut)
    const bool groups = !(flags & (RUN_REVERSE | RUN_CUT_ONLY));
    // the implicit-arrangement lane groups start from the k = 2 table's
    // complex after two planes (RIN_K2_START=0: they insert every plane)
    K2Snap k2s{nullptr, nullptr, nullptr};
    const char* ks = std::getenv("RIN_K2_START");
    if (!MI && groups && !(ks && !std::strcmp(ks, "0"))) {
        const void* pat = nullptr;
        plain_k2_snapshots(s, &pat, &k2s.off, &k2s.bytes);
        k2s.pat = static_cast<const K2Pat*>(pat);
    }
    // the material-interface lane groups start from the 3-material table's
    // complex after three materials (RIN_MI3_START=0: they insert every one)
    MI3Snap m3s{nullptr, nullptr, nullptr};
    const char* ms = std::getenv("RIN_MI3_START");
    if (MI && groups && !(ms && !std::strcmp(ms, "0"))) {
        const MITablesDev TT = mi_tables_dev(s);
        m3s = MI3Snap{TT.pat3, TT.snap_off3, TT.snap3};
    }
    for (int q = 0; q < 2 && NA && groups; ++q) {
        const uint64_t nl = bcnt[1 + q];
        if (!nl) continue;
        auto go = [&](auto kernel, int per_wave, auto snap) {
            const unsigned blocks = (unsigned)std::min<uint64_t>((nl + per_wave - 1) / per_wave, 4096);
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64), 0, s, g.view, (const double*)val.vals.p,
                               (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p,
                               (const uint32_t*)blist[1 + q], Count(nl), (const uint32_t*)out.rec_off.p, out.rec.p,
                               cnt.p, errs.p, 1, snap);
        };
        if (q == 0)
            MI ? go(k_cut_group<CMed, 64, true>, 1, m3s) : go(k_cut_group<CMed, 16, false>, 4, k2s);
        else
            MI ? go(k_cut_group<CBig, 64, true>, 1, m3s) : go(k_cut_group<CBig, 64, false>, 1, k2s);
        RIN_HIP(hipGetLastError());
    }
    // each thread-per-tet cut kernel runs over its own k range only
    DBuf<uint8_t> ws_cut[2];
    for (int q = 0; q < 2 && NA && !groups; ++q) {
        const uint64_t nl = bcnt[1 + q];
        if (!nl) continue;
        // lanes per workgroup (one wavefront): a lane runs a whole cut, whose
        // loops diverge between lanes, so a short list runs faster with few
        // lanes per wavefront as long as its wavefronts all fit on the GPU
        // at once (about 2048 of these)
        unsigned lanes = 2;
        while (lanes < 64 && nl > (uint64_t)lanes * 2048) lanes *= 2;
        // each lane's complex and working arrays in a device workspace; the
        // grid strides over the list beyond kCutLanes lanes
        const uint64_t nb = (nl + lanes - 1) / lanes;
        const unsigned blocks = (unsigned)std::min<uint64_t>(nb, kCutLanes / lanes);
        const int kmin = q == 0 ? 1 : 5, kmax = q == 0 ? 4 : 8;
        const size_t stride = q == 0 ? LaneCut<CMed>::stride : LaneCut<CBig>::stride;
        ws_cut[q].alloc(mem, stride * blocks * lanes);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(lanes), 0, s, g.view, (const double*)val.vals.p,
                               (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p,
                               (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, Count(nl), kmin, kmax, flags,
                               (const uint32_t*)out.rec_off.p, out.rec.p, cnt.p, errs.p,
                               (const uint32_t*)blist[1 + q], ws_cut[q].p);
        };
        if (q == 0)
            MI ? go(k_cut<CMed, true>) : go(k_cut<CMed, false>);
        else
            MI ? go(k_cut<CBig, true>) : go(k_cut<CBig, false>);
        RIN_HIP(hipGetLastError());
    }
    // rare tets: one lane each, complexes in device memory; list: device
    auto cut_large = [&](const uint32_t* list, uint64_t nl, const uint32_t* loc, uint32_t fixed_cap) {
        // ~200 KB of complex per lane: a few blocks
        const unsigned blocks = (unsigned)std::min<uint64_t>((nl + 63) / 64, 4);
        DBuf<uint8_t> ws(mem, LaneCut<CHuge>::stride * 64 * blocks);
        uint8_t* wsp = ws.p;
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64), 0, s, g.view, (const double*)val.vals.p,
                               (const uint32_t*)out.act.p, (const uint8_t*)out.act_k.p, (const uint32_t*)out.csr_off.p,
                               (const uint8_t*)out.fit.p, list, Count(nl), flags, (const uint32_t*)out.rec_off.p, out.rec.p,
                               cnt.p, errs.p, wsp, loc, fixed_cap);
        };
        MI ? go(k_cut_huge<true>) : go(k_cut_huge<false>);
        RIN_HIP(hipGetLastError());
        RIN_HIP(hipStreamSynchronize(s));  // the workspace returns to the pool
    };
    if (bcnt[3]) cut_large(blist[3], bcnt[3], nullptr, 0);
    // record location of every active tet: its slot, or (after a retry) a slot
    // appended for a large complex
    out.rec_loc.alloc(mem, NA + 1);
    if (NA) copy_u32(out.rec_loc.p, out.rec_off.p, Count(NA), s);
    auto read_errs = [&] {
        uint32_t he[64], e = 0;
        RIN_HIP(hipMemcpyAsync(he, errs.p, sizeof(he), hipMemcpyDeviceToHost, s));
        RIN_HIP(hipStreamSynchronize(s));
        for (uint32_t x : he) e |= x;
        return e;
    };
    uint32_t e = read_errs();
    if (e == ERR_RETRY) {
        // complexes that outgrew a bucket kernel: cut again with the
        // large-capacity kernel into slots appended to the record buffer
        std::vector<uint32_t> hc(3 * NA);
        RIN_HIP(hipMemcpyAsync(hc.data(), cnt.p, 4 * hc.size(), hipMemcpyDeviceToHost, s));
        std::vector<uint32_t> hl(NA);
        RIN_HIP(hipMemcpyAsync(hl.data(), out.rec_off.p, 4 * NA, hipMemcpyDeviceToHost, s));
        RIN_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> retry;
        for (uint64_t a = 0; a < NA; ++a)
            if (hc[3 * a] == CNT_RETRY) retry.push_back((uint32_t)a);
        constexpr uint32_t slot = 8 + 4 * CHuge::MV + CHuge::MF * (4 + CHuge::MFV) + 2 * CHuge::MP;
        const uint64_t total = nwords + (uint64_t)slot * retry.size();
        if (total >= (1ull << 31)) throw std::runtime_error("record buffer exceeds 2^31 words");
        DBuf<uint16_t> rec2(mem, total + 1);
        RIN_HIP(hipMemcpyAsync(rec2.p, out.rec.p, 2 * nwords, hipMemcpyDeviceToDevice, s));
        for (size_t i = 0; i < retry.size(); ++i) hl[retry[i]] = (uint32_t)(nwords + slot * i);
        RIN_HIP(hipMemcpyAsync(out.rec_loc.p, hl.data(), 4 * NA, hipMemcpyHostToDevice, s));
        RIN_HIP(hipStreamSynchronize(s));
        out.rec = std::move(rec2);
        fill_u32(errs.p, 0, 64, s);
        DBuf<uint32_t> rl(mem, retry.size());
        RIN_HIP(hipMemcpyAsync(rl.p, retry.data(), 4 * retry.size(), hipMemcpyHostToDevice, s));
        cut_large(rl.p, retry.size(), out.rec_loc.p, slot);
        e = read_errs();
    }
    tm.mark(MI ? "MI cut" : "simp_arr");
    out.err |= e;
    if (out.err) {
        tm.finish(out.times);
        throw std::runtime_error(std::string("device cut failed: ") +
                                 ((out.err & ERR_CAPACITY) ? "complex exceeds capacity or predicate range; " : "") +
                                 ((out.err & ERR_RECORD) ? "record overflow" : ""));
    }
    if (flags & RUN_CUT_ONLY) {
        out.cut_counts = std::move(cnt);
        tm.finish(out.times);
        out.set_views_from_buffers();
        return;
    }
    if (MI) {
        extract_mi(mem, g, val, out, tm, s);
        tm.finish(out.times);
        if (!(flags & RUN_KEEP_RECORDS)) out.rec.reset();
        out.set_views_from_buffers();
        return;
    }

    // ---- extraction
    DBuf<uint32_t> cv(mem, NA + 1), cf(mem, NA + 1), cfv(mem, NA + 1);
    hipLaunchKernelGGL(k_split3, dim3(nblk(NA)), dim3(256), 0, s, NA, (const uint32_t*)cnt.p, cv.p, cf.p, cfv.p);
    scan_u32(mem, cv.p, cv.p, NA, cv.p + NA, s);
    scan_u32(mem, cf.p, cf.p, NA, cf.p + NA, s);
    scan_u32(mem, cfv.p, cfv.p, NA, cfv.p + NA, s);
    const uint64_t NVO = read1(cv.p + NA, s), NFO = read1(cf.p + NA, s), NFV = read1(cfv.p + NA, s);
    out.n_vert_occ = NVO;
    out.n_face_occ = NFO;
    out.n_face_occ_verts = NFV;
    DBuf<uint64_t> okey(mem, NVO + 1);
    DBuf<uint32_t> oa(mem, NVO + 1);
    DBuf<uint16_t> oj(mem, NVO + 1);
    DBuf<uint32_t> fq_a(mem, NFO + 1), fq_vs(mem, NFO + 1), fve(mem, NFV + 1);
    DBuf<uint16_t> fq_l(mem, NFO + 1);
    DBuf<uint8_t> fq_fn(mem, NFO + 1), fq_bnd(mem, NFO + 1);
    if (NA)
        hipLaunchKernelGGL(k_extract_ia, dim3(nblk(NA)), dim3(256), 0, s, g.view, (const uint32_t*)out.act.p,
                           (const uint32_t*)out.csr_off.p, (const uint8_t*)out.fit.p, NA,
                           (const uint32_t*)out.rec_loc.p, (const uint16_t*)out.rec.p, (const uint32_t*)cv.p,
                           (const uint32_t*)cf.p, (const uint32_t*)cfv.p, okey.p, oa.p, oj.p, fq_a.p, fq_l.p,
                           fq_fn.p, fq_bnd.p, fq_vs.p, fve.p);
    RIN_HIP(hipGetLastError());
    tm.mark("extract occurrences");

    // ---- vertex dedup: stable sort by key, first occurrence ranks
    const int kbits = 2 + std::max({bits_for(V), bits_for(g.view.n_edge_ids) + 6, bits_for(g.view.n_face_ids) + 12,
                                    bits_for(NA) + 10});
    Dedup dd;
    dedup_vertices(mem, okey, NVO, kbits, dd, s);
    DBuf<uint32_t>& vid = dd.vid;
    DBuf<uint32_t>& first_occ = dd.first_occ;
    out.n_verts = dd.n;
    const uint64_t NV = out.n_verts;
    tm.mark("dedup verts");
    out.v_tet.alloc(mem, NV + 1);
    out.v_j.alloc(mem, NV + 1);
    out.v_ss.alloc(mem, NV + 1);
    out.v_sv.alloc(mem, 4 * NV + 4);
    out.v_fn.alloc(mem, 4 * NV + 4);
    out.xyz.alloc(mem, 3 * NV + 3);
    {
        VertFieldArgs VF;
        VF.g = g.view;
        VF.vals = (const double*)val.vals.p;
        VF.act = out.act.p;
        VF.off = out.csr_off.p;
        VF.per_rank = false;
        VF.fit = out.fit.p;
        VF.rec_loc = out.rec_loc.p;
        VF.rec = out.rec.p;
        VF.k1_words = nullptr;
        VF.k1_off = nullptr;
        VF.first_occ = first_occ.p;
        VF.oa = oa.p;
        VF.oj = oj.p;
        VF.nverts = Count(NV);
        VF.v_tet = out.v_tet.p;
        VF.v_j = out.v_j.p;
        VF.v_ss = out.v_ss.p;
        VF.v_sv = out.v_sv.p;
        VF.v_fn = out.v_fn.p;
        VF.xyz = out.xyz.p;
        hipLaunchKernelGGL(k_vert_fields_ia, dim3(nblk(NV)), dim3(256), 0, s, VF);
    }    RIN_HIP(hipGetLastError());
    tm.mark("compute xyz");

    // ---- faces: a boundary face may repeat in the tet across its tet face
    ensure_nbr(mem, g, s);
    DBuf<uint32_t> tot3(mem, 4);
    fill_u32(tot3.p, (uint32_t)NFV, 1, s);
    fill_u32(tot3.p + 1, (uint32_t)NFO, 1, s);
    hipLaunchKernelGGL(k_face_map, dim3(nblk(NFV)), dim3(256), 0, s, Count(NFV), fve.p, (const uint32_t*)vid.p);
    DBuf<uint32_t> ffirst(mem, NFO + 1), fmatch(mem, NFO + 1), ftfc(mem, NFO + 1), fvc(mem, NFO + 1),
        frank(mem, NFO + 1), ftpos(mem, NFO + 1), fvpos(mem, NFO + 1);
    FaceArgs FA;
    FA.fq_a = fq_a.p;
    FA.fq_vs = fq_vs.p;
    FA.fve = fve.p;
    FA.act = out.act.p;
    FA.rank = out.rank_of_tet.p;
    FA.foff = cf.p;
    FA.nbr = g.nbr.p;
    FA.fq_bnd = fq_bnd.p;
    FA.fq_l = fq_l.p;
    FA.nfv_total = tot3.p;
    FA.nfo_total = tot3.p + 1;
    FA.nfv_cap = NFV;
    FA.nfo_cap = NFO;
    FA.nq = Count(NFO);
    FA.n_active = Count(NA);
    FA.first = ffirst.p;
    FA.match = fmatch.p;
    FA.tfc = ftfc.p;
    FA.vcnt = fvc.p;
    hipLaunchKernelGGL(k_face_first, dim3(nblk(NFO)), dim3(256), 0, s, FA);
    DBuf<uint32_t> nft(mem, 4);
    scan_u32(mem, ffirst.p, frank.p, NFO, nft.p, s);
    scan_u32(mem, ftfc.p, ftpos.p, NFO, nft.p + 1, s);
    scan_u32(mem, fvc.p, fvpos.p, NFO, nft.p + 2, s);
    uint32_t hn[3];
    RIN_HIP(hipMemcpyAsync(hn, nft.p, 12, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    out.n_faces = hn[0];
    out.n_tet_faces = hn[1];
    out.n_face_verts = hn[2];
    const uint64_t NF = out.n_faces;
    out.f_vstart.alloc(mem, NF + 1);
    out.f_tfstart.alloc(mem, NF + 1);
    out.f_verts.alloc(mem, out.n_face_verts + 1);
    out.f_tf_tet.alloc(mem, out.n_tet_faces + 1);
    out.f_tf_local.alloc(mem, out.n_tet_faces + 1);
    out.f_func.alloc(mem, 2 * NF + 2);
    FaceOutArgs FO;
    FO.fq_a = fq_a.p;
    FO.fq_vs = fq_vs.p;
    FO.fve = fve.p;
    FO.act = out.act.p;
    FO.first = ffirst.p;
    FO.match = fmatch.p;
    FO.frank = frank.p;
    FO.tfpos = ftpos.p;
    FO.vpos = fvpos.p;
    FO.fq_l = fq_l.p;
    FO.fq_fn = fq_fn.p;
    FO.nfv_total = tot3.p;
    FO.nfv_cap = NFV;
    FO.nq = Count(NFO);
    FO.f_vstart = out.f_vstart.p;
    FO.f_verts = out.f_verts.p;
    FO.f_tfstart = out.f_tfstart.p;
    FO.f_tf_tet = out.f_tf_tet.p;
    FO.f_tf_local = out.f_tf_local.p;
    FO.f_func = out.f_func.p;
    hipLaunchKernelGGL(k_face_emit, dim3(nblk(NFO)), dim3(256), 0, s, FO);
    fill_u32(out.f_vstart.p + NF, (uint32_t)out.n_face_verts, 1, s);
    fill_u32(out.f_tfstart.p + NF, (uint32_t)out.n_tet_faces, 1, s);
    RIN_HIP(hipGetLastError());
    tm.mark("faces");
    tm.finish(out.times);
    if (!(flags & RUN_KEEP_RECORDS)) {
        // records stay only when the host graph will ask for them
        out.rec.reset();
    }
    out.set_views_from_buffers();
}

void RunOut::set_views_from_buffers() {
    view.fit = fit.p;
    view.csr_off = csr_off.p;
    view.act = act.p;
    view.rank_of_tet = rank_of_tet.p;
    view.rec_off = rec_off.p;
    view.rec = rec.p;
    view.rec_loc = rec_loc.p ? rec_loc.p : rec_off.p;
    view.k1_rec = nullptr;
    view.k1_rec_off = nullptr;
    view.v_tet = v_tet.p;
    view.v_sv = v_sv.p;
    view.v_j = v_j.p;
    view.v_ss = v_ss.p;
    view.v_fn = v_fn.p;
    view.xyz = xyz.p;
    view.f_vstart = f_vstart.p;
    view.f_verts = f_verts.p;
    view.f_tfstart = f_tfstart.p;
    view.f_tf_tet = f_tf_tet.p;
    view.f_tf_local = f_tf_local.p;
    view.f_func = f_func.p;
    view.cut_counts = cut_counts.p;
}

// ------------------------------------------------------------------ diagnostics

namespace {
template <class C, bool MI>
__global__ void k_debug_cut(const double* rows, const int* ks, const int* starts, int ncase, int cap, uint16_t* words,
                            int* lens, uint8_t* ws) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncase) return;
    const int k = ks[i];
    double P[C::MP - 4][4];
    for (int q = 0; q < k && q < C::MP - 4; ++q)
        for (int c = 0; c < 4; ++c) P[q][c] = rows[4 * (starts[i] + q) + c];
    Complex<C>& X = LaneCut<C>::X(ws, (uint64_t)i);
    if (k > C::MP - 4) {
        lens[i] = -2;
        return;
    }
    if (MI)
        dev::mi_cut<C>(P, k, X, LaneCut<C>::T(ws, (uint64_t)i));
    else
        dev::ia_cut<C>(P, k, X, LaneCut<C>::T(ws, (uint64_t)i));
    if (X.err) {
        lens[i] = -1;
        return;
    }
    lens[i] = MI ? dev::write_record_mi<C>(X, k, words + (size_t)i * cap, cap)
                 : dev::write_record_ia<C>(X, k, words + (size_t)i * cap, cap);
}
template <class C>
__global__ void __launch_bounds__(64) k_debug_cut_wave(const double* rows, const int* ks, const int* starts, int ncase,
                                                       int cap, uint16_t* words, int* lens) {
    __shared__ dev::WaveIA<C> X;
    const int i = blockIdx.x, lane = threadIdx.x;
    const int k = ks[i];
    if (k > C::MP - 4) {
        if (lane == 0) lens[i] = -2;
        return;
    }
    if (lane < 4 * k) X.P[lane >> 2][lane & 3] = rows[4 * starts[i] + lane];
    dev::wave_ia_cut<C>(X, k);
    if (X.err) {
        if (lane == 0) lens[i] = -1;
        return;
    }
    uint32_t a, b, c;
    const int len = dev::wave_ia_record<C>(X, k, words + (size_t)i * cap, cap, a, b, c);
    if (lane == 0) lens[i] = len;
}
// record lookup for a list of tets: source word offset (bit 63: k=1 table)
// and length, 0 for inactive tets
__global__ void k_rec_lookup(const uint32_t* __restrict__ tets, uint64_t n, uint64_t T, uint64_t n_active,
                             const uint32_t* __restrict__ rank, const uint32_t* __restrict__ act,
                             const uint32_t* __restrict__ rec_loc, const uint16_t* __restrict__ rec,
                             const uint16_t* __restrict__ k1_rec, const uint32_t* __restrict__ k1_off,
                             uint64_t* __restrict__ src, uint32_t* __restrict__ len) {
    GRID_STRIDE(i, n) {
        const uint32_t t = tets[i];
        uint32_t L = 0;
        uint64_t o = 0;
        if (t < T) {
            const uint32_t rk = rank[t];
            if (rk < n_active && act[rk] == t) {
                const uint32_t loc = rec_loc ? rec_loc[rk] : 0u;
                if (rec_loc && (loc & 0x80000000u)) {
                    o = (1ull << 63) | k1_off[loc & 0x7FFFFFFFu];
                    L = k1_rec[k1_off[loc & 0x7FFFFFFFu] + 6];
                } else {
                    o = loc;
                    L = rec[loc + 6];
                }
            }
        }
        src[i] = o;
        len[i] = L;
    }
}

__global__ void k_rec_copy(uint64_t n, const uint64_t* __restrict__ src, const uint32_t* __restrict__ len,
                           const uint32_t* __restrict__ dst, const uint16_t* __restrict__ rec,
                           const uint16_t* __restrict__ k1_rec, uint16_t* __restrict__ out) {
    // one wave per tet
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= n) return;
    const uint16_t* s = (src[w] >> 63) ? k1_rec + (src[w] & 0xFFFFFFFFull) : rec + src[w];
    for (uint32_t q = lane; q < len[w]; q += 64) out[dst[w] + q] = s[q];
}
}  // namespace

void gather_records(DevMem& mem, const RunOut& r, const uint64_t* tets, uint64_t n, uint64_t* rec_start,
                    uint16_t* words, uint64_t cap, hipStream_t s) {
    const auto& vw = r.view;
    rec_start[0] = 0;
    if (n == 0) return;
    std::vector<uint32_t> t32(n);
    for (uint64_t i = 0; i < n; ++i) {
        if (tets[i] >= r.T) throw std::invalid_argument("tet index out of range");
        t32[i] = (uint32_t)tets[i];
    }
    DBuf<uint32_t> dt(mem, n), dlen(mem, n), ddst(mem, n);
    DBuf<uint64_t> dsrc(mem, n);
    RIN_HIP(hipMemcpyAsync(dt.p, t32.data(), 4 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_rec_lookup, dim3(nblk(n)), dim3(256), 0, s, (const uint32_t*)dt.p, n, r.T, r.n_active,
                       vw.rank_of_tet, vw.act, vw.rec_loc, vw.rec, vw.k1_rec, vw.k1_rec_off, dsrc.p, dlen.p);
    std::vector<uint32_t> len(n);
    RIN_HIP(hipMemcpyAsync(len.data(), dlen.p, 4 * n, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> dst(n);
    uint64_t o = 0;
    for (uint64_t i = 0; i < n; ++i) {
        dst[i] = (uint32_t)o;
        o += len[i];
        rec_start[i + 1] = o;
    }
    if (o > cap) throw std::invalid_argument("record buffer too small");
    if (o >= (1ull << 32)) throw std::runtime_error("record gather exceeds 2^32 words");
    DBuf<uint16_t> dout(mem, o + 1);
    RIN_HIP(hipMemcpyAsync(ddst.p, dst.data(), 4 * n, hipMemcpyHostToDevice, s));
    const uint64_t threads = n * 64;
    hipLaunchKernelGGL(k_rec_copy, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, n,
                       (const uint64_t*)dsrc.p, (const uint32_t*)dlen.p, (const uint32_t*)ddst.p, vw.rec, vw.k1_rec,
                       dout.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(words, dout.p, 2 * o, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

void compute_edges(DevMem& mem, RunOut& r, hipStream_t s) {
    if (r.have_edges) return;
    const uint64_t NF = r.n_faces, NO = r.n_face_verts, NV = r.n_verts;
    r.e_of_occ.alloc(mem, NO + 1);
    r.e_face_of.alloc(mem, NO + 1);
    r.e_occ.alloc(mem, NO + 1);
    if (NO == 0) {
        r.n_edges = 0;
        r.e_v12.alloc(mem, 2);
        r.e_start.alloc(mem, 1);
        fill_u32(r.e_start.p, 0, 1, s);
        r.have_edges = true;
        return;
    }
    DBuf<uint64_t> key(mem, NO);
    DBuf<uint32_t> sidx(mem, NO), first(mem, NO), head(mem, NO), hexcl(mem, NO), seg_first(mem, NO),
        rank(mem, NO), nt(mem, 2), seg_pos(mem, NO);
    hipLaunchKernelGGL(k_edge_keys, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), r.view.f_vstart, r.view.f_verts,
                       key.p, r.e_face_of.p);
    iota_u32(sidx.p, NO, s);
    sort_pairs(mem, key.p, sidx.p, Count(NO), 32 + bits_for(NV), s);
    hipLaunchKernelGGL(k_heads, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint64_t*)key.p,
                       (const uint32_t*)sidx.p, first.p, head.p);
    scan_u32(mem, head.p, hexcl.p, NO, nullptr, s);
    hipLaunchKernelGGL(k_assign, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, seg_first.p);
    scan_u32(mem, first.p, rank.p, NO, nt.p, s);
    hipLaunchKernelGGL(k_assign2, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, (const uint32_t*)seg_first.p,
                       (const uint32_t*)rank.p, r.e_of_occ.p);
    const uint64_t NE = read1(nt.p, s);
    r.n_edges = NE;
    r.e_v12.alloc(mem, 2 * NE + 2);
    r.e_start.alloc(mem, NE + 1);
    DBuf<uint32_t> cnt(mem, NE + 1);
    hipLaunchKernelGGL(k_edge_segments, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint64_t*)key.p,
                       (const uint32_t*)sidx.p, (const uint32_t*)head.p, (const uint32_t*)hexcl.p,
                       (const uint32_t*)r.e_of_occ.p, seg_pos.p, r.e_v12.p, cnt.p);
    scan_u32(mem, cnt.p, r.e_start.p, NE, r.e_start.p + NE, s);
    hipLaunchKernelGGL(k_edge_fill, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint32_t*)sidx.p,
                       (const uint32_t*)head.p, (const uint32_t*)hexcl.p, (const uint32_t*)seg_pos.p,
                       (const uint32_t*)r.e_of_occ.p, (const uint32_t*)r.e_start.p, r.e_occ.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipStreamSynchronize(s));
    r.have_edges = true;
}

void debug_scan(DevMem& mem, int fields, int stride, const uint32_t* in, uint64_t n, uint64_t count, uint32_t* out,
                uint32_t* totals, hipStream_t s) {
    if (fields != 1 && fields != 3) throw std::invalid_argument("fields must be 1 or 3");
    if (stride < 1 || (fields == 1 && stride != 1)) throw std::invalid_argument("bad stride");
    if (count > n) throw std::invalid_argument("count beyond the capacity");
    const uint64_t in_words = fields == 1 ? n : (stride == 1 ? 3 * n : (uint64_t)stride * n);
    DBuf<uint32_t> din(mem, in_words + 1), dout(mem, (uint64_t)fields * n + 1), dtot(mem, 4), dcnt(mem, 1);
    RIN_HIP(hipMemcpyAsync(din.p, in, 4 * in_words, hipMemcpyHostToDevice, s));
    fill_u32(dout.p, 0xdeadbeefu, (uint64_t)fields * n + 1, s);
    fill_u32(dtot.p, 0xdeadbeefu, 4, s);
    fill_u32(dcnt.p, (uint32_t)count, 1, s);
    // count below the capacity: a device-side count, clamped as the pipeline's are
    const Count c = count == n ? Count(n) : Count(dcnt.p, n);
    if (fields == 1) {
        scan_u32(mem, din.p, dout.p, c, dtot.p, s);
    } else {
        const uint32_t* ins[3];
        uint32_t* outs[3];
        uint32_t* tots[3];
        for (int f = 0; f < 3; ++f) {
            ins[f] = stride == 1 ? din.p + f * n : din.p + f;
            outs[f] = dout.p + f * n;
            tots[f] = dtot.p + f;
        }
        scan3_u32(mem, ins, stride, outs, c, tots, s);
    }
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(out, dout.p, 4 * (uint64_t)fields * n, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipMemcpyAsync(totals, dtot.p, 4 * (uint64_t)fields, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

void compute_patches(DevMem& mem, RunOut& r, hipStream_t s) {
    if (r.have_patches) return;
    compute_edges(mem, r, s);
    const uint64_t NF = r.n_faces, NO = r.n_face_verts;
    r.p_start.alloc(mem, NF + 2);
    r.p_faces.alloc(mem, NF + 1);
    if (NF == 0) {
        r.n_patches = 0;
        fill_u32(r.p_start.p, 0, 1, s);
        RIN_HIP(hipStreamSynchronize(s));
        r.have_patches = true;
        return;
    }
    DBuf<uint32_t> adj(mem, NO + 1), parent(mem, NF), seed(mem, NF + 1), srank(mem, NF + 1), pid(mem, NF),
        level(mem, NF), fa(mem, NF), fb(mem, NF), order(mem, NF), cnt(mem, 2);
    DBuf<unsigned long long> claim(mem, NF);
    DBuf<uint64_t> key(mem, NF);
    hipLaunchKernelGGL(k_face_adj, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint32_t*)r.e_of_occ.p,
                       (const uint32_t*)r.e_start.p, (const uint32_t*)r.e_occ.p, (const uint32_t*)r.e_face_of.p,
                       adj.p);
    iota_u32(parent.p, Count(NF), s);
    hipLaunchKernelGGL(k_uf_hook, dim3(nblk(NO)), dim3(256), 0, s, Count(NO), (const uint32_t*)r.e_face_of.p,
                       (const uint32_t*)adj.p, parent.p);
    hipLaunchKernelGGL(k_uf_roots, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), parent.p, seed.p);
    scan_u32(mem, seed.p, srank.p, Count(NF), srank.p + NF, s);
    hipLaunchKernelGGL(k_patch_seeds, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), (const uint32_t*)parent.p,
                       (const uint32_t*)seed.p, (const uint32_t*)srank.p, pid.p, level.p, fa.p, claim.p);
    const uint64_t NP = read1(srank.p + NF, s);
    r.n_patches = NP;
    copy_u32(order.p, fa.p, Count(NP), s);
    // breadth-first levels over every patch at once
    uint32_t* front = fa.p;
    uint32_t* next = fb.p;
    uint64_t nfront = NP, base = NP;
    uint32_t lv = 1;
    DBuf<uint32_t> sorted(mem, NF);
    while (nfront) {
        fill_u32(cnt.p, (uint32_t)nfront, 1, s);
        fill_u32(cnt.p + 1, 0, 1, s);
        hipLaunchKernelGGL(k_patch_claim, dim3(nblk(nfront)), dim3(256), 0, s, (const uint32_t*)cnt.p,
                           (const uint32_t*)front, (const uint32_t*)r.view.f_vstart, (const uint32_t*)adj.p,
                           (const uint32_t*)level.p, claim.p, next, cnt.p + 1);
        const uint64_t nn = read1(cnt.p + 1, s);
        if (nn == 0) break;
        if (base + nn > NF) throw std::runtime_error("patch search listed a face twice");
        hipLaunchKernelGGL(k_patch_keys, dim3(nblk(nn)), dim3(256), 0, s, (const uint32_t*)(cnt.p + 1),
                           (const uint32_t*)next, (const unsigned long long*)claim.p, key.p, sorted.p);
        sort_pairs(mem, key.p, sorted.p, Count(nn), 16 + bits_for(nfront), s);
        hipLaunchKernelGGL(k_patch_level, dim3(nblk(nn)), dim3(256), 0, s, (const uint32_t*)(cnt.p + 1),
                           (const uint32_t*)sorted.p, lv, level.p, order.p, (uint32_t)base);
        copy_u32(front, sorted.p, Count(nn), s);
        base += nn;
        nfront = nn;
        ++lv;  // front holds the new frontier in order; next takes the next offers
    }
    if (base != NF) throw std::runtime_error("patch search did not reach every face");
    // faces by patch, each patch in its breadth-first order (stable placement)
    DBuf<uint32_t> pc(mem, NP + 1);
    fill_u32(pc.p, 0, NP + 1, s);
    hipLaunchKernelGGL(k_patch_count, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), (const uint32_t*)pid.p, pc.p);
    scan_u32(mem, pc.p, r.p_start.p, Count(NP), r.p_start.p + NP, s);
    hipLaunchKernelGGL(k_patch_keys_pid, dim3(nblk(NF)), dim3(256), 0, s, Count(NF), (const uint32_t*)order.p,
                       (const uint32_t*)pid.p, key.p, r.p_faces.p);
    sort_pairs(mem, key.p, r.p_faces.p, Count(NF), bits_for(NP), s);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipStreamSynchronize(s));
    r.have_patches = true;
}

void compute_chains(DevMem& mem, RunOut& r, hipStream_t s) {
    if (r.have_chains) return;
    compute_edges(mem, r, s);
    const uint64_t NE = r.n_edges, NV = r.n_verts;
    r.c_start.alloc(mem, NE + 2);
    r.c_edges.alloc(mem, NE + 1);
    if (NE == 0) {
        r.n_chains = 0;
        fill_u32(r.c_start.p, 0, 1, s);
        RIN_HIP(hipStreamSynchronize(s));
        r.have_chains = true;
        return;
    }
    DBuf<uint32_t> vdeg(mem, NV + 1), vslot(mem, 2 * NV + 2), nb(mem, 2 * NE), parent(mem, NE), seen(mem, NE),
        seed(mem, NE + 1), srank(mem, NE + 1), size(mem, NE + 1), seeds(mem, NE + 1), err(mem, 1);
    fill_u32(vdeg.p, 0, NV + 1, s);
    fill_u32(err.p, 0, 1, s);
    const uint32_t* v12 = r.e_v12.p;
    const uint32_t* es = r.e_start.p;
    hipLaunchKernelGGL(k_chain_vert, dim3(nblk(NE)), dim3(256), 0, s, Count(NE), v12, es, vdeg.p, vslot.p);
    hipLaunchKernelGGL(k_chain_links, dim3(nblk(NE)), dim3(256), 0, s, Count(NE), v12, es, (const uint32_t*)vdeg.p,
                       (const uint32_t*)vslot.p, nb.p, parent.p, seen.p);
    hipLaunchKernelGGL(k_chain_hook, dim3(nblk(2 * NE)), dim3(256), 0, s, Count(2 * NE), (const uint32_t*)nb.p,
                       parent.p);
    hipLaunchKernelGGL(k_chain_roots, dim3(nblk(NE)), dim3(256), 0, s, Count(NE), es, parent.p, seed.p);
    scan_u32(mem, seed.p, srank.p, Count(NE), srank.p + NE, s);
    const uint64_t NC = read1(srank.p + NE, s);
    r.n_chains = NC;
    if (NC == 0) {
        fill_u32(r.c_start.p, 0, 1, s);
        RIN_HIP(hipStreamSynchronize(s));
        r.have_chains = true;
        return;
    }
    fill_u32(size.p, 0, NC + 1, s);
    hipLaunchKernelGGL(k_chain_count, dim3(nblk(NE)), dim3(256), 0, s, Count(NE), es, (const uint32_t*)parent.p,
                       (const uint32_t*)seed.p, (const uint32_t*)srank.p, size.p, seeds.p);
    scan_u32(mem, size.p, r.c_start.p, Count(NC), r.c_start.p + NC, s);
    hipLaunchKernelGGL(k_chain_walk, dim3(nblk(NC)), dim3(256), 0, s, Count(NC), (const uint32_t*)seeds.p,
                       (const uint32_t*)r.c_start.p, (const uint32_t*)nb.p, seen.p, r.c_edges.p, err.p);
    RIN_HIP(hipGetLastError());
    if (read1(err.p, s) != 0) throw std::runtime_error("chain walk did not cover its component");
    RIN_HIP(hipStreamSynchronize(s));
    r.have_chains = true;
}

namespace {
__global__ void k_debug_sqrt(const double* __restrict__ in, uint64_t n, double* __restrict__ out) {
    GRID_STRIDE(i, n) out[i] = dev::sqrt_rn(in[i]);
}
}  // namespace

void debug_sqrt(DevMem& mem, const double* in, uint64_t n, double* out, hipStream_t s) {
    if (!n) return;
    DBuf<double> din(mem, n), dout(mem, n);
    RIN_HIP(hipMemcpyAsync(din.p, in, 8 * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_debug_sqrt, dim3(nblk(n)), dim3(256), 0, s, (const double*)din.p, n, dout.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(out, dout.p, 8 * n, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

float debug_copy(DevMem& mem, uint64_t bytes, int reps, hipStream_t s) {
    const uint64_t n = bytes / 8;
    if (n == 0 || reps < 1) throw std::invalid_argument("copy size and repetitions must be positive");
    DBuf<uint64_t> a(mem, n), b(mem, n);
    fill_u64(a.p, 0x0123456789abcdefull, n, s);
    copy_u64(b.p, a.p, Count(n), s);  // warm-up
    hipEvent_t e0, e1;
    RIN_HIP(hipEventCreate(&e0));
    RIN_HIP(hipEventCreate(&e1));
    RIN_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) copy_u64(b.p, a.p, Count(n), s);
    RIN_HIP(hipEventRecord(e1, s));
    RIN_HIP(hipEventSynchronize(e1));
    float ms = 0;
    RIN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

int debug_table_record_mi3(const double* rows, uint16_t* out, int cap) {
    const MI3Table& T = mi3_table();
    double P[3][4];
    for (int q = 0; q < 3; ++q)
        for (int c = 0; c < 4; ++c) P[q][c] = rows[4 * q + c];
    int pat = 0;
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) pat += (P[1][c] > P[0][c] ? 2 : (P[1][c] < P[0][c] ? 0 : 1)) * w;
    const MI3Pat& K = T.pat[pat];
    bool err = !K.valid;
    uint32_t slot = K.base, w = 1;
    for (int q = 0; q < K.nv && !err; ++q, w *= 3) {
        const int lab[4] = {K.lab[q][0], K.lab[q][1], K.lab[q][2], K.lab[q][3]};
        slot += (uint32_t)(dev::orient_mi(P, lab, P[2], err) + 1) * w;
    }
    if (err || T.off[slot] == MI3_NONE) return -1;
    const uint16_t* r = T.words.data() + T.off[slot];
    const int len = r[6];
    if (len > cap) return -2;
    std::memcpy(out, r, 2 * (size_t)len);
    return len;
}

int debug_table_record_mi2(const double* rows, uint16_t* out, int cap) {
    const MI2Table& T = mi2_table();
    int pat = 0;
    for (int c = 0, w = 1; c < 4; ++c, w *= 3) {
        const double x = rows[c], y = rows[4 + c];
        pat += (y > x ? 2 : (y < x ? 0 : 1)) * w;
    }
    const MI2Entry& E = T.e[pat];
    if (E.len > cap) return -2;
    std::memcpy(out, T.words.data() + E.off, 2 * (size_t)E.len);
    return E.len;
}

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
    DBuf<uint8_t> dws(mem, mi == 2 ? 1 : LaneCut<CBig>::stride * 64 * nb);
    if (mi == 2)
        hipLaunchKernelGGL((k_debug_cut_wave<CBig>), dim3((unsigned)ncase), dim3(64), 0, s, (const double*)drows.p,
                           (const int*)dks.p, (const int*)dst.p, ncase, cap, dw.p, dlens.p);
    else if (mi)
        hipLaunchKernelGGL((k_debug_cut<CBig, true>), dim3(nb), dim3(64), 0, s, (const double*)drows.p,
                           (const int*)dks.p, (const int*)dst.p, ncase, cap, dw.p, dlens.p, dws.p);
    else
        hipLaunchKernelGGL((k_debug_cut<CBig, false>), dim3(nb), dim3(64), 0, s, (const double*)drows.p,
                           (const int*)dks.p, (const int*)dst.p, ncase, cap, dw.p, dlens.p, dws.p);
    RIN_HIP(hipGetLastError());
    RIN_HIP(hipMemcpyAsync(words, dw.p, 2 * (size_t)ncase * cap, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipMemcpyAsync(lens, dlens.p, 4 * (size_t)ncase, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
}

}  // namespace rin
