// Batched form of the material-interface path (src/material_interface.cpp:53-447).
//
// The plain form's stages (highest material, filter, CSR, cut lists, the
// per-tet cut by table or by the thread cut, boundary-face pairing,
// extraction, radix-sort dedup, coordinates, faces) with every size kept on
// the device: buffers are allocated to capacities, launches are sized by the
// capacities and read the live counts from device memory (clamped, and
// skipped once a count overflowed), scans run over the zero-padded
// capacity.  So the whole path is one captured graph per workspace, replayed
// on every run; an overflow grows the workspace and captures again, and new
// values of the same size re-record the graph and update it in place.  The
// closing kernel publishes the counters into pinned host memory, so a replay
// needs no copies.  Tets the batched cut does not take (more than 8
// materials, or a complex beyond the thread cut's capacity) send the run to
// the plain form.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>

#include "kernels_ia.h"
#include "kernels_mi.h"
#include "pipeline.h"
#include "scan_sort.h"

namespace rin {

namespace {

// device counters of one run
enum MICtr : uint32_t {
    M_DEGEN = 0,   // vertices with tied maxima
    M_ACTIVE,      // active tets
    M_LIST,        // cut list sizes [4], histogram [3], 3..4 / 5..8 counts, largest count (k_bucket_scan)
    M_WORDS = M_LIST + 10,  // record words
    M_NFT,         // faces of the active tets' complexes
    M_VO, M_FO, M_FV,  // vertex occurrences, faces, face corners
    M_NV,          // vertices after dedup
    M_FIT,         // active (tet, material) entries: the CSR's total
    M_ABORT,       // a count exceeded its capacity: later stages process nothing
    M_N
};
// kMiHuge: a tet has more than 8 materials (the large-capacity cut takes it)
constexpr uint32_t kMiOverflow = 1u << 8, kMiPlain = 1u << 9, kMiHuge = 1u << 10;
constexpr unsigned kHugeBlocks = 4;  // lanes of the large-capacity cut: 4 x 64
constexpr int kMiPub = 32;
static_assert(M_N <= kMiPub, "published counters");

struct MICaps {
    uint32_t active = 0, fit = 0, words = 0, nft = 0, vo = 0, fo = 0, fv = 0;
    uint32_t huge = 0;  // the captured path holds the large-capacity cut
};

__global__ void k_mi_zero(uint32_t* ctr, uint32_t* errs) {
    if (threadIdx.x < M_N) ctr[threadIdx.x] = 0;
    if (threadIdx.x < 64) errs[threadIdx.x] = 0;
}

// after each counting stage: a count beyond its capacity stops the later
// stages (and asks the host to grow the workspace); the cut lists' sizes are
// zeroed then, so the cut kernels read nothing out of range
__global__ void k_mi_check(uint32_t* ctr, const uint32_t* csr_off, uint64_t T, MICaps c, uint32_t* errs) {
    if (threadIdx.x != 0) return;
    ctr[M_FIT] = csr_off[T];
    bool over = ctr[M_ACTIVE] > c.active || csr_off[T] > c.fit || ctr[M_WORDS] > c.words ||
                ctr[M_NFT] > c.nft || ctr[M_VO] > c.vo || ctr[M_FO] > c.fo || ctr[M_FV] > c.fv;
    if (over) {
        ctr[M_ABORT] = 1;
        errs[0] |= kMiOverflow;
    }
    // tets beyond 8 materials take the large-capacity cut (a path captured
    // without it stops here and is captured again with it); beyond its
    // capacity the plain form reports them
    if (ctr[M_LIST + 9] > 8) {
        errs[0] |= kMiHuge;
        if (!c.huge) ctr[M_ABORT] = 1;
    }
    if (ctr[M_LIST + 9] > (uint32_t)kMaxHuge) errs[0] |= kMiPlain;
    if (ctr[M_ABORT])
        for (int q = 0; q < 4; ++q) ctr[M_LIST + q] = 0;
}

// the CSR, active list and counts (k_fill with the capacity check)
__global__ void k_mi_fill(uint64_t T, const uint64_t* __restrict__ amask, const uint32_t* __restrict__ csr_off,
                          const uint32_t* __restrict__ rank, const uint32_t* __restrict__ abort,
                          uint8_t* __restrict__ fit, uint32_t* __restrict__ act, uint8_t* __restrict__ act_k) {
    if (*abort) return;
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

// record slot bounds over the whole capacity (zero past the active tets)
__global__ void k_mi_rec_bound(uint64_t cap, const uint32_t* __restrict__ n_active,
                               const uint32_t* __restrict__ abort, const uint8_t* __restrict__ act_k,
                               uint32_t* __restrict__ b) {
    const uint64_t n = *abort ? 0 : *n_active;
    GRID_STRIDE(a, cap + 1) b[a] = a < n ? rec_bound(1, act_k[a]) : 0u;
}

// per active tet its complex's face count; zero past the active tets
__global__ void k_mi_nf_padded(MIArgs A, uint64_t cap, uint32_t* __restrict__ nf) {
    const uint64_t n = count_of(A.n_active);
    GRID_STRIDE(a, cap + 1) {
        uint32_t x = 0;
        if (a < n) {
            const uint16_t* r = A.rec + A.rec_loc[a];
            x = r[0] == 2 ? r[3] : 0u;
        }
        nf[a] = x;
    }
}

// the three per-tet output counts, zero past the active tets
__global__ void k_mi_split_padded(Count n, uint64_t cap, const uint32_t* __restrict__ cnt, uint32_t* a,
                                  uint32_t* b, uint32_t* c) {
    const uint64_t na = count_of(n);
    GRID_STRIDE(i, cap + 1) {
        const bool in = i < na;
        a[i] = in ? cnt[3 * i] : 0u;
        b[i] = in ? cnt[3 * i + 1] : 0u;
        c[i] = in ? cnt[3 * i + 2] : 0u;
    }
}

// closes the face arrays, publishes the counters and the union of the error
// flags into pinned host memory
__global__ void k_mi_close(uint32_t* ctr, uint32_t* errs, uint32_t cap_fo, uint32_t* f_vstart, uint32_t* f_tfstart,
                           uint32_t* pub) {
    const uint32_t nf = ctr[M_FO] < cap_fo ? ctr[M_FO] : cap_fo;
    if (threadIdx.x == 0 && !ctr[M_ABORT]) {
        f_vstart[nf] = ctr[M_FV];
        f_tfstart[nf] = nf;
    }
    if (threadIdx.x < M_N) pub[threadIdx.x] = ctr[threadIdx.x];
    uint32_t e = errs[threadIdx.x];
    for (int d = 32; d > 0; d >>= 1) e |= __shfl_xor(e, d, 64);
    if (threadIdx.x == 0) pub[kMiPub] = e;
    __threadfence_system();
}

unsigned blocks_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65535)); }

int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

}  // namespace

struct MIWorkspace {
    DevMem mem;
    MICaps cap;
    bool in_use = false;
    // per vertex / tet (fixed by the grid)
    DBuf<uint64_t> hm, amask;
    DBuf<uint32_t> kcnt, isact, csr_off, rank, tscan;
    // per active tet
    DBuf<uint8_t> fit, act_k;
    DBuf<uint32_t> act, blists, bc, rbound, rec_off, cnt, fb;
    // each active tet's record location: its slot (rec_off), or for a table
    // tet the table's record, whose words follow the slots in rec (at
    // tab_base2 and tab_base3)
    DBuf<uint32_t> rec_loc;
    uint32_t tab_base2 = 0, tab_base3 = 0;
    DBuf<uint16_t> rec;
    DBuf<uint64_t> emask;
    unsigned cut_blocks[2] = {0, 0};
    DBuf<uint8_t> ws_huge;  // the large-capacity cut's complexes, once a run needs them
    // per complex face, occurrence, face
    DBuf<uint16_t> fl, fw, oj, fq_l;
    DBuf<uint64_t> ft, okey;
    DBuf<uint32_t> cv, cf, cfv, tv, tf, tfv, oa, fq_a, fq_vs, fve;
    DBuf<uint8_t> fq_m;
    // dedup
    DBuf<uint32_t> first, mo, hslot, vrank, vid, first_occ;
    // first-occurrence hash table (key + 1, ~smallest occurrence), zero
    // between runs: the assign pass clears the slots a run used
    DBuf<unsigned long long> hkey;
    DBuf<uint32_t> hmx;
    int hbits = 0;
    bool hash_dirty = false;  // a run stopped with entries left: clear before the next
    // outputs
    DBuf<uint32_t> v_tet, v_sv, f_vstart, f_verts, f_tfstart, f_tf_tet;
    DBuf<uint16_t> v_j, f_tf_local;
    DBuf<uint8_t> v_ss, v_fn, f_func;
    DBuf<double> xyz;
    DBuf<uint32_t> ctr, errs;
    uint32_t* pub_h = nullptr;
    uint32_t* pub_d = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double* bound = nullptr;  // the values the graph reads
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        bound = nullptr;
    }
    ~MIWorkspace() {
        drop_graph();
        if (pub_h) (void)hipHostFree(pub_h);
    }
};

struct MIPlan {
    std::mutex mu;
    std::vector<std::shared_ptr<MIWorkspace>> ws;
};

struct MIPlanCache {
    std::mutex mu;
    std::map<uint32_t, std::shared_ptr<MIPlan>> plans;  // by material count
    std::shared_ptr<MIPlan> get(uint32_t F) {
        std::lock_guard<std::mutex> l(mu);
        auto& p = plans[F];
        if (!p) p = std::make_shared<MIPlan>();
        return p;
    }
};

std::shared_ptr<void> make_mi_plan_cache() { return std::make_shared<MIPlanCache>(); }

namespace {

void alloc_mi(MIWorkspace& W, const Grid& g, const MICaps& c, hipStream_t s) {
    DevMem& m = W.mem;
    W.drop_graph();
    const uint64_t V = g.view.V, T = g.view.T;
    W.cap = c;
    if (!W.hm.p) {
        W.hm.alloc(m, V);
        W.amask.alloc(m, T);
        W.kcnt.alloc(m, T);
        W.isact.alloc(m, T);
        W.tscan.alloc(m, T + 1);
        W.csr_off.alloc(m, T + 1);
        W.rank.alloc(m, T + 1);
        W.ctr.alloc(m, 64);
        W.errs.alloc(m, 64);
    }
    const uint64_t A = c.active;
    W.fit.alloc(m, c.fit + 1);
    W.act_k.alloc(m, A + 1);
    W.act.alloc(m, A + 1);
    W.blists.alloc(m, 4 * (A + 1));
    const uint64_t nch = (A + kBucketChunk - 1) / kBucketChunk + 1;
    W.bc.alloc(m, (uint64_t)kBucketWords * nch + kBucketWords);
    W.rbound.alloc(m, A + 2);
    W.rec_off.alloc(m, A + 2);
    W.rec_loc.alloc(m, A + 2);
    {
        // the slots, then the two tables' record words (copied once per allocation)
        const MITablesRaw TR = mi_tables_device(s);
        if ((uint64_t)c.words + 1 + TR.n_w2 + TR.n_w3 >= (1ull << 32))
            throw std::runtime_error("material-interface record words exceed 32-bit offsets");
        W.tab_base2 = (uint32_t)c.words + 1;
        W.tab_base3 = W.tab_base2 + (uint32_t)TR.n_w2;
        W.rec.alloc(m, (uint64_t)W.tab_base3 + TR.n_w3 + 1);
        RIN_HIP(hipMemcpyAsync(W.rec.p + W.tab_base2, TR.w2, 2 * TR.n_w2, hipMemcpyDeviceToDevice, s));
        RIN_HIP(hipMemcpyAsync(W.rec.p + W.tab_base3, TR.w3, 2 * TR.n_w3, hipMemcpyDeviceToDevice, s));
    }
    W.cnt.alloc(m, 3 * A + 3);
    W.fb.alloc(m, A + 2);
    W.emask.alloc(m, A + 1);
    W.cv.alloc(m, A + 2);
    W.cf.alloc(m, A + 2);
    W.cfv.alloc(m, A + 2);
    W.tv.alloc(m, A + 2);
    W.tf.alloc(m, A + 2);
    W.tfv.alloc(m, A + 2);
    // thread cuts (four and five to eight materials): a lane per tet with
    // its complex in LDS, two lanes (one for the larger complexes) a
    // wavefront, grids striding over their lists
    W.cut_blocks[0] = (unsigned)std::min<uint64_t>((A + 1) / 2 + 1, 4096);
    W.cut_blocks[1] = (unsigned)std::min<uint64_t>(A + 1, 2048);
    W.fl.alloc(m, c.nft + 1);
    W.fw.alloc(m, c.nft + 1);
    W.ft.alloc(m, 3 * (uint64_t)c.nft + 3);
    W.okey.alloc(m, c.vo + 1);
    W.oa.alloc(m, c.vo + 1);
    W.oj.alloc(m, c.vo + 1);
    W.fq_a.alloc(m, c.fo + 1);
    W.fq_l.alloc(m, c.fo + 1);
    W.fq_m.alloc(m, 2 * (uint64_t)c.fo + 2);
    W.fq_vs.alloc(m, c.fo + 1);
    W.fve.alloc(m, c.fv + 1);
    for (auto* b : {&W.first, &W.mo, &W.hslot, &W.vrank, &W.vid, &W.first_occ}) b->alloc(m, c.vo + 1);
    // a table of at least twice the occurrence capacity, cleared once here
    W.hbits = 10;
    while ((1ull << W.hbits) < 2ull * (c.vo + 1)) ++W.hbits;
    W.hkey.alloc(m, 1ull << W.hbits);
    W.hmx.alloc(m, 1ull << W.hbits);
    RIN_HIP(hipMemsetAsync(W.hkey.p, 0, 8ull << W.hbits, s));
    RIN_HIP(hipMemsetAsync(W.hmx.p, 0, 4ull << W.hbits, s));
    W.hash_dirty = false;
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
    if (!W.pub_h) {
        RIN_HIP(hipHostMalloc((void**)&W.pub_h, (kMiPub + 1) * sizeof(uint32_t),
                              hipHostMallocMapped | hipHostMallocCoherent));
        RIN_HIP(hipHostGetDevicePointer((void**)&W.pub_d, W.pub_h, 0));
    }
}

// the whole path on stream s; every launch is sized by capacities and reads
// the live counts from the workspace
void mi_path(MIWorkspace& W, const Grid& g, const Values& val, const MITablesDev& TT, int kbits, hipStream_t s) {
    const GridView& G = g.view;
    const uint64_t V = G.V, T = G.T;
    const uint32_t F = val.F;
    MICaps c = W.cap;
    c.huge = W.ws_huge.p ? 1u : 0u;
    uint32_t* ctr = W.ctr.p;
    const uint32_t* abort = ctr + M_ABORT;
    const double* vals = val.vals.p;
    hipLaunchKernelGGL(k_mi_zero, dim3(1), dim3(64), 0, s, ctr, W.errs.p);
    // highest material and filter (src/material_interface.cpp:53-152)
    hipLaunchKernelGGL(k_highest_mi, dim3(blocks_for(V)), dim3(256), 0, s, vals, V, F, W.hm.p, (uint32_t*)nullptr,
                       ctr + M_DEGEN);
    hipLaunchKernelGGL(k_filter_mi, dim3(blocks_for(T)), dim3(256), 0, s, G, vals, F, (const uint64_t*)W.hm.p,
                       W.amask.p, W.kcnt.p, W.isact.p);
    {
        // CSR offsets and active ranks in one pass (the third field is unused)
        const uint32_t* in[3] = {W.kcnt.p, W.isact.p, W.isact.p};
        uint32_t* out[3] = {W.csr_off.p, W.rank.p, W.tscan.p};
        uint32_t* tot[3] = {W.csr_off.p + T, ctr + M_ACTIVE, nullptr};
        scan3_u32(W.mem, in, 1, out, Count(T), tot, s);
    }
    hipLaunchKernelGGL(k_mi_check, dim3(1), dim3(64), 0, s, ctr, (const uint32_t*)W.csr_off.p, T, c, W.errs.p);
    hipLaunchKernelGGL(k_mi_fill, dim3(blocks_for(T)), dim3(256), 0, s, T, (const uint64_t*)W.amask.p,
                       (const uint32_t*)W.csr_off.p, (const uint32_t*)W.rank.p, abort, W.fit.p, W.act.p,
                       W.act_k.p);
    const Count NA(ctr + M_ACTIVE, c.active, abort);
    // cut lists: the tables take two and three materials, lane groups four
    // to eight.  RIN_CUT_FORM=wave or thread turns the tables off and cuts
    // every tet up to eight materials by lane groups or by a lane each.
    const char* form = std::getenv("RIN_CUT_FORM");
    const bool all_thread = form && !std::strcmp(form, "thread");
    const bool all_wave = form && !std::strcmp(form, "wave");
    const int tables = (all_thread || all_wave) ? 0 : 1;
    uint32_t* bl[4];
    for (int q = 0; q < 4; ++q) bl[q] = W.blists.p + (uint64_t)q * (c.active + 1);
    const uint32_t nch = (uint32_t)((c.active + kBucketChunk - 1) / kBucketChunk) + 1;
    hipLaunchKernelGGL(k_bucket_count, dim3(nch), dim3(256), 0, s, NA, (const uint8_t*)W.act_k.p, 1, tables,
                       W.bc.p);
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, nch, W.bc.p, ctr + M_LIST);
    hipLaunchKernelGGL(k_bucket_fill, dim3(nch), dim3(256), 0, s, NA, (const uint8_t*)W.act_k.p, tables,
                       (const uint32_t*)W.bc.p, bl[0], bl[1], bl[2], bl[3]);
    // record slots
    hipLaunchKernelGGL(k_mi_rec_bound, dim3(blocks_for(c.active + 1)), dim3(256), 0, s, (uint64_t)c.active,
                       (const uint32_t*)(ctr + M_ACTIVE), abort, (const uint8_t*)W.act_k.p, W.rbound.p);
    scan_u32(W.mem, W.rbound.p, W.rec_off.p, Count((uint64_t)c.active + 1), ctr + M_WORDS, s);
    hipLaunchKernelGGL(k_mi_check, dim3(1), dim3(64), 0, s, ctr, (const uint32_t*)W.csr_off.p, T, c, W.errs.p);
    // the cut: tables for two and three materials, thread cuts beyond
    // record locations: each tet's slot, then table tets point at the table's record
    copy_u32(W.rec_loc.p, W.rec_off.p, NA, s);
    hipLaunchKernelGGL(k_mi_table, dim3(blocks_for(c.active)), dim3(256), 0, s, G, vals, (const uint32_t*)W.act.p,
                       (const uint32_t*)W.csr_off.p, (const uint8_t*)W.fit.p, (const uint32_t*)bl[0],
                       (uint64_t)c.active, (const uint32_t*)(ctr + M_LIST), (uint32_t)0, (const uint32_t*)W.rec_off.p,
                       W.rec.p, W.cnt.p, W.errs.p, TT, W.rec_loc.p, W.tab_base2, W.tab_base3);
    const Count n1(ctr + M_LIST + 1, c.active, abort), n2(ctr + M_LIST + 2, c.active, abort);
    if (all_thread) {
        hipLaunchKernelGGL((k_cut_lds<CMed, true, 2>), dim3(W.cut_blocks[0]), dim3(2), 0, s, G, vals,
                           (const uint32_t*)W.act.p, (const uint8_t*)W.act_k.p, (const uint32_t*)W.csr_off.p,
                           (const uint8_t*)W.fit.p, n1, 1, 4, (uint32_t)0, (const uint32_t*)W.rec_off.p, W.rec.p,
                           W.cnt.p, W.errs.p, (const uint32_t*)bl[1]);
        hipLaunchKernelGGL((k_cut_lds<CBig, true, 1>), dim3(W.cut_blocks[1]), dim3(1), 0, s, G, vals,
                           (const uint32_t*)W.act.p, (const uint8_t*)W.act_k.p, (const uint32_t*)W.csr_off.p,
                           (const uint8_t*)W.fit.p, n2, 5, 8, (uint32_t)0, (const uint32_t*)W.rec_off.p, W.rec.p,
                           W.cnt.p, W.errs.p, (const uint32_t*)bl[2]);
    } else {
        // a wavefront per tet: these lists are short and each tet's latency
        // is the kernel's time; the cuts start from the 3-material table's
        // complex after three materials (RIN_MI3_START=0: from the tet)
        const char* ms = std::getenv("RIN_MI3_START");
        const MI3Snap m3s = (ms && !std::strcmp(ms, "0")) ? MI3Snap{nullptr, nullptr, nullptr}
                                                          : MI3Snap{TT.pat3, TT.snap_off3, TT.snap3};
        const unsigned g1 = (unsigned)std::min<uint64_t>(c.active + 1, 2048);
        const unsigned g2 = (unsigned)std::min<uint64_t>(c.active + 1, 1024);
        hipLaunchKernelGGL((k_cut_group<CMed, 64, true>), dim3(g1), dim3(64), 0, s, G, vals,
                           (const uint32_t*)W.act.p, (const uint8_t*)W.act_k.p, (const uint32_t*)W.csr_off.p,
                           (const uint8_t*)W.fit.p, (const uint32_t*)bl[1], n1, (const uint32_t*)W.rec_off.p, W.rec.p,
                           W.cnt.p, W.errs.p, 0, m3s);
        hipLaunchKernelGGL((k_cut_group<CBig, 64, true>), dim3(g2), dim3(64), 0, s, G, vals,
                           (const uint32_t*)W.act.p, (const uint8_t*)W.act_k.p, (const uint32_t*)W.csr_off.p,
                           (const uint8_t*)W.fit.p, (const uint32_t*)bl[2], n2, (const uint32_t*)W.rec_off.p, W.rec.p,
                           W.cnt.p, W.errs.p, 0, m3s);
    }
    if (W.ws_huge.p)
        hipLaunchKernelGGL(k_cut_huge<true>, dim3(kHugeBlocks), dim3(64), 0, s, G, vals, (const uint32_t*)W.act.p,
                           (const uint8_t*)W.act_k.p, (const uint32_t*)W.csr_off.p, (const uint8_t*)W.fit.p,
                           (const uint32_t*)bl[3], Count(ctr + M_LIST + 3, c.active, abort), (uint32_t)0,
                           (const uint32_t*)W.rec_off.p, W.rec.p, W.cnt.p, W.errs.p, W.ws_huge.p,
                           (const uint32_t*)nullptr, (uint32_t)0);
    // extraction (src/extract_mesh.cpp:569-986): boundary-face pairing, then
    // occurrences in the reference's creation order
    MIArgs A;
    A.g = G;
    A.act = W.act.p;
    A.csr_off = W.csr_off.p;
    A.rank = W.rank.p;
    A.rec_loc = W.rec_loc.p;
    A.nbr = g.nbr.p;
    A.fit = W.fit.p;
    A.rec = W.rec.p;
    A.n_active = NA;
    hipLaunchKernelGGL(k_mi_nf_padded, dim3(blocks_for(c.active + 1)), dim3(256), 0, s, A, (uint64_t)c.active, W.fb.p);
    scan_u32(W.mem, W.fb.p, W.fb.p, Count((uint64_t)c.active + 1), ctr + M_NFT, s);
    hipLaunchKernelGGL(k_mi_check, dim3(1), dim3(64), 0, s, ctr, (const uint32_t*)W.csr_off.p, T, c, W.errs.p);
    MIFaceKeys K;
    K.fb = W.fb.p;
    K.fl = W.fl.p;
    K.fw = W.fw.p;
    K.ft = W.ft.p;
    const unsigned nbl = blocks_for((uint64_t)c.active * kBfaceLanes);
    hipLaunchKernelGGL(k_mi_face_keys, dim3(nbl), dim3(256), 0, s, A, K);
    hipLaunchKernelGGL(k_mi_bface, dim3(nbl), dim3(256), 0, s, A, K, W.emask.p);
    hipLaunchKernelGGL(k_mi_count, dim3(blocks_for(c.active)), dim3(256), 0, s, A, (const uint64_t*)W.emask.p,
                       W.cnt.p);
    hipLaunchKernelGGL(k_mi_split_padded, dim3(blocks_for(c.active + 1)), dim3(256), 0, s, NA, (uint64_t)c.active,
                       (const uint32_t*)W.cnt.p, W.tv.p, W.tf.p, W.tfv.p);
    {
        const uint32_t* in[3] = {W.tv.p, W.tf.p, W.tfv.p};
        uint32_t* out[3] = {W.cv.p, W.cf.p, W.cfv.p};
        uint32_t* tot[3] = {ctr + M_VO, ctr + M_FO, ctr + M_FV};
        scan3_u32(W.mem, in, 1, out, Count((uint64_t)c.active + 1), tot, s);
    }
    hipLaunchKernelGGL(k_mi_check, dim3(1), dim3(64), 0, s, ctr, (const uint32_t*)W.csr_off.p, T, c, W.errs.p);
    hipLaunchKernelGGL(k_mi_extract, dim3(blocks_for(c.active)), dim3(256), 0, s, A, (const uint64_t*)W.emask.p,
                       (const uint32_t*)W.cv.p, (const uint32_t*)W.cf.p, (const uint32_t*)W.cfv.p, W.okey.p, W.oa.p,
                       W.oj.p, W.fq_a.p, W.fq_l.p, W.fq_m.p, W.fq_vs.p, W.fve.p);
    // vertex dedup: each key's smallest occurrence by an atomic hash table,
    // ids by a scan of the first-occurrence flags (the table is left empty)
    const Count NVO(ctr + M_VO, c.vo, abort);
    (void)kbits;
    hipLaunchKernelGGL(k_hash_insert, dim3(blocks_for(c.vo)), dim3(256), 0, s, NVO, (const uint32_t*)nullptr,
                       (const uint64_t*)W.okey.p, W.hkey.p, W.hmx.p, W.hbits, W.hslot.p, W.errs.p);
    hipLaunchKernelGGL(k_hash_first, dim3(blocks_for(c.vo)), dim3(256), 0, s, NVO, (const uint32_t*)nullptr,
                       (const uint32_t*)W.hslot.p, (const uint32_t*)W.hmx.p, W.first.p, W.mo.p);
    scan_u32(W.mem, W.first.p, W.vrank.p, NVO, ctr + M_NV, s);
    hipLaunchKernelGGL(k_dedup_assign, dim3(blocks_for(c.vo)), dim3(256), 0, s, NVO, (const uint32_t*)W.mo.p,
                       (const uint32_t*)W.first.p, (const uint32_t*)W.vrank.p, W.vid.p, W.first_occ.p, NVO,
                       (const uint32_t*)nullptr, (const uint32_t*)W.hslot.p, W.hkey.p, W.hmx.p, true);
    // coordinates and faces
    const Count NV(ctr + M_NV, c.vo, abort), NF(ctr + M_FO, c.fo, abort), NFV(ctr + M_FV, c.fv, abort);
    hipLaunchKernelGGL(k_vert_fields_mi, dim3(blocks_for(c.vo)), dim3(256), 0, s, G, vals, (const uint32_t*)W.act.p,
                       (const uint32_t*)W.csr_off.p, (const uint8_t*)W.fit.p, (const uint32_t*)W.rec_loc.p,
                       (const uint16_t*)W.rec.p, (const uint32_t*)W.first_occ.p, (const uint32_t*)W.oa.p,
                       (const uint16_t*)W.oj.p, NV, W.v_tet.p, W.v_j.p, W.v_ss.p, W.v_sv.p, W.v_fn.p, W.xyz.p);
    hipLaunchKernelGGL(k_mi_face_out, dim3(blocks_for(c.fo)), dim3(256), 0, s, NF, (const uint32_t*)W.act.p,
                       (const uint32_t*)W.fq_a.p, (const uint16_t*)W.fq_l.p, (const uint8_t*)W.fq_m.p,
                       (const uint32_t*)W.fq_vs.p, (const uint32_t*)W.fve.p, (const uint32_t*)W.vid.p, NFV,
                       W.f_vstart.p, W.f_verts.p, W.f_tfstart.p, W.f_tf_tet.p, W.f_tf_local.p, W.f_func.p);
    hipLaunchKernelGGL(k_mi_close, dim3(1), dim3(64), 0, s, ctr, W.errs.p, c.fo, W.f_vstart.p, W.f_tfstart.p,
                       W.pub_d);
    RIN_HIP(hipGetLastError());
}

MICaps grow_caps(const MICaps& have, const uint32_t* pub, uint64_t T) {
    // a count past its capacity grows that capacity by a quarter over the
    // count; a stage that never ran (an earlier overflow) doubles its capacity
    auto g = [](uint32_t have_c, uint32_t seen) -> uint32_t {
        const uint64_t want = seen > have_c ? (uint64_t)seen + seen / 4 + 64 : (uint64_t)have_c;
        return (uint32_t)std::min<uint64_t>(want, 0xFFFFFFF0u);
    };
    MICaps c;
    const bool ran = !pub[M_ABORT];
    c.active = g(have.active, pub[M_ACTIVE]);
    c.fit = g(have.fit, pub[M_FIT]);
    c.words = g(have.words, pub[M_WORDS]);
    c.nft = g(have.nft, pub[M_NFT]);
    c.vo = g(have.vo, pub[M_VO]);
    c.fo = g(have.fo, pub[M_FO]);
    c.fv = g(have.fv, pub[M_FV]);
    if (!ran) {
        // later stages did not count: make room in proportion to the tets
        c.fit = std::max<uint32_t>(c.fit, 3 * c.active);
        c.words = std::max<uint32_t>(c.words, 160 * c.active);
        c.nft = std::max<uint32_t>(c.nft, 12 * c.active);
        c.vo = std::max<uint32_t>(c.vo, 8 * c.active);
        c.fo = std::max<uint32_t>(c.fo, 4 * c.active);
        c.fv = std::max<uint32_t>(c.fv, 16 * c.active);
    }
    (void)T;
    return c;
}

}  // namespace

void run_batched_mi(DevMem& ctx_mem, const Grid& g, const Values& val, uint32_t flags, RunOut& out, hipStream_t s) {
    (void)ctx_mem;
    if (flags & (RUN_CUT_ONLY | RUN_REVERSE | RUN_NO_LOOKUP))
        throw NeedsPlainForm("the -R check and lookups off run the plain form");
    if (val.grid_id != g.id) throw std::invalid_argument("values were made on another grid");
    ensure_nbr(ctx_mem, g, s);
    auto* cache = static_cast<MIPlanCache*>(g.mi_plans.get());
    std::shared_ptr<MIPlan> planp = cache->get(val.F);
    MIPlan& P = *planp;
    std::lock_guard<std::mutex> lock(P.mu);
    std::shared_ptr<MIWorkspace> W;
    for (auto& w : P.ws)
        if (!w->in_use) {
            W = w;
            break;
        }
    if (!W) {
        W = std::make_shared<MIWorkspace>();
        P.ws.push_back(W);
    }
    MIWorkspace& w = *W;
    const MITablesDev TT = mi_tables_view(mi_tables_device(s));
    const uint64_t T = g.view.T;
    uint32_t pub[kMiPub + 1];
    bool ok = false;
    for (int attempt = 0; attempt < 8 && !ok; ++attempt) {
        if (!w.ctr.p) {
            MICaps c0;
            c0.active = (uint32_t)std::min<uint64_t>(T, 1024 + T / 16);
            c0.fit = 3 * c0.active;
            c0.words = 160 * c0.active;
            c0.nft = 12 * c0.active;
            c0.vo = 8 * c0.active;
            c0.fo = 4 * c0.active;
            c0.fv = 16 * c0.active;
            alloc_mi(w, g, c0, s);
        }
        if (w.bound != val.vals.p) {
            // (re-)record the path for these values; update the graph in place
            const int kbits = 2 + std::max({bits_of(g.view.V), bits_of(g.view.n_edge_ids) + 12,
                                            bits_of(g.view.n_face_ids) + 18, bits_of(w.cap.active) + 10});
            if (kbits > 63) throw NeedsPlainForm("simplex keys exceed 63 bits for this mesh");
            hipGraph_t gr = nullptr;
            RIN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            try {
                mi_path(w, g, val, TT, kbits, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            RIN_HIP(hipStreamEndCapture(s, &gr));
            bool updated = false;
            if (w.exec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res;
                updated = hipGraphExecUpdate(w.exec, gr, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
                if (!updated) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(w.exec);
                    w.exec = nullptr;
                }
            }
            if (!updated) RIN_HIP(hipGraphInstantiate(&w.exec, gr, nullptr, nullptr, 0));
            if (w.graph) (void)hipGraphDestroy(w.graph);
            w.graph = gr;
            w.bound = val.vals.p;
        }
        hipEvent_t e0, e1;
        RIN_HIP(hipEventCreate(&e0));
        RIN_HIP(hipEventCreate(&e1));
        if (w.hash_dirty) {
            RIN_HIP(hipMemsetAsync(w.hkey.p, 0, 8ull << w.hbits, s));
            RIN_HIP(hipMemsetAsync(w.hmx.p, 0, 4ull << w.hbits, s));
            w.hash_dirty = false;
        }
        RIN_HIP(hipEventRecord(e0, s));
        RIN_HIP(hipGraphLaunch(w.exec, s));
        RIN_HIP(hipEventRecord(e1, s));
        RIN_HIP(hipStreamSynchronize(s));
        float ms = 0;
        RIN_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::memcpy(pub, w.pub_h, sizeof(pub));
        uint32_t e = pub[kMiPub];
        // a stopped or failed run may leave table entries
        if ((e & ~kMiHuge) || ((e & kMiHuge) && !w.ws_huge.p)) w.hash_dirty = true;
        if (e & kMiPlain) throw NeedsPlainForm("a tet has more materials than the device cut holds");
        if ((e & kMiHuge) && !w.ws_huge.p) {
            // first run with such tets: room for the large-capacity cut, and
            // the path captured again with its launch
            w.ws_huge.alloc(w.mem, LaneCut<CHuge>::stride * 64 * kHugeBlocks);
            w.drop_graph();
            continue;
        }
        e &= ~kMiHuge;
        if (e & kMiOverflow) {
            alloc_mi(w, g, grow_caps(w.cap, pub, T), s);
            continue;
        }
        if (e) {
            out.err = e;
            throw NeedsPlainForm("a complex exceeds the batched cut's capacities or the predicate range");
        }
        out.times.ms.push_back({"batched path (one graph)", ms});
        ok = true;
    }
    if (!ok) throw std::runtime_error("batched material-interface workspace did not converge");
    w.in_use = true;
    out.mi = true;
    out.flags = flags;
    out.T = T;
    out.F = val.F;
    out.num_degenerate = pub[M_DEGEN];
    out.n_active = out.num_intersecting = pub[M_ACTIVE];
    uint32_t fit_total = 0;
    RIN_HIP(hipMemcpyAsync(&fit_total, w.csr_off.p + T, 4, hipMemcpyDeviceToHost, s));
    RIN_HIP(hipStreamSynchronize(s));
    out.n_fit = fit_total;
    for (int q = 0; q < 3; ++q) out.nb[q] = pub[M_LIST + 4 + q];
    out.n_cut34 = pub[M_LIST + 7];
    out.n_cut58 = pub[M_LIST + 8];
    out.n_verts = pub[M_NV];
    out.n_faces = pub[M_FO];
    out.n_face_verts = pub[M_FV];
    out.n_tet_faces = pub[M_FO];
    out.n_vert_occ = pub[M_VO];
    out.n_face_occ = pub[M_FO];
    out.n_face_occ_verts = pub[M_FV];
    auto& vw = out.view;
    vw.fit = w.fit.p;
    vw.csr_off = w.csr_off.p;
    vw.act = w.act.p;
    vw.rank_of_tet = w.rank.p;
    vw.rec_off = w.rec_off.p;
    vw.rec = w.rec.p;
    vw.rec_loc = w.rec_loc.p;
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
    std::weak_ptr<MIPlan> wp = planp;
    MIWorkspace* raw = W.get();
    out.hold = std::shared_ptr<void>(nullptr, [W, wp, raw](void*) {
        if (auto p = wp.lock()) {
            std::lock_guard<std::mutex> l(p->mu);
            raw->in_use = false;
        }
    });
}

}  // namespace rin
