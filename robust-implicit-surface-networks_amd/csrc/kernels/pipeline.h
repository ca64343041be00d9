// Device hot path: values -> signs -> active-function CSR -> per-tet cut ->
// iso-vertex / iso-face extraction with first-occurrence dedup -> xyz.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../devmem.h"
#include "eval_dev.h"
#include "grid_dev.h"

namespace rin {

constexpr uint32_t kMaxFuncs = 64;   // function ids are packed in 6 bits
constexpr uint32_t kMaxActive = 10;  // active functions per tet on the device path

struct PlanCache;  // batched form: workspaces and captured graphs per function count (batched.hip)
std::shared_ptr<PlanCache> make_plan_cache();
std::shared_ptr<void> make_mi_plan_cache();  // the same for material interfaces (batched_mi.hip)

struct Grid {
    uint64_t id = 0;  // unique per grid made in this process (values record it)
    GridView view{};
    DBuf<double> pts;
    DBuf<uint32_t> tets, tedge, tface;
    // tet across each tet face (4 t + local face, or all ones on the mesh
    // boundary); built on the first material-interface run
    mutable DBuf<uint32_t> nbr;
    // the batched form's plans live with the grid, so value sets of the same
    // size share workspaces, tuning and captured graphs
    std::shared_ptr<PlanCache> plans;
    std::shared_ptr<void> mi_plans;
};

struct Values {
    uint64_t grid_id = 0;  // the grid the values were made on
    uint32_t F = 0;
    uint64_t V = 0;
    DBuf<double> vals;  // F x V, function-major
};

enum RunFlags : uint32_t {
    RUN_MI = 1,
    RUN_REVERSE = 2,     // insert functions in descending order (cut only)
    RUN_NO_LOOKUP = 4,   // general cut for every tet
    RUN_BATCHED = 8,     // bucketed fused form
    RUN_KEEP_RECORDS = 16,
    RUN_CUT_ONLY = 32,   // stop after the per-tet cut (robustness check)
    RUN_STAGE_TIMES = 64,  // batched form: time every stage (one captured graph per stage)
    RUN_FIXED_FORMS = 128, // batched form: no tuning launches (table for k = 2, wave form beyond)
};

struct StageTimes {
    std::vector<std::pair<std::string, float>> ms;
};

struct RunOut {
    bool mi = false;
    uint32_t flags = 0;
    uint64_t T = 0, F = 0;
    uint64_t num_degenerate = 0, num_intersecting = 0;
    uint64_t nb[3] = {0, 0, 0};
    uint64_t n_active = 0, n_fit = 0;
    uint32_t err = 0;  // bit 0 capacity, bit 1 predicate range, bit 2 record overflow
    DBuf<uint8_t> fit;          // CSR function ids
    DBuf<uint32_t> csr_off;     // T + 1
    DBuf<uint32_t> act;         // tet of each active rank
    DBuf<uint8_t> act_k;
    DBuf<uint32_t> rank_of_tet; // T (exclusive scan of the active flags)
    DBuf<uint32_t> rec_off;     // n_active + 1 (words): each tet's slot and its capacity
    DBuf<uint32_t> rec_loc;     // n_active: where each tet's record is (its slot, or a retry slot)
    DBuf<uint16_t> rec;
    DBuf<uint32_t> cut_counts;  // n_active x 3: (vertices, faces, cells) for the robustness check
    // mesh
    uint64_t n_verts = 0, n_faces = 0, n_face_verts = 0, n_tet_faces = 0;
    uint64_t n_vert_occ = 0, n_face_occ = 0, n_face_occ_verts = 0, n_cut34 = 0, n_cut58 = 0, n_rec_words = 0;
    uint64_t cut_k_sum[2] = {0, 0}, cut_words[2] = {0, 0};  // batched form, cut buckets 3..4 and 5..8
    DBuf<uint32_t> v_tet;
    DBuf<uint16_t> v_j;
    DBuf<uint8_t> v_ss;
    DBuf<uint32_t> v_sv;  // 4 per vertex
    DBuf<uint8_t> v_fn;   // 4 per vertex (0xFF unused)
    DBuf<double> xyz;     // 3 per vertex
    DBuf<uint32_t> f_vstart, f_verts, f_tfstart, f_tf_tet;
    DBuf<uint16_t> f_tf_local;
    DBuf<uint8_t> f_func;  // 2 per face (0xFF unused)
    StageTimes times;

    // read-only views of the outputs (owned by the DBufs above in the plain
    // form, by a batched workspace otherwise)
    struct View {
        const uint8_t* fit = nullptr;
        const uint32_t *csr_off = nullptr, *act = nullptr, *rank_of_tet = nullptr, *rec_off = nullptr;
        const uint16_t* rec = nullptr;
        const uint32_t* rec_loc = nullptr;       // batched: word offset, or 0x80000000|pattern for k=1
        const uint16_t* k1_rec = nullptr;        // batched: k=1 records by pattern
        const uint32_t* k1_rec_off = nullptr;
        const uint32_t *v_tet = nullptr, *v_sv = nullptr;
        const uint16_t* v_j = nullptr;
        const uint8_t *v_ss = nullptr, *v_fn = nullptr;
        const double* xyz = nullptr;
        const uint32_t *f_vstart = nullptr, *f_verts = nullptr, *f_tfstart = nullptr, *f_tf_tet = nullptr;
        const uint16_t* f_tf_local = nullptr;
        const uint8_t* f_func = nullptr;
        const uint32_t* cut_counts = nullptr;
    } view;
    std::shared_ptr<void> hold;  // keeps a batched workspace checked out
    void set_views_from_buffers();
    // iso-edges of the faces (compute_edges), on first request
    bool have_edges = false;
    uint64_t n_edges = 0;
    DBuf<uint32_t> e_v12;       // 2 per edge: the sorted endpoints
    DBuf<uint32_t> e_start;     // n_edges + 1: each edge's occurrences in e_occ
    DBuf<uint32_t> e_occ;       // face-vertex positions (face f, position j <-> f_vstart[f] + j)
    DBuf<uint32_t> e_of_occ;    // edge of each face-vertex position
    DBuf<uint32_t> e_face_of;   // face of each face-vertex position
    // patches (compute_patches), on first request
    bool have_patches = false;
    uint64_t n_patches = 0;
    DBuf<uint32_t> p_start;  // n_patches + 1
    DBuf<uint32_t> p_faces;  // each patch's faces in breadth-first order
    // chains of non-manifold edges (compute_chains), on first request
    bool have_chains = false;
    uint64_t n_chains = 0;
    DBuf<uint32_t> c_start;  // n_chains + 1
    DBuf<uint32_t> c_edges;  // each chain's edges in breadth-first order
};

void grid_regular(DevMem& mem, int res, const double bmin[3], const double bmax[3], Grid& g, hipStream_t s);
void grid_explicit(DevMem& mem, const double* pts, uint64_t V, const uint64_t* tets, uint64_t T, Grid& g,
                   hipStream_t s);
void values_eval(DevMem& mem, const Grid& g, const FuncDesc* funcs, uint32_t F, Values& v, hipStream_t s);
void values_upload(DevMem& mem, const Grid& g, const double* rowmajor, uint32_t F, Values& v, hipStream_t s);
// function values at arbitrary points, row-major n x F, into host memory
void eval_points(DevMem& mem, const double* pts, uint64_t n, const FuncDesc* funcs, uint32_t F, double* rowmajor,
                 hipStream_t s);
void run_pipeline(DevMem& mem, const Grid& g, const Values& v, uint32_t flags, RunOut& out, hipStream_t s);

// thrown by the batched form when a tet needs more than its buckets hold
// (more than 8 active functions, or a complex beyond the bucket capacities);
// the caller runs the plain form, which handles those tets
struct NeedsPlainForm : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// batched form: buckets by active-function count, k=1 lookup table in LDS,
// persistent per-bucket cut kernels, whole path captured in a hipGraph
void run_batched(DevMem& ctx_mem, const Grid& g, const Values& v, uint32_t flags, RunOut& out, hipStream_t s,
                 int n_cu);

// the material-interface lookup tables (2- and 3-material) on the current
// device, uploaded once per process and device; plain pointers (the kernels'
// view of them is MITablesDev in kernels_mi.h)
struct MITablesRaw {
    const void *e2, *w2, *pat3, *off3, *cnt3, *cut3, *w3, *snap_off3, *snap3;
    uint64_t n_w2, n_w3;  // record words of the two- and three-material tables
};
MITablesRaw mi_tables_device(hipStream_t s);

// allocation-free point evaluation (noalloc.hip): values that one call may
// stage, the device address of the staging array on the current device
// (cached per device), and the evaluation itself (caller's device buffers)
constexpr uint64_t kNoallocValues = 2u << 20;
void* noalloc_scratch_address();
void eval_points_noalloc(void* scratch_dev, const double* pts, uint64_t n, const FuncDesc* funcs, uint32_t F,
                         double* rowmajor, hipStream_t s);

// batched form of the material interface (batched_mi.hip): the plain form's
// stages with device-side counts and capacities, captured as one graph per
// workspace and replayed; throws NeedsPlainForm for what it does not handle
void run_batched_mi(DevMem& ctx_mem, const Grid& g, const Values& v, uint32_t flags, RunOut& out, hipStream_t s);

// the tet across every tet face (Grid::nbr), built on first use; synchronises
// the stream, so call it outside graph capture
void ensure_nbr(DevMem& mem, const Grid& g, hipStream_t s);

// copy the cut records of the given tets to the host, packed (rec_start has
// n + 1 word offsets; inactive tets get empty ranges)
void gather_records(DevMem& mem, const RunOut& r, const uint64_t* tets, uint64_t n, uint64_t* rec_start,
                    uint16_t* words, uint64_t cap, hipStream_t s);

// iso-edges of the result's faces, on the device (compute_mesh_edges,
// src/mesh_connectivity.cpp:10-56); fills the e_* buffers of the result
void compute_edges(DevMem& mem, RunOut& r, hipStream_t s);
// the result's patches on the device (src/mesh_connectivity.cpp:58-93), computing the edges first
void compute_patches(DevMem& mem, RunOut& r, hipStream_t s);
// chains of the result's non-manifold iso-edges (computes the edges first)
void compute_chains(DevMem& mem, RunOut& r, hipStream_t s);

// diagnostics: a plain device copy of `bytes` (reference kernel for
// hardware-counter scaling); returns the average milliseconds per copy
float debug_copy(DevMem& mem, uint64_t bytes, int reps, hipStream_t s);
// the function evaluation's correctly rounded square root on host data
void debug_sqrt(DevMem& mem, const double* in, uint64_t n, double* out, hipStream_t s);
// the plain form's k = 1 and k = 2 implicit-arrangement tets (list of active
// ranks, n of them) by the lookup tables: each table record is copied into
// the tet's record slot (batched.hip)
void table_cut_plain(const GridView& g, const double* vals, const uint32_t* act, const uint8_t* act_k,
                     const uint32_t* csr_off, const uint8_t* fit, const uint32_t* list, uint64_t n,
                     const uint32_t* rec_off, uint16_t* rec, uint32_t* cnt, uint32_t* errs, hipStream_t s);
// the k = 2 table's patterns and complexes after two planes on this device
// (the plain form's lane-group cuts start from them; batched.hip)
void plain_k2_snapshots(hipStream_t s, const void** pat, const uint32_t** off, const uint8_t** bytes);
// the device scans (scan_u32; scan3_u32 over three fields) on host data
void debug_scan(DevMem& mem, int fields, int stride, const uint32_t* in, uint64_t n, uint64_t count, uint32_t* out,
                uint32_t* totals, hipStream_t s);

// diagnostics: run the device cut on independent plane/material sets
void debug_cuts(DevMem& mem, int mi, const double* rows, const int* ks, int ncase, int cap, uint16_t* words,
                int* lens, hipStream_t s);

// diagnostics (host only, no device): the record a lookup table gives for one
// tet — IA k = 1 or k = 2 plane rows (k_cube_fill's and k_k2_lookup's
// indexing), or two or three material rows (k_mi_table's); returns the record length
// or a negative code (-1: no table entry, -2: record longer than cap)
int debug_table_record_ia(int k, const double* rows, uint16_t* out, int cap);
int debug_table_record_mi2(const double* rows, uint16_t* out, int cap);
int debug_table_record_mi3(const double* rows, uint16_t* out, int cap);

}  // namespace rin
