/*
 * rin.h — C ABI of the MI355X per-tet implicit-arrangement / material-interface
 * hot path.
 *
 * The reference runs this path inside implicit_arrangement() and
 * material_interface() (src/implicit_arrangement.h:39-62,
 * src/material_interface.h:38-61): function signs, the active-function filter,
 * one simplicial_arrangement cut per active tet, extract_iso_mesh /
 * extract_MI_mesh and the vertex coordinates.  This library performs those
 * stages on one GPU and hands back plain arrays; the drop-in C++ entry points
 * (include/rin/implicit_arrangement.h) call it and run the host graph stages.
 *
 * Conventions: every call returns RIN_OK (0) or a negative status; no C++
 * exception crosses the ABI; rin_last_error() describes the last failure of a
 * context.  The library owns device memory until the matching *_release; the
 * caller owns every host buffer.  One context per host thread; contexts are
 * independent (no process-global state).  Results are deterministic and do not
 * depend on scheduling.
 */
#ifndef RIN_H
#define RIN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rin_ctx rin_ctx;
typedef struct rin_grid rin_grid;
typedef struct rin_values rin_values;
typedef struct rin_result rin_result;

enum rin_status {
    RIN_OK = 0,
    RIN_ERR_ARG = -1,         /* bad argument */
    RIN_ERR_DEVICE = -2,      /* HIP runtime error */
    RIN_ERR_MEMORY = -3,      /* device allocation failed */
    RIN_ERR_CAPACITY = -4,    /* input exceeds a device capacity (active functions, keys) */
    RIN_ERR_UNSUPPORTED = -5, /* feature not available on this path */
    RIN_ERR_INTERNAL = -6
};

/* Function descriptor (the reference reads these from a function-file JSON,
 * app/implicit_arrangement.cpp:59).  type: 0 plane (p0..2 point, p3..5 normal),
 * 1 sphere (p0..2 center, p3 radius, p4 squared flag), 2 cylinder (p0..2 axis
 * point, p3..5 unit axis, p6 radius, p7 flipped), 3 torus (p0..2 center, p3..5
 * unit axis, p6 major radius, p7 minor radius, p8 flipped), 4 zero. */
typedef struct rin_func_desc {
    int32_t type;
    int32_t reserved;
    double params[12];
} rin_func_desc;

/* rin_run flags */
#define RIN_FLAG_MI 1u           /* material interface instead of implicit arrangement */
#define RIN_FLAG_REVERSE 2u      /* insert active functions in descending order (with CUT_ONLY) */
#define RIN_FLAG_NO_LOOKUP 4u    /* general cut for every tet (no k=1/k=2 or 2-material tables;
                                    with BATCHED the run takes the plain form) */
#define RIN_FLAG_BATCHED 8u      /* bucketed fused form */
#define RIN_FLAG_KEEP_RECORDS 16u /* keep per-tet cut records for rin_result_records */
#define RIN_FLAG_CUT_ONLY 32u    /* stop after the per-tet cut (the -R robustness check) */
#define RIN_FLAG_STAGE_TIMES 64u /* with BATCHED: per-stage device times (one graph per stage; slower) */
#define RIN_FLAG_FIXED_FORMS 128u /* with BATCHED: no tuning launches on a new plan; the k=2 table and the
                                     wave-per-tet cuts (for one-off jobs, e.g. serving) */

typedef struct rin_sizes {
    uint64_t n_tets, n_funcs;
    uint64_t num_degenerate;   /* IA: zero (vertex,function) entries; MI: vertices with tied maxima */
    uint64_t num_intersecting; /* active tets */
    uint64_t bucket[3];        /* IA: 1, 2, >=3 active functions; MI: 2, 3, >=4 */
    uint64_t n_func_in_tet;    /* CSR length */
    uint64_t n_verts, n_faces, n_face_verts, n_tet_faces;
    /* work sizes of the extraction (for throughput accounting) */
    uint64_t n_vert_occ;      /* vertex occurrences before dedup */
    uint64_t n_face_occ;      /* face occurrences */
    uint64_t n_face_occ_verts;
    uint64_t n_cut[2];        /* active tets with 3..4 and 5..8 functions */
    uint64_t n_rec_words;     /* cut-record words read by the extraction (batched form; 0 otherwise) */
    /* batched form, cut buckets 3..4 and 5..8 (0 otherwise): active functions
     * summed over the bucket's tets, and the words its cut wrote (records and
     * their face offsets) */
    uint64_t cut_k_sum[2];
    uint64_t cut_words[2];
} rin_sizes;

/* Context: one HIP stream and a device-memory cache on `device`.
 * Replaces the process-global lookup-table state of the reference
 * (app/implicit_arrangement.cpp:29-42). */
int rin_create(rin_ctx** out, int32_t device);
/* Free and total device memory of the context's GPU, in bytes. */
int rin_device_memory(rin_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
int rin_destroy(rin_ctx* ctx);
const char* rin_last_error(const rin_ctx* ctx);

/* Background mesh.  rin_grid_regular restates generate_tet_mesh
 * (src/io.cpp:95-152): (res+1)^3 vertices, 5*res^3 tets, tets implicit.
 * rin_grid_explicit takes the arrays load_tet_mesh produces
 * (src/io.cpp:65-93): V x 3 points, T x 4 corner indices. */
int rin_grid_regular(rin_ctx* ctx, int32_t res, const double bmin[3], const double bmax[3], rin_grid** out);
int rin_grid_explicit(rin_ctx* ctx, const double* pts, uint64_t n_pts, const uint64_t* tets, uint64_t n_tets,
                      rin_grid** out);
int rin_grid_points(rin_ctx* ctx, const rin_grid* grid, double* pts_out);
int rin_grid_release(rin_ctx* ctx, rin_grid* grid);

/* Function values at the grid vertices, kept on the device.
 * rin_values_eval evaluates descriptors on the GPU (replaces load_functions,
 * app/implicit_arrangement.cpp:59); rin_values_upload takes the reference's
 * row-major V x F funcVals (src/implicit_arrangement.h:47). */
int rin_values_eval(rin_ctx* ctx, const rin_grid* grid, const rin_func_desc* funcs, uint32_t n_funcs,
                    rin_values** out);
int rin_values_upload(rin_ctx* ctx, const rin_grid* grid, const double* vals_rowmajor, uint32_t n_funcs,
                      rin_values** out);
int rin_values_download(rin_ctx* ctx, const rin_values* vals, double* vals_rowmajor);
/* Function values at arbitrary points, evaluated on the GPU into host memory:
 * n x 3 points in, n x n_funcs row-major values out.  This is the evaluation
 * inside load_functions (external implicit_functions, called at
 * app/implicit_arrangement.cpp:59 and tests/test_implicit_networks.cpp:62). */
int rin_eval_points(rin_ctx* ctx, const double* pts, uint64_t n, const rin_func_desc* funcs, uint32_t n_funcs,
                    double* vals_rowmajor);
/* No counterpart in the reference: rin_eval_points for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops).
 * pts (n x 3), funcs (n_funcs descriptors) and vals_rowmajor (n x n_funcs)
 * are DEVICE pointers on the context's GPU.  The call allocates nothing: its
 * one intermediate (the function-major values) lives in a fixed device array
 * of 2^21 doubles (16 MiB) per GPU, held until the process ends, so
 * n * n_funcs may not exceed 2097152 (RIN_ERR_ARG beyond).  Runs on the
 * context's stream and returns when the values are written.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other threads,
 * other contexts) race. */
int rin_eval_points_noalloc(rin_ctx* ctx, const double* pts, uint64_t n, const rin_func_desc* funcs,
                            uint32_t n_funcs, double* vals_rowmajor);
int rin_values_release(rin_ctx* ctx, rin_values* vals);

/* The hot path: src/implicit_arrangement.cpp:53-402 (IA) or
 * src/material_interface.cpp:53-447 (MI, with RIN_FLAG_MI). */
int rin_run(rin_ctx* ctx, const rin_grid* grid, const rin_values* vals, uint32_t flags, rin_result** out);

int rin_result_sizes(rin_ctx* ctx, const rin_result* res, rin_sizes* sizes);
/* CSR of active functions: func_in_tet[n_func_in_tet], start_index[n_tets+1]
 * (src/implicit_arrangement.cpp:82-116). */
int rin_result_csr(rin_ctx* ctx, const rin_result* res, uint64_t* func_in_tet, uint64_t* start_index);
/* Vertices in first-occurrence order: 11 words each — tet_index,
 * tet_vert_index, simplex_size, simplex_vert_indices[4],
 * func_indices[4] (MI: material_indices) with UINT64_MAX for unused slots —
 * and n_verts x 3 coordinates (src/mesh.h:28-57, src/extract_mesh.cpp:1446-1637). */
int rin_result_verts(rin_ctx* ctx, const rin_result* res, uint64_t* rec11, double* xyz);
/* Faces in first-occurrence order: vert_start[n_faces+1], verts[n_face_verts],
 * tf_start[n_faces+1], tf[2*n_tet_faces] (tet, local face) pairs,
 * func[2*n_faces] (func_index; UINT64_MAX when unused) (src/mesh.h:15-25). */
int rin_result_faces(rin_ctx* ctx, const rin_result* res, uint64_t* vert_start, uint64_t* verts, uint64_t* tf_start,
                     uint64_t* tf, uint64_t* func);
/* Iso-edges of the result's faces, computed on the device on first request
 * (compute_mesh_edges, src/mesh_connectivity.cpp:10-56): edges in order of
 * first occurrence over (face, j), where face edge j joins face vertices j
 * and j + 1 (cyclically).  rin_result_edges fills v12[2 * n_edges] (each
 * edge's vertices, smaller first), start[n_edges + 1] and occ[2 * n_face_verts]
 * (edge e's (face, j) pairs are occ[2 * start[e] .. 2 * start[e + 1]), in
 * (face, j) order: face_edge_indices), and edge_of[n_face_verts] (the edge
 * of face f's edge j at vert_start[f] + j: edges_of_face). */
int rin_result_edge_count(rin_ctx* ctx, const rin_result* res, uint64_t* n_edges);
int rin_result_edges(rin_ctx* ctx, const rin_result* res, uint64_t* v12, uint64_t* start, uint64_t* occ,
                     uint64_t* edge_of);
/* Patches of the result's faces, computed on the device on first request
 * (compute_patches, src/mesh_connectivity.cpp:58-93): components of the faces
 * joined across manifold edges (edges with exactly two face occurrences),
 * numbered by their smallest face, each listed in breadth-first order from
 * that face (neighbours in edge order).  rin_result_patches fills
 * start[n_patches + 1] and faces[n_faces] (patch p is faces[start[p] ..
 * start[p + 1])). */
int rin_result_patch_count(rin_ctx* ctx, const rin_result* res, uint64_t* n_patches);
int rin_result_patches(rin_ctx* ctx, const rin_result* res, uint64_t* start, uint64_t* faces);
/* Chains of the result's non-manifold iso-edges, computed on the device on
 * first request (compute_chains, src/mesh_connectivity.cpp:196-241): edges
 * with more than two face occurrences, joined across vertices that carry
 * exactly two such edges; chains numbered by their smallest edge, each listed
 * in breadth-first order from that edge (the neighbour across its first
 * vertex, then across its second).  rin_result_chains fills
 * start[n_chains + 1] and edges[start[n_chains]] (chain c is edges[start[c] ..
 * start[c + 1]), edge ids as rin_result_edges numbers them). */
int rin_result_chain_count(rin_ctx* ctx, const rin_result* res, uint64_t* n_chains);
int rin_result_chains(rin_ctx* ctx, const rin_result* res, uint64_t* start, uint64_t* edges);
/* Per-tet cut records (layout in include/rin/cut_record.h) of the given tets,
 * packed: rec_start[n+1] word offsets into words[cap_words].  Needs
 * RIN_FLAG_KEEP_RECORDS.  Empty tets get an empty record. */
int rin_result_records(rin_ctx* ctx, const rin_result* res, const uint64_t* tets, uint64_t n, uint64_t* rec_start,
                       uint16_t* words, uint64_t cap_words);
/* With RIN_FLAG_CUT_ONLY: (vertices, faces, cells) of every active tet's
 * complex, in tet order, n_active x 3 (the -R comparison,
 * src/implicit_arrangement.cpp:136-243). */
int rin_result_cut_counts(rin_ctx* ctx, const rin_result* res, uint64_t* counts);
/* Device stage times in milliseconds; labels follow the reference's timing
 * labels where a stage has one.  Returns the number of stages. */
int rin_result_stage_times(rin_ctx* ctx, const rin_result* res, double* ms, int32_t cap);
const char* rin_result_stage_label(rin_ctx* ctx, const rin_result* res, int32_t i);
int rin_result_release(rin_ctx* ctx, rin_result* res);

/* Test hook with no counterpart in the reference: run the device per-tet cut
 * on independent plane (IA) or material (MI) sets.  mode: 0 IA thread-per-tet
 * cut, 1 MI thread-per-tet cut, 2 IA wave-cooperative cut.  rows holds
 * sum(ks) rows of 4 corner values; each case writes at most cap words at
 * words + i*cap and its length (or a negative error) to lens[i]. */
int rin_debug_cuts(rin_ctx* ctx, int32_t mi, const double* rows, const int32_t* ks, int32_t n_cases, int32_t cap,
                   uint16_t* words, int32_t* lens);

/* Test hook with no counterpart in the reference, host only (no context, no
 * device): the per-tet record the lookup tables give for independent tets.
 * mode 0: IA one-function table (1 plane row per case), 1: IA two-function
 * table (2 rows), 2: MI two-material table (2 material rows), 3: MI
 * three-material table (3 material rows).  Each case
 * writes at most cap words at words + i*cap and its length to lens[i]
 * (-1: the table has no entry for it, -2: longer than cap). */
int rin_debug_table_cuts(int32_t mode, const double* rows, int32_t n_cases, int32_t cap, uint16_t* words,
                         int32_t* lens);

/* Test hook with no counterpart in the reference: `reps` plain device copies
 * of `bytes` (a reference kernel of known traffic for hardware-counter
 * scaling); ms receives the average duration of one copy. */
int rin_debug_copy(rin_ctx* ctx, uint64_t bytes, int32_t reps, double* ms);

/* Test hook with no counterpart in the reference: the device's exclusive
 * scans on host data.  fields = 1: out[i] = sum(in[0..i)); fields = 3: three
 * scans in one pass, field f of element i at in[f * n + i] (stride 1) or
 * in[i * stride + f], out[f * n + i].  count < n passes the element count
 * through device memory with capacity n (as the pipeline's stages do);
 * totals[f] receives each field's sum. */
int rin_debug_scan(rin_ctx* ctx, int32_t fields, int32_t stride, const uint32_t* in, uint64_t n, uint64_t count,
                   uint32_t* out, uint32_t* totals);
/* Test hook with no counterpart in the reference: the correctly rounded
 * square root the device's function evaluation uses, on host data (n
 * doubles in, n out), so that its exact fallback near rounding midpoints can
 * be checked against the host's. */
int rin_debug_sqrt(rin_ctx* ctx, const double* in, uint64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* RIN_H */
