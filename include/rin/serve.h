/*
 * rin/serve.h — serving loop over the per-tet hot path (no counterpart in the
 * reference, whose drivers run one arrangement per process).
 *
 * A server owns one GPU and n_slots worker slots; each slot has its own
 * context (HIP stream and device-memory pool) and keeps the background grids
 * it has built resident, keyed by resolution.  Jobs (an implicit arrangement
 * or a material interface over a regular grid on [-1,1]^3 with the given
 * functions) queue in arrival order and start as soon as a slot is free and
 * their estimated device memory fits the server's HBM budget
 * (rin_server_admits), so jobs that arrive while others run are admitted as
 * slots and memory free up, and their work overlaps on the device (one
 * stream per slot).  Implicit arrangements run the batched form, material
 * interfaces the plain form.  Each finished job keeps its mesh (the
 * rin_result_verts / rin_result_faces arrays of include/rin.h) in host memory
 * until released.  Times are seconds since the server was created.
 */
#ifndef RIN_SERVE_H
#define RIN_SERVE_H
#include <stdint.h>

#include "../rin.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rin_server rin_server;

typedef struct rin_job_desc {
    int32_t mode;    /* 0 implicit arrangement, 1 material interface */
    int32_t res;     /* grid resolution: (res+1)^3 vertices, 5 res^3 tets */
    uint32_t n_funcs;
    uint32_t flags;  /* extra rin_run flags (e.g. RIN_FLAG_NO_LOOKUP) */
    const rin_func_desc* funcs;  /* copied at submission */
} rin_job_desc;

typedef struct rin_job_info {
    uint64_t id;
    int32_t status;  /* RIN_OK, a negative rin_status, or 1 while queued / running */
    int32_t slot;
    double t_submit, t_start, t_end;
    uint64_t hbm_estimate;
    rin_sizes sizes;
} rin_job_info;

typedef struct rin_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t max_in_flight;    /* most jobs running at once */
    uint64_t tets;             /* tets of completed jobs */
    double first_submit, last_end;
    double p50_latency, p99_latency, max_latency;  /* t_end - t_submit over completed jobs */
    uint64_t hbm_budget;
} rin_server_stats;

/* hbm_budget = 0: 80% of the device memory free at creation. */
int rin_server_create(int32_t device, int32_t n_slots, uint64_t hbm_budget, rin_server** out);
int rin_server_destroy(rin_server* server);
const char* rin_server_last_error(const rin_server* server);
int rin_server_submit(rin_server* server, const rin_job_desc* job, uint64_t* job_id);
/* Blocks until the job has finished (or timeout_s elapses: returns RIN_ERR_INTERNAL). */
int rin_server_wait(rin_server* server, uint64_t job_id, double timeout_s, rin_job_info* info);
/* Mesh of a finished job, with the layout of rin_result_verts / rin_result_faces. */
int rin_server_job_verts(rin_server* server, uint64_t job_id, uint64_t* rec11, double* xyz);
int rin_server_job_faces(rin_server* server, uint64_t job_id, uint64_t* vert_start, uint64_t* verts,
                         uint64_t* tf_start, uint64_t* tf, uint64_t* func);
int rin_server_job_release(rin_server* server, uint64_t job_id);
int rin_server_stats_get(rin_server* server, rin_server_stats* stats);
/* seconds since the server was created (the clock of rin_job_info) */
double rin_server_now(const rin_server* server);
/* The admission rule: the head of the queue starts when nothing runs, or when
 * the memory in use plus its estimate fits the budget (1 = admit). */
int rin_server_admits(uint64_t in_flight, uint64_t in_use, uint64_t estimate, uint64_t budget);

#ifdef __cplusplus
}
#endif
#endif /* RIN_SERVE_H */
