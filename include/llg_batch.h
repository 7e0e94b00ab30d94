/*
 * llg_batch.h - many independent spin systems advanced together on one GPU, and the
 * continuous-batching request loop built on them.
 *
 * No reference counterpart: the reference integrates one system per process
 * (src/cuda.cu:278-319).  A member of a batch is an FP32 handle (llg_desc with
 * precision LLG_FP32, scheme LLG_HEUN, counter noise); its couplings, constants,
 * temperature, field, seed and step counter are its own.  One kernel launch per half
 * step covers the tiles of every member, and with LLG_BATCH_GRAPH a chunk of steps is
 * captured once into a hipGraph and replayed.  Members are admitted against an HBM
 * budget and can join or leave between chunks; a member's trajectory does not depend on
 * which other members share its launches (counter noise keyed by its own seed/stream,
 * no cross-member reductions).
 */
#ifndef LLG_BATCH_H
#define LLG_BATCH_H
#include <stdint.h>

#include "llg_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LLG_BATCH_GRAPH 0x1u      /* replay captured hipGraphs of whole chunks */
#define LLG_BATCH_PERSISTENT 0x2u /* advance a chunk in persistent launches (llg_batch_set_persistent) */
#define LLG_BATCH_SPIN_DEFAULT 0x1000000u /* polls per grid-barrier wait of a persistent launch */

typedef struct llg_batch llg_batch;

typedef struct llg_batch_info {
    uint32_t members;      /* live members */
    uint32_t tiles;        /* 256-spin tiles per half-step launch */
    uint64_t spins;        /* total spins */
    uint64_t device_bytes; /* bytes held by live members */
    uint64_t budget_bytes;
    uint32_t graphs_built; /* hipGraph instantiations so far */
    uint32_t persist_grid; /* blocks of the last persistent launch (0: none yet) */
} llg_batch_info;

/* Upper bound of the device bytes llg_create would allocate for d. */
uint64_t llg_estimate_bytes(const llg_desc* d);

int llg_batch_create(int device, uint64_t hbm_budget_bytes, uint32_t flags, llg_batch** out);
void llg_batch_destroy(llg_batch* b);

/* Admit a system with initial spins (SoA, n doubles each) at step t0.  Returns
 * LLG_E_NOMEM without allocating when the estimate exceeds the remaining budget. */
int llg_batch_add(llg_batch* b, const llg_desc* d, const double* sx, const double* sy, const double* sz,
                  double T, const double B[3], uint64_t t0, int* member);
/* Admit a member that shares the couplings of member `proto_member` (same n; d->exch
 * ignored); only its state counts against the budget. */
int llg_batch_add_like(llg_batch* b, const llg_desc* d, int proto_member, const double* sx, const double* sy,
                       const double* sz, double T, const double B[3], uint64_t t0, int* member);
int llg_batch_remove(llg_batch* b, int member);
int llg_batch_set_temperature(llg_batch* b, int member, double T);
int llg_batch_set_applied(llg_batch* b, int member, const double B[3]);

/* Advance every member nsteps steps (each from its own step counter).  Members created with
 * LLG_FLAG_CHECK_FINITE are checked for NaN/Inf spins after the steps; a member that fails is
 * taken out of later launches (its spins and step counter stay as they were found, its status
 * becomes LLG_E_NONFINITE) and the call returns LLG_E_NONFINITE naming it.  The other members
 * have advanced normally and go on unaffected: no member reads another's data.  This replaces
 * the reference's error exit (src/error.cpp:271-281) with a per-member status. */
int llg_batch_step(llg_batch* b, uint32_t nsteps);
/* LLG_OK, or LLG_E_NONFINITE for a member taken out by a failed finite check. */
int llg_batch_member_status(llg_batch* b, int member, int* status);

/* Launch shape of the batched kernels: tpb threads per tile (one block per tile: 128, 256
 * or 512), the XCD-aware block order (1) or plain order (0), for members with a line
 * encoding the line reader (1) or their packed entries (0), and gather_copy: neighbours
 * (and a row's own spin) read from AoS3 copies of S and Y that the kernels keep current (1)
 * or from the SoA arrays (0; packed and line readers only).  Tile size, block order and the
 * gather copy do not change results (each thread owns one spin of one member, same sums);
 * the reader changes the order of the sums to FP32 rounding. */
int llg_batch_set_launch(llg_batch* b, uint32_t tpb, int xcd, int line_mode, int gather_copy);

/* Persistent chunks (on = 1; LLG_BATCH_PERSISTENT at creation sets it too): llg_batch_step
 * advances each group of members through all steps of a chunk in ONE cooperative launch of
 * as many blocks as fit on the GPU at once; each block takes a fixed set of tiles, and a grid
 * barrier (agent-scope atomics) separates the half steps.  Dipolar refreshes split the chunk:
 * a refresh runs between two persistent launches, before the predictor of its step; a batch
 * with a member that refreshes at the predicted spins (every stage) uses per-half-step
 * launches.  Results are those of the per-half-step launches, bit for bit.  With
 * LLG_BATCH_GRAPH as well, each chunk (barrier resets, the persistent launches and the
 * refreshes between them) is captured once into a hipGraph and replayed; the residency probe
 * runs before the capture, and a runtime that refuses to capture a cooperative launch gets
 * direct launches instead.  Every barrier wait ends after spin_limit polls of about a
 * microsecond each (LLG_BATCH_SPIN_DEFAULT, the value LLG_BATCH_PERSISTENT sets, is 2^24; 0
 * abandons the launch at the first barrier that is not complete on arrival, which tests the
 * failure path): a wait that runs out abandons the launch, llg_batch_step returns
 * LLG_E_DEVICE, the members' spins are left partly advanced and their step counters
 * unchanged.  No reference counterpart (the reference's step launches a kernel
 * per field term and per half step, src/cuda.cu:43-240). */
int llg_batch_set_persistent(llg_batch* b, int on, uint32_t spin_limit);

/* The batch's own autotuner (no reference counterpart; the reference takes
 * threadsperblock from its config, src/cuda.cu:322-324): times nsteps steps of the live
 * members for every tile size, block order and gather copy (with hipGraph chunks of nsteps steps when
 * the batch replays graphs) and, when a member has a line encoding, for both readers;
 * keeps the fastest, and restores every member's spins and step counter. */
typedef struct llg_batch_tune {
    int tpb, xcd;
    double ms_per_step;          /* chosen */
    double ms_per_step_default;  /* 256 threads, XCD order, line reader */
    int candidates;
    int line_mode;
    int gather_copy;
} llg_batch_tune;
int llg_batch_autotune(llg_batch* b, uint32_t nsteps, llg_batch_tune* out);

/* FP64 sublattice magnetisation of every member listed, m[k*nsub_max*3 ...]: each member
 * writes n_sublattice*3 values at offset k*stride. */
int llg_batch_reduce(llg_batch* b, const int* members, uint32_t count, uint32_t stride, double* m);
int llg_batch_get_spins(llg_batch* b, int member, double* sx, double* sy, double* sz);
int llg_batch_get_step(llg_batch* b, int member, uint64_t* t);
int llg_batch_get_info(llg_batch* b, llg_batch_info* info);

/* ---------------------------------------------------------------- request loop */

/* Request types: the semantics of the reference drivers (src/mvt.cpp, src/ramp_field.cpp,
 * src/timeseries.cpp):
 *   LLG_REQ_MT    fixed T and B: run `steps`, sample <M> every `update` steps;
 *   LLG_REQ_RAMP  field ramp B(t) = B0 + (B1 - B0) t/steps, B updated only on sampling
 *                 steps (src/ramp_field.cpp:209-224), T fixed;
 *   LLG_REQ_SW    spin-wave time series at fixed T and B0: every `update` steps, <M> and the
 *                 transverse structure factor S(k) (quantisation axis z, llg_set_sq with
 *                 ua = ub = 1) at k = (q, 0, 0) mesh points, q = 0 .. L/2, as the
 *                 time-series driver samples (src/timeseries.cpp:256-378). */
enum { LLG_REQ_MT = 0, LLG_REQ_RAMP = 1, LLG_REQ_SW = 2 };

typedef struct llg_request {
    uint32_t type;
    uint32_t L;        /* simple-cubic L x L x L, periodic */
    uint32_t steps;
    uint32_t update;   /* sampling interval */
    double T;
    double J;          /* nearest-neighbour exchange (J) */
    double K;          /* uniaxial anisotropy along z (J) */
    double mu;         /* moment (mu_B) */
    double lambda;
    double dt;
    double B0[3], B1[3];
    double arrival_s;  /* arrival time relative to the start of the run */
    uint64_t seed;
} llg_request;

typedef struct llg_result {
    double arrival_s, admit_s, done_s, latency_s;
    double m_final[3];  /* <M> at the last sample */
    double m_mean[3];   /* mean of the samples of the second half of the run */
    uint32_t samples;
    uint32_t pad_;
} llg_result;

typedef struct llg_serve_config {
    int device;
    uint32_t flags;           /* LLG_BATCH_GRAPH, LLG_BATCH_PERSISTENT */
    uint64_t hbm_budget_bytes;
    uint32_t max_members;     /* 0: unlimited */
    uint32_t chunk;           /* steps per batched launch chunk; 0: the smallest update interval */
    int sequential;           /* 1: no batching - one request at a time in arrival order */
    int precision;            /* LLG_FP32 (batched) or LLG_FP64_PLAIN (sequential only) */
    /* optional: called on the serving thread after every progress_every-th completion and
     * after the last one, with the completions so far, nreq and the seconds since the start
     * (NULL or progress_every 0: never) */
    void (*progress)(uint32_t completed, uint32_t nreq, double wall_s, void* user);
    void* progress_user;
    uint32_t progress_every;
    uint32_t max_live_spins;  /* 0: unlimited; else a request waits (FIFO) while admitting it would take
                                 the live requests' spins above this, unless none is live */
} llg_serve_config;

typedef struct llg_serve_stats {
    double wall_s;
    double solves_per_s;
    double p50_latency_s, p99_latency_s, mean_latency_s;
    double spin_updates_per_s;
    uint32_t completed, max_live;
    uint64_t spin_updates;
} llg_serve_stats;

/* Run the request stream in real time: requests are admitted when they have arrived
 * and memory is free, advance in shared launches, and leave when done.  results[i]
 * belongs to reqs[i]; m_series (may be NULL) receives for request i its samples
 * (3 doubles each) at series_offset[i] (computed by the caller as the prefix sum of
 * ceil(steps/update)).  sq_series (may be NULL) receives for spin-wave request i, sample s,
 * the L/2 + 1 complex S(k) values (re, im) at doubles sq_offset[i] + s (L/2 + 1) 2
 * (the caller's prefix sum of ceil(steps/update) (L/2 + 1) 2 over spin-wave requests). */
int llg_serve_run(const llg_serve_config* cfg, const llg_request* reqs, uint32_t nreq, llg_result* results,
                  double* m_series, const uint64_t* series_offset, double* sq_series, const uint64_t* sq_offset,
                  llg_serve_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
