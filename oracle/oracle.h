/*
 * oracle.h - FP64 CPU oracle of the stochastic-LLG time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or the timed CPU
 * baseline; the product (libsllg.so) never links or calls it.
 *
 * The oracle restates, single-threaded and in FP64, the reference algorithm:
 *   llgCPU::llgCPU            src/llgCPU.cpp:56-361 (loop structure, field terms)
 *   cuint::CHeun1/CHeun2      src/cuint.cu:60-242 (GPU Heun semantics = parity target)
 *   cufields::CSpMV_*         src/cufields.cu:97-246 (CSR/DIA order, val[a][b]*S_b)
 *   RK4 / Cash-Karp tableau   src/cuintRK4.cu:436-456, src/cuintRK5.cu:44-54 (stage formation
 *                             corrected to the textbook methods)
 *   fields::bfdip             src/fields.cpp:199-251 (brute-force dipolar sum)
 *   util::calc_mag method 0   src/util_mag.cpp:44-63
 *   exch::exchunitcell        src/exch_unitcell.cpp:265-491 (CSR construction)
 *   Random::MWC1616/normal    src/random.cpp:17-64, spins::setSpinsRandom src/spins.cpp:265-281
 * Each function below cites the lines it follows.
 */
#ifndef SLLG_ORACLE_H
#define SLLG_ORACLE_H
#include <stdint.h>
#include "../include/llg_capi.h"
#include "../include/llg_lattice.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ctx or_ctx;

/* Oracle options beyond the descriptor. */
#define OR_COMPAT_LLGCPU 0x1u /* reproduce llgCPU quirks A1 (corrector exchange accumulates onto
                                 the predictor field, src/llgCPU.cpp:253-274) and A2 (z noise uses
                                 the x draw, src/llgCPU.cpp:180,302) */

int or_create(const llg_desc* d, uint32_t options, or_ctx** out);
void or_destroy(or_ctx* c);
int or_set_spins(or_ctx* c, const double* sx, const double* sy, const double* sz);
int or_get_spins(or_ctx* c, double* sx, double* sy, double* sz);
int or_set_temperature(or_ctx* c, double T);
int or_set_applied(or_ctx* c, const double B[3]);
int or_set_staggered(or_ctx* c, const double* hx, const double* hy, const double* hz);
int or_scale_staggered(or_ctx* c, double r);
int or_push_noise(or_ctx* c, uint32_t nsteps, const double* xi);
int or_step(or_ctx* c, uint64_t t0, uint32_t nsteps);
int or_get_det_field(or_ctx* c, double* hx, double* hy, double* hz);
int or_get_field(or_ctx* c, double* hx, double* hy, double* hz);
int or_get_dipolar(or_ctx* c, double* hx, double* hy, double* hz); /* brute force at current spins */
int or_reduce(or_ctx* c, double* m, double* energy);
int or_noise_sample(or_ctx* c, uint64_t t, double* xi);
int or_run_sampled(or_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t every, double* m_series);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), one block. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The three N(0,1) values the FP64 path uses for (seed, stream, spin, step). */
void or_counter_normals_f64(uint64_t seed, uint64_t stream, uint32_t spin, uint64_t t, double z[3]);

/* MWC1616 + Marsaglia polar normal (src/random.cpp:17-64). */
typedef struct or_mwc { uint32_t x, y; int have; double n2; } or_mwc;
void or_mwc_seed(or_mwc* r, uint32_t a, uint32_t b);
uint32_t or_mwc_next(or_mwc* r);
double or_mwc_uniform(or_mwc* r);
double or_mwc_normal(or_mwc* r);
/* setSpinsRandom (src/spins.cpp:265-281) driven by MWC1616 seeded (a, b). */
void or_random_spins(uint32_t a, uint32_t b, uint32_t n, double* sx, double* sy, double* sz);

/* exchunitcell restatement (src/exch_unitcell.cpp:265-491): CSR of the lattice in the
 * reference spin order.  Two-pass: with adjncy == NULL only *nnz is returned. */
int or_build_csr(const llg_unitcell* uc, const uint32_t dims[3], const uint32_t pbc[3],
                 uint32_t n_int, const llg_interaction* ints, const double* mu_atom, int ncomp,
                 uint32_t* xadj, uint32_t* adjncy, double* vals, uint64_t* nnz);

#ifdef __cplusplus
}
#endif
#endif
