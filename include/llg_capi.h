/*
 * llg_capi.h - C ABI of the MI355X stochastic Landau-Lifshitz-Gilbert time step.
 *
 * One opaque handle (llg_ctx) owns the device state of one spin system: spins in
 * structure-of-arrays layout, exchange couplings, per-species constants, the
 * thermal-noise source, the optional dipolar convolution, and a HIP stream.
 * Every entry point returns LLG_OK (0) or a negative llg_status; nothing calls
 * exit().  llg_last_error() returns a message for the last failure on the
 * calling thread.  A handle is not thread-safe.  All host buffers belong to the
 * caller; the library copies what it needs and keeps no host pointer.
 *
 * Reference interfaces replaced (tomostler13/gpu-atomistic-code):
 *   llg_create          <- llg::initLLG (src/llg.cpp:33-638), cullg::cuinit (src/cuda.cu:261-402),
 *                          allocate_memory_on_card (src/cuda_util.cu:253-562),
 *                          cuint::copyConstData (src/cuint.cu:44-59), llgCPU::initLLG (src/llgCPU.cpp:32-54)
 *   llg_step            <- llg::integrate(unsigned int&) (inc/llg.h:17, src/llg.cpp:639-664), called nsteps times
 *   llg_set_spins /     <- spins::Sx/Sy/Sz (inc/spins.h:21) and the D2H copy every spins::update steps
 *   llg_get_spins          (src/cuda.cu:234-246)
 *   llg_get_det_field   <- fields::Hx/Hy/Hz = CDetFields (src/cuint.cu:235, src/cuda.cu:247-253)
 *   llg_set_temperature <- global llg::T (inc/llg.h:11)
 *   llg_set_applied     <- global llg::applied[3] (inc/llg.h:11)
 *   llg_set_staggered   <- fields::setStagField / cullg::CsetStagFields / CsetStagFieldsZero
 *                          (src/fields.cpp:162-198, src/cuda_util.cu:563-588)
 *   llg_scale_staggered <- cullg::rampStag (src/cuda.cu:38-41), fields::incStagFields
 *   llg_reduce          <- util::calc_mag method 0 (src/util_mag.cpp:44-63); energy has no reference counterpart
 *   llg_set_bins /      <- util::calc_mag methods 1-10 (src/util_mag.cpp:65-245) as device sums
 *   llg_reduce_bins
 *   llg_set_sq /        <- the time-series S(k, t) sampler (src/timeseries.cpp:262-380): FFTW of the
 *   llg_sample_sq          transverse spin grid read at chosen k, here a direct sum at those k
 */
#ifndef LLG_CAPI_H
#define LLG_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LLG_ABI_VERSION 1u
#define LLG_MAX_SPECIES 64u

/* Physical constants used by the reference (src/llg.cpp:25). */
#define LLG_GAMMA_E 1.76e11 /* rad s^-1 T^-1 */
#define LLG_MU_B 9.27e-24   /* J/T */
#define LLG_K_B 1.38e-23    /* J/K */

typedef enum llg_status {
    LLG_OK = 0,
    LLG_E_INVALID = -1,     /* malformed descriptor or argument */
    LLG_E_NOMEM = -2,       /* device allocation failed, or over the HBM budget */
    LLG_E_DEVICE = -3,      /* HIP runtime error */
    LLG_E_UNSUPPORTED = -4, /* valid request that this library does not implement */
    LLG_E_NODEVICE = -5,    /* no GPU visible */
    LLG_E_STATE = -6,       /* call out of order (e.g. no injected noise left) */
    LLG_E_NONFINITE = -7    /* NaN/Inf detected in spins */
} llg_status;

/* Integration scheme: reference llg::intscheme (src/llg.cpp:309-341). */
typedef enum llg_scheme { LLG_HEUN = 0, LLG_RK4 = 1, LLG_RK5 = 2 } llg_scheme;

/* Arithmetic of the time step.
 * LLG_FP64_PLAIN: FP64 everywhere, reference structure (separate SpMV and integrator
 *   passes, per-nnz FP64 couplings, per-spin parameter arrays); the correctness path.
 * LLG_FP32: FP32 fused field assembly + integrator, |S|=1 renormalised each step,
 *   FP64 reductions and FP64 dipolar-kernel setup; the performance path. */
typedef enum llg_precision { LLG_FP64_PLAIN = 0, LLG_FP32 = 1 } llg_precision;

/* Exchange storage; numbering follows the reference config::exchm (src/config.cpp:139-159). */
typedef enum llg_exch_format {
    LLG_EXCH_NONE = 0,
    LLG_EXCH_DIA = 1,    /* Bell-Garland DIA, bounds-checked, no wrap (src/cufields.cu:97-178) */
    LLG_EXCH_CSR = 2,    /* CSR (src/cufields.cu:180-246) */
    LLG_EXCH_STENCIL = 3 /* periodic stencil: per-site offsets with wrap arithmetic (no reference counterpart) */
} llg_exch_format;

/* Source of the thermal field xi ~ N(0,1) per component per step. */
typedef enum llg_noise {
    LLG_NOISE_COUNTER = 0, /* Philox4x32-10 keyed by (seed, stream), counter (spin, step); Box-Muller */
    LLG_NOISE_INJECTED = 1 /* caller-provided arrays, consumed one per step (llg_push_noise) */
} llg_noise;

/* Flags (llg_desc.flags). */
#define LLG_FLAG_CHECK_FINITE 0x1u /* llg_step checks spins for NaN/Inf after the run */
/* Encoding switches of the FP32 path, for measurement; results do not depend on them. */
#define LLG_FLAG_NO_XCD_REMAP 0x100u        /* tiles in launch order, no per-XCD row ranges */
#define LLG_FLAG_NO_SUBLATTICE_ORDER 0x200u /* keep the reference spin order on the device */
#define LLG_FLAG_NO_SLICE_SHARING 0x400u    /* store every 64-row slice's entries, even repeats (no line mode) */
#define LLG_FLAG_NO_LINE_MODE 0x800u        /* no line encoding: per-lane gathers only */
/* FP32 dipolar convolution by hipFFT 3D transforms of the whole padded grid instead of 2D
 * transforms of its x < Gx/2 half plus the fused x-transform / k-space product kernel
 * (same result to FP32 rounding). */
#define LLG_FLAG_DIPOLAR_FULL_FFT 0x1000u
/* llg_autotune keeps the reader set by llg_set_launch (line reader or packed entries) instead
 * of timing both: the two sum a row's couplings in different orders, so with this flag the
 * FP32 trajectory does not depend on which reader a timing run happened to favour. */
#define LLG_FLAG_TUNE_FIXED_READER 0x2000u

/* Per-species (unit-cell atom) constants; reference geom::ucm (inc/unitcell.h) and
 * the per-spin arrays filled in src/geom.cpp:239-247 and src/llg.cpp:591-592,626-631. */
typedef struct llg_species {
    double mu;         /* moment in Bohr magnetons */
    double lambda;     /* Gilbert damping */
    double gamma_rel;  /* gyromagnetic ratio / gamma_e; enters sigma only (src/llg.cpp:591-592) */
    double k1u;        /* first-order uniaxial constant K (J); field = 2K/(mu mu_B)(S.e)e */
    double k1u_dir[3]; /* unit axis e */
    uint32_t sublattice; /* magnetisation group of this species (geom::sublattice) */
    uint32_t pad_;
} llg_species;

/* Global anisotropy terms (reference anis::k2par.. , src/anis.cpp:82-165, applied in
 * src/cuint.cu:96-137 with prefactor 2K/(mu_0 mu_B), mu_0 = moment of species 0).
 * kc is a true cubic term with no reference counterpart: E = -(kc/2) sum_a S_a^4 per spin,
 * field (2 kc/(mu mu_B)) (Sx^3, Sy^3, Sz^3). */
typedef struct llg_anisotropy {
    double k2perp, k2par, k4perp, k4par; /* J */
    double k2perp_dir[3];
    double k2par_dir[3];
    double k4perp_dirs[9]; /* [3 axes][3] */
    double k4par_dirs[9];
    double kc; /* J */
} llg_anisotropy;

/* Exchange couplings in tesla, J_ij/(mu_i mu_B) as the reference stores them
 * (src/exch_unitcell.cpp:399).  ncomp = 1 (isotropic: one value used for xx, yy, zz),
 * 3 (xx, yy, zz) or 9 (row-major xx xy xz yx yy yz zx zy zz; H_a += sum_b J_ab S_b).
 * Values are component-major, like the reference exch::dataxx.. arrays. */
typedef struct llg_exchange {
    int32_t format; /* llg_exch_format */
    int32_t ncomp;  /* 1, 3 or 9 */
    /* CSR: row i sums entries xadj[i]..xadj[i+1]-1 in stored order. */
    uint64_t nnz;
    const uint32_t* xadj;   /* [n+1] */
    const uint32_t* adjncy; /* [nnz] */
    const double* csr_vals; /* [ncomp][nnz] */
    /* DIA: col = row + offset[d]; skipped when outside [0, n). */
    uint32_t ndiag;
    const int32_t* dia_offsets; /* [ndiag] */
    const double* dia_vals;     /* [ncomp][ndiag][n] */
    /* Periodic stencil on the spin order idx = ((x*dims1 + y)*dims2 + z)*n_atoms + t
     * (src/geom.cpp:214-266).  Basis atom t has stencil_count[t] entries
     * (dx, dy, dz, t'); neighbours across a non-periodic face are skipped.
     * With LLG_EXCH_CSR, n_atoms > 1 (<= 16) declares that spins follow that order
     * (basis atom fastest); the FP32 path then stores spins sublattice-major on the
     * device.  Results are returned in the caller's order either way. */
    uint32_t dims[3];
    uint32_t n_atoms;
    uint32_t pbc[3];
    uint32_t stencil_k;              /* max entries per site */
    const uint32_t* stencil_count;   /* [n_atoms] */
    const int32_t* stencil_entries;  /* [n_atoms][stencil_k][4] */
    const double* stencil_vals;      /* per_row ? [ncomp][stencil_k][n] : [ncomp][n_atoms][stencil_k] */
    uint32_t stencil_vals_per_row;
    uint32_t pad_;
} llg_exchange;

/* Dipole-dipole field by zero-padded FFT convolution (src/intmat.cpp:263-331,
 * src/cuda.cu:61-90).  The padded grid is (2*dim[a]*nk[a]) per axis; the kernel is
 * built in FP64 at create time and kept as 6 real k-space components. */
typedef struct llg_dipolar {
    int32_t enabled;
    int32_t every_stage; /* 0: refresh every `refresh` steps before the predictor only
                            (reference semantics, src/cuda.cu:65); 1: before every field evaluation */
    uint32_t refresh;    /* refresh interval in steps; 0 means desc.update */
    uint32_t pad_;
    uint32_t dim[3];     /* unit cells per axis */
    uint32_t nk[3];      /* mesh points per unit cell (geom::Nk) */
    double abc[3];       /* lattice constants in metres */
    const uint32_t* mesh; /* [n][3] mesh coordinates lu(i,0..2) */
} llg_dipolar;

/* Four-spin exchange, reference cufields::CSpMV_CSR_FourSpin (src/cufields.cu:59-90; host form
 * src/llgCPU.cpp:142-173).  Spin i has quartets (j, k, l) at xadj[i]..xadj[i+1]-1 and, after the
 * pair exchange, receives
 *   H_i -= (jq (S_j (S_k.S_l) + S_k (S_j.S_l) + S_l (S_k.S_j))) / 3
 * with jq = JQ/(mu_0 mu_B) in tesla, one constant as in the GPU kernel (species 0's value,
 * src/cufields.cu:38-40).  cpu_form = 1 drops the /3 as the reference CPU path does
 * (src/llgCPU.cpp:163-165).  nq = 0 turns the term off.  Energy: -1/4 mu mu_B S.H4 per spin. */
typedef struct llg_fourspin {
    uint64_t nq;
    const uint32_t* xadj; /* [n+1] */
    const uint32_t* jkl;  /* [nq][3] */
    double jq;
    int32_t cpu_form;
    uint32_t pad_;
} llg_fourspin;

/* FFT exchange, the reference's Exchange_method "fft" (exchm = 0): the exchange (and, with
 * dip.enabled, the dipole-dipole interaction) as one species-resolved 3x3 convolution on the
 * zero-padded mesh of desc.dip (dim, nk, abc, mesh; open boundaries), applied before every
 * field evaluation (src/exch_intmat.cpp:28-160 fills N(s1, s2, a, b, r) with
 * J_ab / (mu_s1 mu_B); src/intmat.cpp:188-262 adds 1e-7 (3 e_a e_b - delta_ab) mu_s2 mu_B / r^3;
 * src/cuda.cu:47-60,162-175 and src/cufields.cu:248-292,361-408,432-446 apply it per stage).
 * Interaction q: a spin i of species s1 at mesh m receives
 *   H_a(i) += sum_b J[q][3 a + b] S_b(j) / (mu_s1 mu_B)
 * from the spin j of species s2 at mesh m + d[q] when there is one (J in joules; the caller
 * lists every vector, e.g. all symmetry images of a shell).  exch.format must be
 * LLG_EXCH_NONE; n_int = 0 turns it off. */
typedef struct llg_fft_interaction {
    uint32_t s1, s2;
    int32_t d[3];
    uint32_t pad_;
    double J[9];
} llg_fft_interaction;

typedef struct llg_fftexch {
    uint32_t n_int;
    uint32_t pad_;
    const llg_fft_interaction* ints;
} llg_fftexch;

typedef struct llg_desc {
    uint32_t abi_version; /* LLG_ABI_VERSION */
    uint32_t n;           /* number of spins */
    uint32_t n_species;   /* 1..LLG_MAX_SPECIES */
    uint32_t n_sublattice;
    const llg_species* species;  /* [n_species] */
    const uint32_t* spin_species; /* [n] species of each spin (lu(i,4)); NULL: species = i % n_species */
    llg_anisotropy anis;
    llg_exchange exch;
    llg_dipolar dip;
    double dt;            /* s */
    int32_t scheme;       /* llg_scheme */
    int32_t precision;    /* llg_precision */
    int32_t noise;        /* llg_noise */
    uint32_t update;      /* sampling / dipolar-refresh interval (spins::update), >= 1 */
    uint64_t seed;        /* counter-noise key */
    uint64_t stream;      /* counter-noise stream id (replica / request) */
    uint32_t flags;
    int32_t device;       /* HIP device ordinal */
    llg_fourspin four;    /* optional four-spin term (zeroed by llg_desc_init) */
    llg_fftexch fftx;     /* optional FFT exchange (zeroed by llg_desc_init) */
} llg_desc;

typedef struct llg_ctx llg_ctx;

/* Version and device queries. */
uint32_t llg_abi_version(void);
const char* llg_last_error(void);
int llg_device_count(int* count);

/* Fill a descriptor with defaults (zeroes, abi_version, update = 1, scheme Heun, FP32). */
void llg_desc_init(llg_desc* d);

/* Create / destroy a handle.  The descriptor and every array it points to are
 * copied; they may be freed after llg_create returns. */
int llg_create(const llg_desc* d, llg_ctx** out);
void llg_destroy(llg_ctx* ctx);
/* Create a handle that shares the device-resident couplings of `proto` (same n, precision
 * and device; d->exch is ignored).  The couplings live until the last sharing handle is
 * destroyed.  No reference counterpart: for replicas and serving requests of one lattice. */
int llg_create_like(const llg_desc* d, llg_ctx* proto, llg_ctx** out);

/* Spin state, SoA host arrays of n doubles.  set_spins normalises nothing: the
 * caller passes unit vectors (zero-length vectors are rejected). */
int llg_set_spins(llg_ctx* ctx, const double* sx, const double* sy, const double* sz);
int llg_get_spins(llg_ctx* ctx, double* sx, double* sy, double* sz);

/* Inputs that drivers change between steps. */
int llg_set_temperature(llg_ctx* ctx, double T);
int llg_set_applied(llg_ctx* ctx, const double B[3]);
/* Staggered field (tesla), SoA n doubles each; all NULL sets it to zero.
 * llg_scale_staggered(r) sets H_stag = r * (last field given to llg_set_staggered). */
int llg_set_staggered(llg_ctx* ctx, const double* hx, const double* hy, const double* hz);
int llg_scale_staggered(llg_ctx* ctx, double r);

/* Injected noise (LLG_NOISE_INJECTED): appends nsteps arrays of 3n doubles laid out
 * [step][component][spin]; each step consumes one array (all stages of a step use
 * the same xi, as the reference does, src/cuda.cu:148, src/cudaRK4.cu:41). */
int llg_push_noise(llg_ctx* ctx, uint32_t nsteps, const double* xi);

/* Advance steps t0 .. t0+nsteps-1, i.e. nsteps calls of llg::integrate(t).  The
 * dipolar field refreshes at steps with t % refresh == 0.  Asynchronous: returns
 * after enqueueing on the handle's stream; the getters synchronise. */
int llg_step(llg_ctx* ctx, uint64_t t0, uint32_t nsteps);
int llg_sync(llg_ctx* ctx);

/* Deterministic field = total field minus the thermal part, evaluated where the
 * reference evaluates CDetFields: at the predicted spins of the last Heun step
 * (src/cuint.cu:235).  After RK4/RK5 steps, or before any step, it is evaluated at
 * the current spins. */
int llg_get_det_field(llg_ctx* ctx, double* hx, double* hy, double* hz);
/* Deterministic field at the current spins (testing aid for field assembly). */
int llg_get_field(llg_ctx* ctx, double* hx, double* hy, double* hz);

/* FP64 reductions on the device: m[s*3+c] = mean of S_c over sublattice s
 * (src/util_mag.cpp:44-63); energy (J, may be NULL) = exchange/2 + anisotropy +
 * Zeeman + staggered + dipolar/2 at the current spins, thermal part excluded. */
int llg_reduce(llg_ctx* ctx, double* m, double* energy);

/* llg_reduce(ctx, m, NULL) without any device allocation: the per-block partial sums and the
 * results go through a fixed device array of the library (197,120 bytes per GPU, held until the
 * process ends), so callers that cannot allocate (a tight serving loop, a harness that forbids
 * hipMalloc) can sample <M>.  No counterpart in the reference.  Same kernels and block count
 * as llg_reduce, so the same bits.  At most 16 sublattices (LLG_E_UNSUPPORTED beyond the
 * array).  Calls that overlap on one GPU (other host threads, unordered streams) race for
 * the array. */
int llg_reduce_mag_noalloc(llg_ctx* ctx, double* m);

/* Binned FP64 sums on the device: out[b*3+c] = sum over spins i with bin[i] == b of
 * weight[i] * S_c(i) (weight NULL = 1; bin[i] >= nbins leaves spin i out).  bin and weight
 * are in the reference spin order and are uploaded once by llg_set_bins.  This is the
 * device form of util::calc_mag methods 1-10 (src/util_mag.cpp:65-245): layer sums along
 * x/y/z weighted by mu (1-3), layer x species (5-7), chequerboard (8), discretised cells
 * (9), per unit-cell atom (10); the caller divides by counts where the reference does. */
int llg_set_bins(llg_ctx* ctx, uint32_t nbins, const uint32_t* bin, const double* weight);
int llg_reduce_bins(llg_ctx* ctx, double* out);

/* Transverse spin structure factor at mesh wave-vectors: the device form of the spin-wave
 * sampler of the time-series driver (src/timeseries.cpp:262-380, settings src/sf_glob.cpp).
 * Per spin c_i = S_a(i) ua[i] + i S_b(i) ub[i], (a, b) the components transverse to the
 * quantisation axis (axis 0: y, z; 1: x, z; 2: x, y); ua/ub carry the reference's
 * per-sublattice sign transform uo (NULL = 1).  For k-point k (mesh indices kpts[k][3]):
 *   out[(g*nk + k)*2 + {0,1}] = sum_{i: group[i] == g} c_i exp(-2 pi i sum_d kpts[k][d] mesh[i][d] / grid[d]),
 * the FFTW_FORWARD convention of the reference's plan, unnormalised (group NULL: one group of
 * every spin; group[i] >= ngroups leaves spin i out).  mesh (n x 3), group, ua, ub are in the
 * reference spin order; mesh[i][d] < grid[d] and kpts[k][d] < grid[d]. */
int llg_set_sq(llg_ctx* ctx, const uint32_t grid[3], const uint32_t* mesh, uint32_t nk, const uint32_t* kpts,
               int axis, const double* ua, const double* ub, uint32_t ngroups, const uint32_t* group);
int llg_sample_sq(llg_ctx* ctx, double* out);

/* Run nsteps steps and record the sublattice magnetisation on the device every
 * `every` steps, before the step, as the drivers sample (src/simple.cpp:87-95).
 * m_series receives ceil(nsteps/every) records of n_sublattice*3 doubles. */
int llg_run_sampled(llg_ctx* ctx, uint64_t t0, uint32_t nsteps, uint32_t every, double* m_series);

/* The thermal noise xi (before scaling by sqrt(T) sigma) that step t uses, as
 * 3n doubles [component][spin], generated on the device (counter mode only). */
int llg_noise_sample(llg_ctx* ctx, uint64_t t, double* xi);

/* The HIP stream (hipStream_t) every kernel of this handle runs on. */
void* llg_stream(llg_ctx* ctx);

/* Per-kernel timing: runs nsteps steps with HIP events around every launch and
 * returns the mean duration (ms) of each kernel kind.  names receives up to
 * max_kinds NUL-terminated names of up to 47 chars (48 bytes each). */
int llg_profile_kernels(llg_ctx* ctx, uint64_t t0, uint32_t nsteps, int max_kinds,
                        char* names, double* mean_ms, int* n_kinds);

/* Launch-shape autotuner (no reference counterpart): times the FP32 Heun kernel pair for
 * every candidate block size (256, 512, 1024 threads) with and without the XCD-aware
 * tile remap, for the packed encoding 2 or 4 coupling chunks loaded ahead per wave, and
 * with the thermal field drawn in both passes or handed from predictor to corrector,
 * (packed and line encodings) with neighbours gathered from the SoA spins or from an AoS3
 * copy the kernels keep beside them (bit-identical either way), and (line encodings) with
 * the line reader or the packed entries kept beside the line stream (same couplings, sums
 * in another order), on this handle's own couplings, for nsteps steps each; keeps the
 * fastest.
 * The spin state is restored afterwards.  Results are cached per (device, encoding, n,
 * entries) in the process and, if the environment variable SLLG_TUNE_FILE names a file,
 * across processes.  Other paths return LLG_OK with the defaults. */
typedef struct llg_tune_result {
    int tpb, xcd;
    double ms_per_step;     /* chosen */
    double ms_per_step_default; /* 256 threads, remap on, 2 chunks ahead, hand-off on, gather copy on, line reader */
    int candidates, from_cache;
    int unroll;
    int thermal_cache;  /* 1: the predictor hands its thermal field to the corrector */
    int gather_copy;    /* 1: neighbours are gathered from an AoS3 copy of the spins */
    int line_mode;      /* line encodings: 1 line reader, 2 LDS-staged line reader (when the encoding has
                           tile streams; else as 1), 0 the packed entries kept beside them */
} llg_tune_result;
int llg_autotune(llg_ctx* ctx, uint32_t nsteps, llg_tune_result* out);
int llg_set_launch(llg_ctx* ctx, int tpb, int xcd, int unroll, int thermal_cache, int gather_copy, int line_mode);

/* Algorithmic HBM bytes per spin-update of the configured path, split into the
 * kernels of one step, and the device bytes the handle holds. */
typedef struct llg_info {
    uint64_t device_bytes;
    double bytes_per_spin_update; /* whole step */
    double dominant_kernel_bytes_per_spin;
    char dominant_kernel[48];
    char exch_encoding[64];
} llg_info;
int llg_get_info(llg_ctx* ctx, llg_info* info);

#ifdef __cplusplus
}
#endif
#endif /* LLG_CAPI_H */
