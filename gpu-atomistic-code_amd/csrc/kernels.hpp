// kernels.hpp - host launchers of the device kernels (each returns 0 or an llg_status).
#pragma once
#include "ctx.hpp"

namespace sllg {

// FP64 plain path (k_plain.hip)
int plain_spmv(llg_ctx* c, const double* S, double* H);
int plain_noise(llg_ctx* c, uint64_t t);
int plain_heun1(llg_ctx* c);
int plain_heun2(llg_ctx* c);
int plain_rk_init_tables();
int plain_rk_stage(llg_ctx* c, int m);
int plain_rk_comb(llg_ctx* c);
int plain_detfield(llg_ctx* c, const double* Sp, double* out);
int stag_scale(llg_ctx* c, double r);  // STG = r * STG0 (FP64 H0 kept on the device)

// FP32 fused path (k_fused.hip)
struct StepSched {
    uint64_t t;        // step number (counter noise, dipolar refresh)
    const double* xi;  // injected noise for this step (device, 3n doubles) or null
};
int fused_heun_pred(llg_ctx* c, const StepSched& s);
int fused_heun_corr(llg_ctx* c, const StepSched& s);
int fused_y_soa(llg_ctx* c);  // rebuild the SoA predictor output from its AoS3 copy
int sync_spins(llg_ctx* c);
int reduce_mag_noalloc(llg_ctx* c, double* m);  // <M> per sublattice through the device scratch array   // rebuild the SoA spins from their AoS3 copy when only that is current
int fused_rk_stage(llg_ctx* c, const StepSched& s, int m);
int fused_detfield(llg_ctx* c, const float* Sp, double* out);
int fused_rk_init_tables();
// Batched Heun over several FP32 handles of one (format, ncomp) class (k_fused.hip).
size_t batch_member_size();
// g3: the kernels gather from AoS3 copies of S and Y (built here where stale)
int batch_build_table(const std::vector<llg_ctx*>& members, const std::vector<uint64_t>& t0, uint32_t tpb,
                      bool g3, DevBuf& dmem, DevBuf& dmap, uint32_t& ntiles, hipStream_t s);
int batch_heun_launch(hipStream_t s, int fmt, int nc, const void* dmem, const void* dmap, uint32_t ntiles,
                      uint32_t step_off, int phase, uint32_t tpb, int xcd, bool g3);
bool batch_needs_y_soa(const llg_ctx* c);  // the member's predictor must also write the SoA Y
// Grid barrier of a persistent batched launch (device memory).  Blocks arrive on one of 8
// group counters (block % 8, each on its own 128-byte line); the last block of a group
// arrives on `count`, waits for all groups there and then raises its group's flag, which the
// group's other blocks wait on.  Agent-scope atomics on one address serialise, and 1,280
// blocks polling one word contend, so each address sees at most a group.  The counters and
// flags are cleared before every launch; `abort` is raised by a block whose wait ran out of
// polls and stays raised until the host clears it; `polls` is the longest wait of a group's
// last block.
struct GridBar {
    uint32_t count;
    uint32_t abort;
    uint32_t polls;
    uint32_t pad_[29];
    uint32_t group[8][32];
    uint32_t flag[8][32];
};
// Persistent batched Heun: nsteps steps (from step offset step_off) of every member of the
// table in one cooperative launch of at most as many blocks as fit on the GPU at once, with a
// grid barrier between half steps whose every wait ends after spin_limit polls.  Clears
// bar->count first; *grid receives the launch's block count.  nsteps == 0: only find how
// many blocks per CU are resident together (probe launches, synchronous, once per kernel
// shape and device), so a later launch can be captured into a hipGraph.
int batch_heun_persistent(hipStream_t s, int fmt, int nc, const void* dmem, const void* dmap, uint32_t ntiles,
                          uint32_t step_off, uint32_t nsteps, uint32_t tpb, int xcd, bool g3, GridBar* bar,
                          uint32_t spin_limit, uint32_t* grid);
int dipolar_set_stream(llg_ctx* c);
// Handle over an existing coupling store (capi.hip).
int create_with_store(const llg_desc* d, const std::shared_ptr<ExchStore>& store, llg_ctx** out);

// Reductions (k_reduce.hip): FP64 sublattice magnetisation and energy.
int reduce_mag(llg_ctx* c, double* dev_out /* nsub*3 */);
int reduce_energy(llg_ctx* c, const void* dem_now /* dipolar field at the current spins or null */,
                  double* dev_out /* 1 */);
int noise_sample(llg_ctx* c, uint64_t t, double* dev_out);
int check_finite(llg_ctx* c, int* bad);
int set_bins(llg_ctx* c, uint32_t nbins, const uint32_t* bin, const double* weight);
int reduce_bins(llg_ctx* c, double* dev_out /* nbins*3 */);
int set_sq(llg_ctx* c, const uint32_t grid[3], const uint32_t* mesh, uint32_t nk, const uint32_t* kpts, int axis,
           const double* ua, const double* ub, uint32_t ngroups, const uint32_t* group);
int sample_sq(llg_ctx* c, double* dev_out /* ngroups*nk*2 */);

// Dipolar convolution (k_dipolar.hip)
// fx != null: FFT exchange (species-resolved 3x3 kernel; dipolar terms added when dip_terms)
int dipolar_create(llg_ctx* c, const llg_dipolar& d, const uint32_t* spin_species, const llg_fftexch* fx,
                   bool dip_terms);
void dipolar_destroy(llg_ctx* c);
uint64_t dipolar_bytes(const llg_ctx* c);  // plan buffers and FFT work areas (a shared kernel counted in full)
bool dipolar_is_exchange(const llg_ctx* c);  // the plan is an FFT-exchange plan
bool dipolar_capturable(const llg_ctx* c);   // a refresh is this library's kernels only (no hipFFT plan)
int dipolar_refresh(llg_ctx* c, const void* S, void* out);  // spins of the path's precision -> out (3n)

// Encodings (encode.cpp)
int build_exchange_fp64(llg_ctx* c, const llg_exchange& e);
int build_exchange_fp32(llg_ctx* c, const llg_exchange& e);
int build_fourspin(llg_ctx* c, const llg_fourspin& f);

}  // namespace sllg
