// ctx.hpp - device-side description of one spin system and the host handle.
//
// HBM layout (all SoA, component-major: component c of spin i at [c*n + i]):
//   FP64 plain path : S, Y (predicted / stage input), H (exchange + demag), FN (f(S_n)),
//                     DET, XI (noise), per-spin sigma/llgpf/lambda/k1u/k1udir, CSR with
//                     FP64 values per nnz - the reference GPU data structures, SoA.
//   FP32 fused path : S, Y, A (= S + rdt/2 f(S)), per-species constants, a compact
//                     exchange encoding (sliced ELL whose u32 entry packs the neighbour's
//                     row offset with a coupling-dictionary index, repeated slices stored
//                     once; raw u32/FP32 when it does not fit; periodic stencil; DIA),
//                     noise regenerated in registers.
#pragma once
#include <hip/hip_runtime.h>

#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace sllg {

constexpr int kTPB = 256;       // threads per block of the per-spin kernels
constexpr int kSlice = 64;      // sliced-ELL slice height = one wavefront
constexpr int kMaxSub = 16;     // sublattices reduced on the device
constexpr int kDictLds = 2048;  // floats of coupling dictionary staged in LDS (8 KB)

// Line-mode group records: line_group(nc) neighbour windows per record, line_words(nc) words
// (the group's shift mask, its window starts, 3 shifts x nc couplings per window, the next
// group's window starts last; padded to whole 16-byte units).
constexpr uint32_t kLineFallback = 0xFFFFFFFFu;  // line-mode slice read from the packed entries
constexpr int line_group(int nc) { return nc == 1 ? 3 : (nc == 3 ? 2 : 1); }
constexpr int line_words(int nc) { return (1 + 2 * line_group(nc) + 3 * line_group(nc) * nc + 3) / 4 * 4; }

// LDS-staged line reader ("tile" reader, beside a 64-cell line encoding without packed
// slices): a workgroup of tnw slices stages the union of their neighbour lines in LDS as
// [line][component][64] floats, 768 bytes per line, and reads its windows from there.
constexpr uint32_t kTileLineBytes = 3 * 64 * 4;
constexpr uint32_t kTileLdsMax = 64 * 1024;  // dynamic LDS of one workgroup
constexpr int tile_head(int nw) { return (2 + 3 * nw + 7) / 8 * 8; }  // start of the staged-line list
// Tile records: line_group(nc) windows (their LDS byte offsets, the shift mask in the low 3
// bits of the first) and their couplings, tile_words(nc) words (12 for isotropic couplings).
constexpr int tile_words(int nc) { return (line_group(nc) + 3 * line_group(nc) * nc + 3) / 4 * 4; }

enum ExchFmt : int {
    FMT_NONE = 0,
    FMT_DIA = 1,
    FMT_CSR = 2,        // FP64 plain: CSR rows stored by 64-row slices, u32 columns + FP64 values
    FMT_STENCIL = 3,    // periodic stencil (shared or per-row values)
    FMT_SELL_PACK = 4,  // FP32: sliced ELL, u32 = (row offset << dbits) | dictionary index
    FMT_SELL_RAW = 5,   // FP32: sliced ELL, u32 col + FP32 values [ncomp][entries]
    FMT_SELL_LINE = 6,   // FP32: slices that are periodic 64-cell lines, lane-uniform scalar coupling stream
    FMT_SELL_LINE16 = 7,      // FP32: slices of four periodic 16-cell lines, same stream (rotations within 16 lanes)
    FMT_SELL_LINE_OPEN = 8,   // FP32: open 64-cell lines (zero-filling shifts instead of rotations)
    FMT_SELL_LINE16_OPEN = 9  // FP32: four open 16-cell lines per slice
};

constexpr bool is_line_fmt(int f) {
    return f == FMT_SELL_LINE || f == FMT_SELL_LINE16 || f == FMT_SELL_LINE_OPEN || f == FMT_SELL_LINE16_OPEN;
}
constexpr int line_len(int f) { return (f == FMT_SELL_LINE || f == FMT_SELL_LINE_OPEN) ? 64 : 16; }
constexpr bool line_open(int f) { return f == FMT_SELL_LINE_OPEN || f == FMT_SELL_LINE16_OPEN; }

// Per-species constants, precision R.
template <class R> struct Species {
    R llgpf, lambda, sigma, k1u, kc;
    R k1dir[3];
};

// Global anisotropy with prefactors 2K/(mu_0 mu_B) precomputed on the host in FP64.
// K2 perp / par and K4 per axis (src/cuint.cu:98-137): each term is linear in S,
// h += pf (S.d) d, pf = 2K/(mu_0 mu_B); the terms present, in the reference's order (K2 perp,
// K2 par, then per axis K4 perp and K4 par), as one list, so that absent terms cost nothing.
template <class R> struct AnisG {
    int nlin;
    int pad_;
    R lin_pf[8];
    R lin_dir[8][3];
};

// Exchange operands on the device (one struct for every format; unused fields null).
template <class V> struct Exch {
    int fmt = FMT_NONE, ncomp = 1;
    uint32_t n = 0;
    // CSR
    const uint32_t* xadj = nullptr;
    const uint32_t* adj = nullptr;
    const V* vals = nullptr;  // [ncomp][nnz] (CSR), [ncomp][ndiag][n] (DIA), stencil values
    uint64_t nnz = 0;
    // DIA
    const int32_t* doff = nullptr;
    uint32_t ndiag = 0;
    // stencil
    uint32_t dims[3] = {0, 0, 0}, na = 1, pbc[3] = {0, 0, 0}, sk = 0, per_row = 0;
    const int4* sent = nullptr;       // [na][sk] (dx, dy, dz, t')
    const uint32_t* scount = nullptr; // [na]
    // sliced ELL
    // [nslices] {first entry / 64, entries per row}; slices with identical packed entries
    // point at one stored copy
    const uint2* slice = nullptr;
    const uint32_t* ent = nullptr;        // packed entries or columns, [..][64]
    const V* dict = nullptr;              // [ndict][ncomp] (SELL_PACK) or raw values
    uint32_t ndict = 0;
    uint64_t nent = 0;                    // total entries (SELL_RAW value stride)
    uint32_t ncell = 0;                   // rows per sublattice block of the device order
    uint32_t dbits = 0;                   // dictionary-index bits of a packed entry
    // line mode (FMT_SELL_LINE / LINE16): per slice {first word, groups}; groups ==
    // kLineFallback sends the slice to the packed entries above
    const uint2* lslice = nullptr;
    const uint32_t* lent = nullptr;       // group records
    uint32_t lfall = 0;                   // slices that use the packed entries
    // tile reader: per workgroup the first word of its stream in tent (streams of equal
    // content stored once); tnw == 0: no tile streams
    const uint32_t* tgrp = nullptr;
    const uint32_t* tent = nullptr;
    uint32_t tnw = 0;                     // slices (waves) per workgroup
    uint32_t tumax = 0;                   // most lines any workgroup stages
    uint32_t tlps = 0;                    // slices per sublattice block
    // four-spin quartets of each row (same row order as the spins of this path)
    const uint32_t* fs_xadj = nullptr;    // [n+1]; null = term off
    const uint4* fs_jkl = nullptr;        // (j, k, l, 0)
    V fs_jq = V(0);
    int fs_div3 = 1;
};

// Device order of the FP32 sliced-ELL path.  The reference orders spins basis-atom
// fastest, i = cell * na + t (src/geom.cpp:214-266); on the device spin t of cell c sits
// at d = t * ncell + c ("sublattice-major").  Rows of one slice then share t, and the
// neighbours a wavefront gathers for one coupling entry are consecutive cells of one
// sublattice: contiguous loads instead of na-strided ones.
struct Perm {
    uint32_t na = 1, ncell = 0;
    __host__ __device__ uint32_t to_dev(uint32_t i) const { return (i % na) * ncell + i / na; }
    __host__ __device__ uint32_t to_ref(uint32_t d) const { return (d % ncell) * na + d / ncell; }
};

// Packed sliced-ELL entry: the signed device-row offset of the neighbour (j - i, exact, no
// wrap) in the high 32 - dbits bits and the coupling-dictionary index in the low dbits bits.
// The offset is the same for every row whose neighbourhood does not cross a periodic
// boundary, so slices away from the boundary planes encode identically.
__host__ __device__ __forceinline__ uint32_t pack_word(int32_t drow, uint32_t idx, uint32_t dbits) {
    return ((uint32_t)drow << dbits) | idx;
}
__host__ __device__ __forceinline__ uint32_t pack_dict(uint32_t w, uint32_t dbits) {
    return w & ((1u << dbits) - 1u);
}
// Neighbour device row of entry w in row i.
__host__ __device__ __forceinline__ uint32_t pack_nb(uint32_t w, uint32_t i, uint32_t dbits) {
    return i + (uint32_t)((int32_t)w >> dbits);
}
// Position of entry q of lane `lane` within a packed slice block of len (even) entries per
// row: full chunks of 4 stored [chunk][lane][4], then a tail of 2 stored [lane][2].
__host__ __device__ __forceinline__ size_t pack_pos(uint32_t q, uint32_t lane, uint32_t len) {
    const uint32_t full = len & ~3u;
    return q < full ? (size_t)(q >> 2) * (kSlice * 4) + lane * 4 + (q & 3)
                    : (size_t)full * kSlice + lane * 2 + (q - full);
}

struct DipolarPlan;  // k_dipolar.hip

// Host-side snapshot of the exchange given at create time (FP64, reference layout).
struct HostExch {
    int fmt = 0, ncomp = 1;
    std::vector<uint32_t> xadj, adj;
    std::vector<double> vals;
    uint64_t nnz = 0;
    std::vector<int32_t> doff;
    uint32_t ndiag = 0;
    uint32_t dims[3] = {0, 0, 0}, na = 1, pbc[3] = {0, 0, 0}, sk = 0, per_row = 0;
    std::vector<uint32_t> scount;
    std::vector<int32_t> sent;
};

struct KernelTimer;  // capi.hip

// Device-resident exchange encoding, immutable after llg_create and shared by every
// handle created from it (llg_create_like): replicas and serving requests of one lattice
// hold one copy of the couplings.
struct ExchStore {
    int device = 0, prec = 0;
    uint32_t n = 0;
    DevBuf a, b, c, d, e, f, g, dict64, fs_x, fs_q;
    Exch<float> ex32;
    Exch<double> ex64;
    Perm perm;
    std::string encoding;
    HostExch h;  // scalars kept; the large host vectors are released after upload
    uint64_t bytes() const {
        return a.bytes + b.bytes + c.bytes + d.bytes + e.bytes + f.bytes + g.bytes + dict64.bytes + fs_x.bytes +
               fs_q.bytes;
    }
};

}  // namespace sllg

struct llg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false while the handle is a member of a batch (batch stream)
    uint32_t n = 0, ns = 1, nsub = 1;
    int scheme = LLG_HEUN, prec = LLG_FP32, noise_mode = LLG_NOISE_COUNTER;
    uint64_t seed = 0, stream_id = 0;
    uint32_t update = 1, flags = 0;
    double dt = 0, rdt = 0, T = 0, B[3] = {0, 0, 0};
    double basemm = 0;
    bool stag_set = false, det_valid = false, dip = false;
    int dip_every_stage = 0;
    uint32_t dip_refresh = 1;
    std::vector<double> mu_sp;      // per species (mu_B)
    std::vector<uint32_t> sub_sp;   // per species
    std::vector<uint64_t> sub_count;
    llg_anisotropy anis_in{};
    sllg::AnisG<double> an64{};
    sllg::AnisG<float> an32{};
    std::vector<sllg::Species<double>> sp64;
    bool multi_species = false;
    sllg::Perm perm;  // device order of the FP32 sliced-ELL path (identity elsewhere)
    bool permuted() const { return perm.na > 1; }
    int tune_tpb = 256;  // block size of the FP32 Heun kernels (llg_autotune)
    int tune_xcd = 1;    // XCD-aware tile remap on/off (llg_autotune)
    int tune_unr = 2;    // packed-exchange chunks loaded ahead per wave (llg_autotune)
    int tune_thc = 1;    // predictor hands its thermal field to the corrector (llg_autotune)
    int tune_g3 = 1;     // gather neighbours from the AoS3 copy of the spins (llg_autotune)
    int tune_line = 1;   // line encodings: 0 packed entries, 1 line reader, 2 tile reader (llg_autotune)
    bool g3_valid = false;   // S3 holds the current S
    bool s_soa_valid = true;  // S (SoA) holds the current spins (else only S3 does: see sync_spins)
    bool y_soa_valid = true;  // Y (SoA) holds the last predictor output (else only Y3 does)

    // device state
    sllg::DevBuf S, Y, A, H, FN, DET, XI, STG, DEM;  // SoA 3n (type per path)
    sllg::DevBuf STG0;  // staggered field as last given, FP64 SoA 3n in device order (ramped by llg_scale_staggered)
    sllg::DevBuf HT;  // FP32 Heun: thermal field of the current step, predictor -> corrector
    sllg::DevBuf S3, Y3;  // FP32 Heun: AoS3 gather copies [3 i + c] of S and Y
    sllg::DevBuf K[6];                                     // RK stage increments
    sllg::DevBuf spec;      // u8 species per spin (FP32 path, multi-species)
    sllg::DevBuf sub;       // u8 sublattice per spin
    sllg::DevBuf sp32, sp64d;  // Species tables
    sllg::DevBuf p_sigma, p_llgpf, p_lambda, p_k1u, p_k1udir, p_kc;  // FP64 plain per-spin arrays
    // exchange: shared encoded store + by-value views of it for the kernels
    std::shared_ptr<sllg::ExchStore> xs;
    sllg::Exch<double> ex64;  // FP64 plain couplings, or FP64 dictionary/values of the FP32 path
    sllg::Exch<float> ex32;
    // reductions
    sllg::DevBuf red_part, red_out, m_series;
    // binned sums (llg_set_bins): device rows per chunk of one bin, weights, chunk table
    sllg::DevBuf bin_idx, bin_w, bin_cs, bin_cb, bin_part;
    uint32_t nbins = 0, nbin_chunks = 0;
    // transverse structure factor (llg_set_sq): per listed spin {device row, mesh x, y, z},
    // weights (ua, ub), chunk table, wave-vectors
    sllg::DevBuf sq_spin, sq_w, sq_cs, sq_cg, sq_k, sq_part;
    uint32_t sq_nk = 0, sq_ngroups = 0, sq_nch = 0, sq_grid[3] = {0, 0, 0};
    int sq_axis = 2;
    // dipolar
    sllg::DipolarPlan* dplan = nullptr;
    // injected noise queue (host)
    std::deque<std::vector<double>> noise_q;
    // timing
    sllg::KernelTimer* timer = nullptr;
};

namespace sllg {
// Encoding the FP32 kernels run with: a line encoding keeps the packed entries beside its
// line stream, so the packed reader can run instead (tuned per handle).
inline int exec_fmt(const llg_ctx* c) {
    const int f = c->ex32.fmt;
    return (is_line_fmt(f) && !c->tune_line) ? FMT_SELL_PACK : f;
}
}  // namespace sllg
