// k_reduce.hip - on-device FP64 observables, replacing the reference's D2H copy of the
// spins every spins::update steps and the host loop of util::calc_mag method 0
// (src/cuda.cu:234-256, src/util_mag.cpp:44-63).
//
// Determinism: each block reduces its grid-stride share with 64-wide wavefront
// shuffles and a fixed-order combine of its 4 waves; a single block then sums the
// per-block partials in a fixed order.  No floating-point atomics, so the result is
// bitwise reproducible for a given n and launch shape.
#include <mutex>

#include "field.hpp"
#include "kernels.hpp"
#include "philox.hpp"

namespace sllg {

constexpr int kRedBlocks = 512;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum of v in fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kTPB / 64; ++k) r += lds[k];
    return r;
}

template <class SpinT>
__global__ __launch_bounds__(kTPB) void k_mag_partial(uint32_t n, uint32_t nsub, const SpinT* __restrict__ S,
                                                      const uint8_t* __restrict__ sub, double* __restrict__ part) {
    __shared__ double lds[kTPB / 64];
    const uint32_t stride = gridDim.x * kTPB;
    for (uint32_t s = 0; s < nsub; ++s) {
        double a[3] = {0.0, 0.0, 0.0};
        for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < n; i += stride) {
            if (sub && sub[i] != s) continue;
            a[0] += (double)S[i];
            a[1] += (double)S[(size_t)n + i];
            a[2] += (double)S[2 * (size_t)n + i];
        }
        for (int c = 0; c < 3; ++c) {
            const double r = block_sum(a[c], lds);
            if (threadIdx.x == 0) part[(size_t)blockIdx.x * nsub * 3 + s * 3 + c] = r;
        }
    }
}

// out[v] = scale[v / 3] * sum_b part[b][v]
__global__ __launch_bounds__(kTPB) void k_sum_partials(uint32_t nblocks, uint32_t nval, const double* __restrict__ part,
                                                       const double* __restrict__ scale, double* __restrict__ out) {
    __shared__ double lds[kTPB / 64];
    for (uint32_t v = 0; v < nval; ++v) {
        double a = 0.0;
        for (uint32_t b = threadIdx.x; b < nblocks; b += kTPB) a += part[(size_t)b * nval + v];
        const double r = block_sum(a, lds);
        if (threadIdx.x == 0) out[v] = r * (scale ? scale[v / 3] : 1.0);
    }
}

// Per-spin energy (J) in FP64:
//   mu mu_B [ -1/2 S.(H_ex + H_dip + H_quad) - 1/4 S.(H_cubic + H_4spin) - S.(B + H_stag) ]
// where H_quad collects the terms linear in S (uniaxial, K2, K4).
struct EnergyArgs {
    uint32_t n;
    int multi;
    double B[3];
    const void* dem;  // float (FP32 path) or double (plain), current dipolar field or null
    const void* stg;
    int f64;
    const uint8_t* spec;
    const Species<double>* sp;
    const double* mu;  // per species, mu_B units
    AnisG<double> an;
    Exch<double> ex;
};

template <int FMT, int NC, class SpinT>
__device__ __forceinline__ void exch_dispatch(const Exch<double>& e, uint32_t i, const SpinT* S, double dot[3]);

template <int FMT, int NC, class SpinT>
__global__ __launch_bounds__(kTPB) void k_energy_partial(EnergyArgs a, const SpinT* __restrict__ S,
                                                         double* __restrict__ part);

}  // namespace sllg

// The sliced-ELL row reader lives in k_fused.hip; the energy kernel needs the same
// traversal in FP64, so it is restated here for the reduction's own translation unit.
namespace sllg {

template <int NC, class SpinT>
__device__ __forceinline__ void exch_row_sell64(const Exch<double>& e, uint32_t i, const SpinT* __restrict__ S,
                                                double dot[3]) {
    const uint32_t n = e.n;
    dot[0] = dot[1] = dot[2] = 0.0;
    const uint32_t sl = i >> 6, lane = i & 63u;
    const uint2 sd = e.slice[sl];
    const uint32_t q0 = sd.x, len = sd.y;
    double v[NC];
    for (uint32_t q = 0; q < len; ++q) {
        uint32_t w, j;
        if (e.fmt == FMT_SELL_PACK)
            w = e.ent[(size_t)q0 * kSlice + pack_pos(q, lane, len)];
        else
            w = e.ent[((size_t)q0 + q) * kSlice + lane];
        if (e.fmt == FMT_SELL_PACK) {
            j = pack_nb(w, i, e.dbits);
            const double* dv = e.dict + (size_t)pack_dict(w, e.dbits) * NC;
            for (int k = 0; k < NC; ++k) v[k] = dv[k];
        } else {
            j = w;
            for (int k = 0; k < NC; ++k) v[k] = e.dict[(size_t)k * e.nent + ((size_t)q0 + q) * kSlice + lane];
        }
        const double s[3] = {(double)S[j], (double)S[n + j], (double)S[2 * (size_t)n + j]};
        accum<NC, double>(v, s, dot);
    }
}

template <int FMT, int NC, class SpinT>
__device__ __forceinline__ void exch_dispatch(const Exch<double>& e, uint32_t i, const SpinT* S, double dot[3]) {
    if constexpr (FMT == FMT_SELL_PACK || FMT == FMT_SELL_RAW)
        exch_row_sell64<NC, SpinT>(e, i, S, dot);
    else if constexpr (FMT == FMT_NONE)
        dot[0] = dot[1] = dot[2] = 0.0;
    else
        exch_row_ref<NC, double, double, SpinT>(e, i, S, dot);
}

template <int FMT, int NC, class SpinT>
__global__ __launch_bounds__(kTPB) void k_energy_partial(EnergyArgs a, const SpinT* __restrict__ S,
                                                         double* __restrict__ part) {
    __shared__ double lds[kTPB / 64];
    const uint32_t n = a.n;
    const uint32_t stride = gridDim.x * kTPB;
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < n; i += stride) {
        double dot[3];
        exch_dispatch<FMT, NC, SpinT>(a.ex, i, S, dot);
        const double s[3] = {(double)S[i], (double)S[(size_t)n + i], (double)S[2 * (size_t)n + i]};
        const uint32_t spx = a.multi ? a.spec[i] : 0;
        const Species<double> sp = a.sp[spx];
        double hp[3] = {dot[0], dot[1], dot[2]};
        if (a.dem) {
            for (int c = 0; c < 3; ++c)
                hp[c] += a.f64 ? ((const double*)a.dem)[(size_t)c * n + i] : (double)((const float*)a.dem)[(size_t)c * n + i];
        }
        double hq[3] = {0.0, 0.0, 0.0};
        local_terms<double>(sp.k1u, sp.k1dir, 0.0, a.an, s, hq);
        double hz[3];
        for (int c = 0; c < 3; ++c) {
            hz[c] = a.B[c];
            if (a.stg)
                hz[c] += a.f64 ? ((const double*)a.stg)[(size_t)c * n + i] : (double)((const float*)a.stg)[(size_t)c * n + i];
        }
        const double ep = s[0] * hp[0] + s[1] * hp[1] + s[2] * hp[2];
        const double eq = s[0] * hq[0] + s[1] * hq[1] + s[2] * hq[2];
        const double ec = sp.kc * (s[0] * s[0] * s[0] * s[0] + s[1] * s[1] * s[1] * s[1] + s[2] * s[2] * s[2] * s[2]);
        const double ez = s[0] * hz[0] + s[1] * hz[1] + s[2] * hz[2];
        double e4 = 0.0;  // four-spin: quartic in the spins, -1/4 S.H4
        if (a.ex.fs_xadj) {
            double h4[3];
            fourspin_row<double, double, SpinT>(a.ex, i, S, h4);
            e4 = s[0] * h4[0] + s[1] * h4[1] + s[2] * h4[2];
        }
        acc += a.mu[spx] * LLG_MU_B * (-0.5 * ep - 0.5 * eq - 0.25 * ec - 0.25 * e4 - ez);
    }
    const double r = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ __launch_bounds__(kTPB) void k_noise_f64(uint32_t n, uint64_t seed, uint64_t stream, uint64_t t, double* out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    double z[3];
    normals3_f64(seed, stream, i, t, z);
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = z[c];
}

__global__ __launch_bounds__(kTPB) void k_noise_f32(uint32_t n, Perm perm, uint64_t seed, uint64_t stream, uint64_t t,
                                                    double* out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    float z[3];
    normals3_f32(seed, stream, perm.to_ref(i), t, z);
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = (double)z[c];
}

template <class SpinT>
__global__ __launch_bounds__(kTPB) void k_check_finite(uint32_t n3, const SpinT* __restrict__ S, int* flag) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n3) return;
    const double v = (double)S[i];
    if (!(v == v) || v > 1e300 || v < -1e300) atomicOr(flag, 1);
}

// Binned FP64 sums out[b] = sum over the spins of bin b of w_i S_i.  The spins of a bin are
// listed in ascending device order and cut into chunks of at most kBinChunk; one block
// reduces one chunk (fixed-order block sum), then one thread per bin adds its chunks in
// order.  Deterministic, no atomics.
constexpr uint32_t kBinChunk = 4096;

template <class SpinT>
__global__ __launch_bounds__(kTPB) void k_bin_partial(uint32_t n, const uint32_t* __restrict__ cs,
                                                      const uint32_t* __restrict__ idx, const double* __restrict__ w,
                                                      const SpinT* __restrict__ S, double* __restrict__ part) {
    __shared__ double lds[kTPB / 64];
    const uint32_t b = cs[blockIdx.x], e = cs[blockIdx.x + 1];
    double a[3] = {0.0, 0.0, 0.0};
    for (uint32_t q = b + threadIdx.x; q < e; q += kTPB) {
        const uint32_t i = idx[q];
        const double wq = w ? w[q] : 1.0;
        a[0] += wq * (double)S[i];
        a[1] += wq * (double)S[(size_t)n + i];
        a[2] += wq * (double)S[2 * (size_t)n + i];
    }
    for (int c = 0; c < 3; ++c) {
        const double r = block_sum(a[c], lds);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * 3 + c] = r;
    }
}

__global__ __launch_bounds__(kTPB) void k_bin_finish(uint32_t nbins, const uint32_t* __restrict__ cb,
                                                     const double* __restrict__ part, double* __restrict__ out) {
    const uint32_t b = blockIdx.x * kTPB + threadIdx.x;
    if (b >= nbins) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = cb[b]; k < cb[b + 1]; ++k)
        for (int c = 0; c < 3; ++c) a[c] += part[(size_t)k * 3 + c];
    for (int c = 0; c < 3; ++c) out[(size_t)b * 3 + c] = a[c];
}

// Transverse structure factor at a list of mesh wave-vectors k:
//   out[g][k] = sum_{i in g} (S_a(i) ua_i + i S_b(i) ub_i) exp(-2 pi i sum_d k_d x_d(i) / N_d).
// The spins of each group are cut into chunks of kSqChunk; a block stages one chunk's
// (c_i, mesh position) in LDS and each thread accumulates one k over it in FP64; a second
// pass adds the chunks of each group in order.  Phase: frac(sum_d k_d x_d / N_d) in FP64,
// then sincospi(2 frac).
constexpr uint32_t kSqChunk = 512;

struct SqArgs {
    uint32_t n, nk, axis;
    double inv[3];             // 1 / N_d
    const uint4* spin;         // {device row, x, y, z} per listed spin
    const double2* w;          // (ua, ub)
    const uint32_t* cs;        // chunk starts [nch + 1]
    const uint32_t* kp;        // [nk][3]
};

template <class SpinT>
__global__ __launch_bounds__(kTPB) void k_sq_partial(SqArgs a, const SpinT* __restrict__ S, double* __restrict__ part) {
    __shared__ double sre[kSqChunk], sim[kSqChunk], px[kSqChunk], py[kSqChunk], pz[kSqChunk];
    const uint32_t ch = blockIdx.x, b = a.cs[ch], e = a.cs[ch + 1];
    const uint32_t ca = a.axis == 0 ? 1u : 0u, cb = a.axis == 2 ? 1u : 2u;
    for (uint32_t q = b + threadIdx.x; q < e; q += kTPB) {
        const uint4 sp = a.spin[q];
        const double2 w = a.w[q];
        const uint32_t l = q - b;
        sre[l] = (double)S[(size_t)ca * a.n + sp.x] * w.x;
        sim[l] = (double)S[(size_t)cb * a.n + sp.x] * w.y;
        px[l] = (double)sp.y * a.inv[0];
        py[l] = (double)sp.z * a.inv[1];
        pz[l] = (double)sp.w * a.inv[2];
    }
    __syncthreads();
    const uint32_t k = blockIdx.y * kTPB + threadIdx.x;
    if (k >= a.nk) return;
    const double kx = a.kp[3 * k], ky = a.kp[3 * k + 1], kz = a.kp[3 * k + 2];
    double re = 0.0, im = 0.0;
    for (uint32_t l = 0; l < e - b; ++l) {
        double t = kx * px[l] + ky * py[l] + kz * pz[l];
        t -= floor(t);
        double sn, cn;
        sincospi(2.0 * t, &sn, &cn);
        // (c_re + i c_im)(cos - i sin)
        re += sre[l] * cn + sim[l] * sn;
        im += sim[l] * cn - sre[l] * sn;
    }
    part[((size_t)ch * a.nk + k) * 2] = re;
    part[((size_t)ch * a.nk + k) * 2 + 1] = im;
}

__global__ __launch_bounds__(kTPB) void k_sq_finish(uint32_t ngroups, uint32_t nk, const uint32_t* __restrict__ cg,
                                                    const double* __restrict__ part, double* __restrict__ out) {
    const uint32_t v = blockIdx.x * kTPB + threadIdx.x;
    if (v >= ngroups * nk) return;
    const uint32_t g = v / nk, k = v % nk;
    double re = 0.0, im = 0.0;
    for (uint32_t ch = cg[g]; ch < cg[g + 1]; ++ch) {
        re += part[((size_t)ch * nk + k) * 2];
        im += part[((size_t)ch * nk + k) * 2 + 1];
    }
    out[(size_t)v * 2] = re;
    out[(size_t)v * 2 + 1] = im;
}

// ---------------------------------------------------------------- host launchers

static int ensure_red(llg_ctx* c) {
    const size_t need = (size_t)kRedBlocks * (size_t)std::max<uint32_t>(c->nsub * 3, 1) * sizeof(double);
    if (c->red_part.bytes < need) SLLG_TRY(c->red_part.alloc(need));
    if (c->red_out.bytes < 64 * sizeof(double) + kMaxSub * sizeof(double) + 64 * sizeof(double))
        SLLG_TRY(c->red_out.alloc(64 * sizeof(double) + kMaxSub * sizeof(double) + 64 * sizeof(double)));
    return 0;
}

static unsigned red_blocks(uint32_t n) { return std::min<unsigned>(kRedBlocks, std::max<unsigned>(1, blocks_for(n, kTPB))); }

int reduce_mag(llg_ctx* c, double* dev_out) {
    SLLG_TRY(sync_spins(c));
    SLLG_TRY(ensure_red(c));
    const unsigned nb = red_blocks(c->n);
    const uint8_t* sub = (c->nsub > 1) ? c->sub.as<uint8_t>() : nullptr;
    // inverse sublattice counts live after the first 64 doubles of red_out
    double* inv = c->red_out.as<double>() + 64;
    std::vector<double> hinv(c->nsub);
    for (uint32_t s = 0; s < c->nsub; ++s) hinv[s] = c->sub_count[s] ? 1. / (double)c->sub_count[s] : 0.0;
    SLLG_HIP(hipMemcpyAsync(inv, hinv.data(), c->nsub * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->prec == LLG_FP64_PLAIN)
        k_mag_partial<double><<<nb, kTPB, 0, c->stream>>>(c->n, c->nsub, c->S.as<double>(), sub, c->red_part.as<double>());
    else
        k_mag_partial<float><<<nb, kTPB, 0, c->stream>>>(c->n, c->nsub, c->S.as<float>(), sub, c->red_part.as<double>());
    SLLG_HIP(hipGetLastError());
    k_sum_partials<<<1, kTPB, 0, c->stream>>>(nb, c->nsub * 3, c->red_part.as<double>(), inv, dev_out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- allocation-free <M>

// Device scratch of llg_reduce_mag_noalloc: one file-scope array per GPU (each GPU that runs
// the library holds its own copy until the process ends).  Layout of one call:
//   [0, nb nsub 3)          per-block partial sums (every slot written by its block)
//   [kPartDoubles, + 16)    1 / (spins per sublattice)
//   [kPartDoubles + 16, +48) the results, nsub x 3
// Every call on one GPU uses the same array: calls that overlap on one GPU (two host
// threads, or two streams without ordering between them) race.
constexpr size_t kPartDoubles = (size_t)kRedBlocks * kMaxSub * 3;
constexpr size_t kScratchDoubles = kPartDoubles + kMaxSub + kMaxSub * 3;  // 197,120 bytes
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};  // device address of each GPU's copy of `scratch`
}  // namespace

// Address of `scratch` on device dev (current), looked up once per device.
static int scratch_of(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) SLLG_FAIL(LLG_E_INVALID, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            SLLG_FAIL(LLG_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

int reduce_mag_noalloc(llg_ctx* c, double* m) {
    const uint32_t nsub = c->nsub;
    const unsigned nb = red_blocks(c->n);
    if (nsub == 0 || nsub > kMaxSub || (size_t)nb * nsub * 3 > kPartDoubles)
        SLLG_FAIL(LLG_E_UNSUPPORTED, "allocation-free reduction: %u sublattices x %u blocks exceed its %zu-double "
                  "device scratch", nsub, nb, kScratchDoubles);
    double* base = nullptr;
    SLLG_TRY(scratch_of(c->device, &base));
    double* part = base;
    double* inv = base + kPartDoubles;
    double* out = inv + kMaxSub;
    SLLG_TRY(sync_spins(c));  // rebuilds S in place when only its AoS3 copy is current
    double hinv[kMaxSub];
    for (uint32_t s = 0; s < nsub; ++s) hinv[s] = c->sub_count[s] ? 1. / (double)c->sub_count[s] : 0.0;
    SLLG_HIP(hipMemcpyAsync(inv, hinv, nsub * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const uint8_t* sub = (nsub > 1) ? c->sub.as<uint8_t>() : nullptr;
    if (c->prec == LLG_FP64_PLAIN)
        k_mag_partial<double><<<nb, kTPB, 0, c->stream>>>(c->n, nsub, c->S.as<double>(), sub, part);
    else
        k_mag_partial<float><<<nb, kTPB, 0, c->stream>>>(c->n, nsub, c->S.as<float>(), sub, part);
    SLLG_HIP(hipGetLastError());
    k_sum_partials<<<1, kTPB, 0, c->stream>>>(nb, nsub * 3, part, inv, out);
    SLLG_HIP(hipGetLastError());
    SLLG_HIP(hipMemcpyAsync(m, out, nsub * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int reduce_energy(llg_ctx* c, const void* dem_now, double* dev_out) {
    SLLG_TRY(sync_spins(c));
    SLLG_TRY(ensure_red(c));
    const unsigned nb = red_blocks(c->n);
    EnergyArgs a{};
    a.n = c->n;
    a.multi = c->multi_species ? 1 : 0;
    for (int k = 0; k < 3; ++k) a.B[k] = c->B[k];
    a.f64 = c->prec == LLG_FP64_PLAIN;
    a.dem = dem_now;
    a.stg = c->stag_set ? c->STG.p : nullptr;
    a.spec = c->spec.as<uint8_t>();
    a.sp = c->sp64d.as<Species<double>>();
    a.mu = c->red_out.as<double>() + 64 + kMaxSub;
    SLLG_HIP(hipMemcpyAsync((void*)a.mu, c->mu_sp.data(), c->mu_sp.size() * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    a.an = c->an64;
    a.ex = c->ex64;
    double* part = c->red_part.as<double>();
    const int fmt = c->ex64.fmt;  // the FP32 path's FP64 view is the packed or raw encoding
    const int nc = c->ex64.ncomp;
#define SLLG_E(F, N)                                                                                 \
    do {                                                                                             \
        if (a.f64)                                                                                   \
            k_energy_partial<F, N, double><<<nb, kTPB, 0, c->stream>>>(a, c->S.as<double>(), part);  \
        else                                                                                         \
            k_energy_partial<F, N, float><<<nb, kTPB, 0, c->stream>>>(a, c->S.as<float>(), part);    \
    } while (0)
#define SLLG_E_NC(F)                         \
    do {                                     \
        if (nc == 1) SLLG_E(F, 1);           \
        else if (nc == 3) SLLG_E(F, 3);      \
        else SLLG_E(F, 9);                   \
    } while (0)
    switch (fmt) {
        case FMT_NONE: SLLG_E(FMT_NONE, 1); break;
        case FMT_CSR: SLLG_E_NC(FMT_CSR); break;
        case FMT_DIA: SLLG_E_NC(FMT_DIA); break;
        case FMT_STENCIL: SLLG_E_NC(FMT_STENCIL); break;
        case FMT_SELL_PACK: SLLG_E_NC(FMT_SELL_PACK); break;
        default: SLLG_E_NC(FMT_SELL_RAW); break;
    }
#undef SLLG_E_NC
#undef SLLG_E
    SLLG_HIP(hipGetLastError());
    k_sum_partials<<<1, kTPB, 0, c->stream>>>(nb, 1, part, nullptr, dev_out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

int noise_sample(llg_ctx* c, uint64_t t, double* dev_out) {
    if (c->prec == LLG_FP64_PLAIN)
        k_noise_f64<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, c->seed, c->stream_id, t, dev_out);
    else
        k_noise_f32<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, c->perm, c->seed, c->stream_id, t, dev_out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

int check_finite(llg_ctx* c, int* bad) {
    SLLG_TRY(sync_spins(c));
    SLLG_TRY(ensure_red(c));
    int* flag = reinterpret_cast<int*>(c->red_out.as<double>() + 63);
    SLLG_HIP(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    const uint32_t n3 = 3 * c->n;
    if (c->prec == LLG_FP64_PLAIN)
        k_check_finite<double><<<blocks_for(n3, kTPB), kTPB, 0, c->stream>>>(n3, c->S.as<double>(), flag);
    else
        k_check_finite<float><<<blocks_for(n3, kTPB), kTPB, 0, c->stream>>>(n3, c->S.as<float>(), flag);
    SLLG_HIP(hipGetLastError());
    SLLG_HIP(hipMemcpyAsync(bad, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int set_bins(llg_ctx* c, uint32_t nbins, const uint32_t* bin, const double* weight) {
    const uint32_t n = c->n;
    // spins of each bin in ascending device order (bin and weight are given in reference order)
    std::vector<uint32_t> cnt(nbins + 1, 0);
    for (uint32_t i = 0; i < n; ++i)
        if (bin[i] < nbins) ++cnt[bin[i] + 1];
    for (uint32_t b = 0; b < nbins; ++b) cnt[b + 1] += cnt[b];
    std::vector<uint32_t> idx(std::max<uint32_t>(1, cnt[nbins])), fill(cnt.begin(), cnt.end() - 1);
    std::vector<double> w(weight ? idx.size() : 0);
    for (uint32_t d = 0; d < n; ++d) {
        const uint32_t i = c->perm.to_ref(d);
        const uint32_t b = bin[i];
        if (b >= nbins) continue;
        if (weight) w[fill[b]] = weight[i];
        idx[fill[b]++] = d;
    }
    std::vector<uint32_t> cs(1, 0), cb(nbins + 1, 0);
    for (uint32_t b = 0; b < nbins; ++b) {
        cb[b] = (uint32_t)cs.size() - 1;
        for (uint32_t q = cnt[b]; q < cnt[b + 1]; q += kBinChunk) cs.push_back(std::min(cnt[b + 1], q + kBinChunk));
    }
    cb[nbins] = (uint32_t)cs.size() - 1;
    const uint32_t nch = (uint32_t)cs.size() - 1;
    auto up = [&](DevBuf& d, const void* h, size_t bytes) -> int {
        SLLG_TRY(d.alloc(std::max<size_t>(bytes, 8)));
        if (bytes) SLLG_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream));
        return 0;
    };
    SLLG_TRY(up(c->bin_idx, idx.data(), idx.size() * 4));
    if (weight)
        SLLG_TRY(up(c->bin_w, w.data(), w.size() * 8));
    else
        c->bin_w.release();
    SLLG_TRY(up(c->bin_cs, cs.data(), cs.size() * 4));
    SLLG_TRY(up(c->bin_cb, cb.data(), cb.size() * 4));
    SLLG_TRY(c->bin_part.alloc(std::max<size_t>(1, nch) * 3 * sizeof(double)));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    c->nbins = nbins;
    c->nbin_chunks = nch;
    return 0;
}

int reduce_bins(llg_ctx* c, double* dev_out) {
    SLLG_TRY(sync_spins(c));
    if (c->nbins == 0) return 0;
    const double* w = c->bin_w.p ? c->bin_w.as<double>() : nullptr;
    if (c->nbin_chunks) {
        if (c->prec == LLG_FP64_PLAIN)
            k_bin_partial<double><<<c->nbin_chunks, kTPB, 0, c->stream>>>(
                c->n, c->bin_cs.as<uint32_t>(), c->bin_idx.as<uint32_t>(), w, c->S.as<double>(), c->bin_part.as<double>());
        else
            k_bin_partial<float><<<c->nbin_chunks, kTPB, 0, c->stream>>>(
                c->n, c->bin_cs.as<uint32_t>(), c->bin_idx.as<uint32_t>(), w, c->S.as<float>(), c->bin_part.as<double>());
        SLLG_HIP(hipGetLastError());
    }
    k_bin_finish<<<blocks_for(c->nbins, kTPB), kTPB, 0, c->stream>>>(c->nbins, c->bin_cb.as<uint32_t>(),
                                                                     c->bin_part.as<double>(), dev_out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

int set_sq(llg_ctx* c, const uint32_t grid[3], const uint32_t* mesh, uint32_t nk, const uint32_t* kpts, int axis,
           const double* ua, const double* ub, uint32_t ngroups, const uint32_t* group) {
    const uint32_t n = c->n;
    std::vector<uint32_t> cnt(ngroups + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t g = group ? group[i] : 0u;
        if (g < ngroups) ++cnt[g + 1];
    }
    for (uint32_t g = 0; g < ngroups; ++g) cnt[g + 1] += cnt[g];
    std::vector<uint4> sp(std::max<uint32_t>(1, cnt[ngroups]));
    std::vector<double2> w(sp.size());
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (uint32_t d = 0; d < n; ++d) {
        const uint32_t i = c->perm.to_ref(d);
        const uint32_t g = group ? group[i] : 0u;
        if (g >= ngroups) continue;
        const uint32_t q = fill[g]++;
        sp[q] = make_uint4(d, mesh[3 * (size_t)i], mesh[3 * (size_t)i + 1], mesh[3 * (size_t)i + 2]);
        w[q] = make_double2(ua ? ua[i] : 1.0, ub ? ub[i] : 1.0);
    }
    std::vector<uint32_t> cs(1, 0), cg(ngroups + 1, 0);
    for (uint32_t g = 0; g < ngroups; ++g) {
        cg[g] = (uint32_t)cs.size() - 1;
        for (uint32_t q = cnt[g]; q < cnt[g + 1]; q += kSqChunk) cs.push_back(std::min(cnt[g + 1], q + kSqChunk));
    }
    cg[ngroups] = (uint32_t)cs.size() - 1;
    const uint32_t nch = (uint32_t)cs.size() - 1;
    auto up = [&](DevBuf& d, const void* h, size_t bytes) -> int {
        SLLG_TRY(d.alloc(std::max<size_t>(bytes, 8)));
        if (bytes) SLLG_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c->stream));
        return 0;
    };
    SLLG_TRY(up(c->sq_spin, sp.data(), sp.size() * sizeof(uint4)));
    SLLG_TRY(up(c->sq_w, w.data(), w.size() * sizeof(double2)));
    SLLG_TRY(up(c->sq_cs, cs.data(), cs.size() * 4));
    SLLG_TRY(up(c->sq_cg, cg.data(), cg.size() * 4));
    SLLG_TRY(up(c->sq_k, kpts, (size_t)nk * 3 * 4));
    SLLG_TRY(c->sq_part.alloc(std::max<size_t>(1, (size_t)nch * nk) * 2 * sizeof(double)));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    c->sq_nk = nk;
    c->sq_ngroups = ngroups;
    c->sq_nch = nch;
    c->sq_axis = axis;
    for (int d = 0; d < 3; ++d) c->sq_grid[d] = grid[d];
    return 0;
}

int sample_sq(llg_ctx* c, double* dev_out) {
    SLLG_TRY(sync_spins(c));
    if (c->sq_nk == 0 || c->sq_ngroups == 0) return 0;
    if (c->sq_nch) {
        SqArgs a{};
        a.n = c->n;
        a.nk = c->sq_nk;
        a.axis = (uint32_t)c->sq_axis;
        for (int d = 0; d < 3; ++d) a.inv[d] = 1.0 / (double)c->sq_grid[d];
        a.spin = c->sq_spin.as<uint4>();
        a.w = c->sq_w.as<double2>();
        a.cs = c->sq_cs.as<uint32_t>();
        a.kp = c->sq_k.as<uint32_t>();
        const dim3 grid(c->sq_nch, blocks_for(c->sq_nk, kTPB));
        if (c->prec == LLG_FP64_PLAIN)
            k_sq_partial<double><<<grid, kTPB, 0, c->stream>>>(a, c->S.as<double>(), c->sq_part.as<double>());
        else
            k_sq_partial<float><<<grid, kTPB, 0, c->stream>>>(a, c->S.as<float>(), c->sq_part.as<double>());
        SLLG_HIP(hipGetLastError());
    }
    k_sq_finish<<<blocks_for(c->sq_ngroups * c->sq_nk, kTPB), kTPB, 0, c->stream>>>(
        c->sq_ngroups, c->sq_nk, c->sq_cg.as<uint32_t>(), c->sq_part.as<double>(), dev_out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace sllg
