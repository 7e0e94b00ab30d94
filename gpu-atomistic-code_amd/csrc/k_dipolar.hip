// k_dipolar.hip - dipole-dipole field by zero-padded FFT convolution.
//
// Reference: kernel setup intmat::fillDipIntmat (src/intmat.cpp:263-331) + FP64 FFTW
// (src/intmat.cpp:373-398) cast to float (src/cuda_util.cu:149-178); per refresh
// CZero4DRSArrays, CdipCopySpin, cuFFT C2C forward, CdipFConv (9 complex MACs per
// k-point), cuFFT C2C inverse, CdipCopyFields (/zps) (src/cuda.cu:65-90,
// src/cufields.cu:294-358,409-462).
//
// Here: the padded grid G_a = 2 dim_a nk_a per axis; the kernel
//   N_ab(r) = 1e-7 mu_B (3 e_a e_b - delta_ab) / |r|^3,  r = d * abc/nk,
// with d_a = g_a for g_a < D_a and g_a - G_a for g_a > D_a (D_a = G_a/2), the origin
// excluded, is built in FP64 on the host.  The plane g_a = D_a never couples two spins
// (positions lie in [0, D_a)), so it is left zero; N is then even under r -> -r and its
// transform is real: 6 real components (xx xy xz yy yz zz) are kept after an FP64 D2Z
// transform on the device.  Each refresh: scatter mu_i S_i into the persistent real grid
// (empty points stay zero, so no per-refresh zero fill), R2C (3 batched transforms),
// H_k = N_k S_k (6 real x complex per k-point), C2R, gather / P.  FP32 path: FP32 grids
// and transforms with the FP64-built kernel cast to FP32; FP64 plain path: FP64 throughout.
//
// FFT exchange (the reference's exchm = 0, src/exch_intmat.cpp:28-160, src/intmat.cpp:29-262,
// applied per stage src/cuda.cu:47-60,162-175): one grid per species s and component, the
// kernel N(s1, s2, a, b, r) complex (exchange tensors need not be symmetric, so N is not
// even in r): J_ab / (mu_s1 mu_B) at r = -d for the interaction (s1, s2, d, J), plus the
// dipolar 1e-7 (3 e_a e_b - delta_ab) mu_s2 mu_B / r^3 when enabled.  Each evaluation:
// scatter S of every spin into its species' grids, R2C (3 nms transforms), H(s1, a) =
// sum_{s2, b} N(s1, s2, a, b) S(s2, b) per k-point, C2R, gather / P.
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "kernels.hpp"

namespace sllg {

struct DipolarPlan {
    bool f64 = false;
    int G[3] = {0, 0, 0};
    size_t P = 0, Ph = 0;
    DevBuf lin;    // u32 per spin: linear index in the padded grid
    DevBuf mu;     // per spin moment (mu_B), path precision
    std::shared_ptr<DevBuf> Nk;  // [6][Ph] real, path precision (shared by equal geometries)
    DevBuf grid;   // [3][P] real input grid
    DevBuf spec;   // [3][Ph] complex
    DevBuf hgrid;  // [3][P] real output grid
    hipfftHandle fwd = 0, inv = 0;
    bool have_fwd = false, have_inv = false;
    // FFT exchange (general species-resolved kernel)
    bool general = false;
    uint32_t nms = 1;
    DevBuf sp;     // u32 species of each spin (device order)
    DevBuf Nkc;    // [nms][nms][9][Ph] complex, path precision
    DevBuf hspec;  // [nms][3][Ph] complex
    // FP32 dipolar with a power-of-two Gx ("x half"): the input is zero for x >= Gx/2 and
    // only x < Gx/2 is read back, so hipFFT transforms the (y, z) planes of that half only
    // ([3][Gx/2][Gy][Gz] real <-> [3][Gx/2][Mc] complex, Mc = Gy (Gz/2 + 1) columns) and
    // k_dip_xfft does the x transforms and the k-space product in one pass.
    bool xh = false;
    size_t Pc = 0, Mc = 0;  // real points per component of the half grid; columns
    DevBuf tw;              // exp(-2 pi i m / Gx), m < Gx / 2 (FP32 pairs from FP64)
    std::shared_ptr<DevBuf> NkX;  // N(k) in the x kernel's block order [group][kx][6][COLS]
    DevBuf NkXc;                  // FFT exchange, half grid: [group][kx][nms^2 9][COLS] complex
    // half grid with the (y, z) plane kernels in place of hipFFT's 2D transforms and the
    // scatter / gather (power-of-two Gy, Gz at most 64): no real grids, no hipFFT plans
    bool yz = false, yz_dense = false;  // yz_dense: the map below is the identity
    DevBuf map;       // [species][x][y][z] over the physical extents: device spin or ~0
    DevBuf twy, twz;  // exp(-2 pi i m / Gy), m < Gy / 2; exp(-2 pi i m / Gz), m < Gz
    ~DipolarPlan() {
        if (have_fwd) hipfftDestroy(fwd);
        if (have_inv) hipfftDestroy(inv);
    }
};

// Process-wide cache of k-space kernels by (device, precision, grid, mesh, lattice constants).
struct KernelKey {
    int device;
    bool f64;
    int G[3];
    uint32_t nk[3];
    double abc[3];
    bool operator<(const KernelKey& o) const {
        return std::memcmp(this, &o, sizeof(KernelKey)) < 0;
    }
};
static std::mutex g_nk_mu;
static std::map<KernelKey, std::weak_ptr<DevBuf>> g_nk_cache;
static std::map<KernelKey, std::weak_ptr<DevBuf>> g_nkx_cache;  // the same kernels in k_dip_xfft's order

#define SLLG_FFT(call)                                                                       \
    do {                                                                                     \
        hipfftResult r_ = (call);                                                            \
        if (r_ != HIPFFT_SUCCESS) SLLG_FAIL(LLG_E_DEVICE, "%s failed: hipfft %d", #call, (int)r_); \
    } while (0)

template <class R>
__global__ __launch_bounds__(kTPB) void k_dip_scatter(uint32_t n, size_t P, const uint32_t* __restrict__ lin,
                                                      const R* __restrict__ mu, const R* __restrict__ S,
                                                      R* __restrict__ grid) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    const size_t g = lin[i];
    const R m = mu[i];
    grid[g] = m * S[i];
    grid[P + g] = m * S[(size_t)n + i];
    grid[2 * P + g] = m * S[2 * (size_t)n + i];
}

// H_a(k) = sum_b N_ab(k) S_b(k), N real symmetric: 6 comps (xx xy xz yy yz zz).
template <class R, class C>
__global__ __launch_bounds__(kTPB) void k_dip_mult(size_t Ph, const R* __restrict__ Nk, C* __restrict__ spec) {
    const size_t k = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (k >= Ph) return;
    const C sx = spec[k], sy = spec[Ph + k], sz = spec[2 * Ph + k];
    const R nxx = Nk[k], nxy = Nk[Ph + k], nxz = Nk[2 * Ph + k];
    const R nyy = Nk[3 * Ph + k], nyz = Nk[4 * Ph + k], nzz = Nk[5 * Ph + k];
    C hx, hy, hz;
    hx.x = nxx * sx.x + nxy * sy.x + nxz * sz.x;
    hx.y = nxx * sx.y + nxy * sy.y + nxz * sz.y;
    hy.x = nxy * sx.x + nyy * sy.x + nyz * sz.x;
    hy.y = nxy * sx.y + nyy * sy.y + nyz * sz.y;
    hz.x = nxz * sx.x + nyz * sy.x + nzz * sz.x;
    hz.y = nxz * sx.y + nyz * sy.y + nzz * sz.y;
    spec[k] = hx;
    spec[Ph + k] = hy;
    spec[2 * Ph + k] = hz;
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

constexpr int kXT = 512;   // threads of k_dip_xfft
constexpr int kYF = 512;   // threads of the forward (y, z) plane kernel
constexpr int kYI = 1024;  // threads of the inverse one (one block per plane: 1024 measured 13.4 us against
                           // 15.1 at 512 on the 256x256x32 grid; the forward's 768 blocks 16.5 against 15.8)

// (y, kz) columns per block of the x kernels: 3 nms COLS sequences of Gx, about 12.5 KB of
// LDS, so blocks outnumber the resident slots and one block's loads overlap another's LDS
// stages (2048 / Gx columns per block: 33 us, 1024 / Gx: 27 us, 512 / Gx: 25 us for the
// 256 x 256 x 32 grid); at least one column and whole (column, kx) pairs per thread.  Host
// permutes use the same function.
constexpr int xf_cols(int gx, int nms) { return 512 / (gx * nms) > 0 ? 512 / (gx * nms) : 1; }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 w, float2 v) {
    return make_float2(w.x * v.x - w.y * v.y, w.x * v.y + w.y * v.x);
}
// v conj(w)
__device__ __forceinline__ float2 cmulc(float2 w, float2 v) {
    return make_float2(w.x * v.x + w.y * v.y, w.x * v.y - w.y * v.x);
}

// Half-grid x transforms (FP32).  A block takes COLS (y, kz) columns of NS spectrum
// sequences each (3 components, times the species for FFT exchange), zero-extends x from
// Gx/2 to Gx, runs a radix-2 DIT forward FFT along x (bit-reversed load, natural-order
// result, two stages per LDS pass), forms the k-space product, runs a radix-2 DIF inverse
// FFT (natural-order input, bit-reversed result) and writes back x < Gx/2 only.
// Unnormalised both ways (the gather scales by 1/P).  NS COLS sequences of Gx, COLS from
// xf_cols: about 12.5 KB of LDS (more where one column is already that much).  A sequence's element i sits at i + i / 32 (one pad slot
// per 32): the bit-reversed load and store and the column-major product would otherwise
// hit one LDS bank pair.
template <int GX>
__device__ __forceinline__ uint32_t xf_at(uint32_t sq, uint32_t i) {
    return sq * (GX + GX / 32) + i + (i >> 5);
}

// neighbouring column groups share 128-byte lines of the spectrum: each XCD takes a
// contiguous range of groups (blocks are dealt to the 8 XCDs round-robin)
__device__ __forceinline__ uint32_t xf_group() {
    const uint32_t nb = gridDim.x, per = nb >> 3, rem = nb & 7u, xb = blockIdx.x & 7u;
    return xb * per + (xb < rem ? xb : rem) + (blockIdx.x >> 3);
}

// spectrum [sequence][x < Gx/2][M] -> buf (bit-reversed, zero for x >= Gx/2), twiddles -> w
template <int GX, int NS, int COLS>
__device__ __forceinline__ void xf_load(float2* buf, float2* w, const float2* __restrict__ tw,
                                        const float2* __restrict__ spec, uint32_t M, uint32_t j0) {
    constexpr int LX = GX / 2, LG = ilog2c(GX);
    for (uint32_t t = threadIdx.x; t < (uint32_t)(GX / 2); t += kXT) w[t] = tw[t];
    for (uint32_t e = threadIdx.x; e < (uint32_t)(NS * COLS * GX); e += kXT) {
        const uint32_t col = e % COLS, x = (e / COLS) % GX, sc = e / (COLS * GX), j = j0 + col;
        float2 v = make_float2(0.f, 0.f);
        if (x < (uint32_t)LX && j < M) v = spec[((size_t)sc * LX + x) * M + j];
        buf[xf_at<GX>(sc * COLS + col, __brev(x) >> (32 - LG))] = v;
    }
    __syncthreads();
}

// nseq sequences of GX elements (a compile-time constant at the x kernels' call sites)
template <int GX, int NT = kXT>
__device__ __forceinline__ void xf_single_stage(float2* buf, uint32_t nseq) {  // length 2: w[0] = 1
    for (uint32_t q = threadIdx.x; q < nseq * (GX / 2); q += NT) {
        const uint32_t sq = q / (GX / 2), i0 = xf_at<GX>(sq, 2 * (q % (GX / 2))), i1 = i0 + 1;
        const float2 u = buf[i0], v = buf[i1];
        buf[i0] = cadd(u, v);
        buf[i1] = csub(u, v);
    }
    __syncthreads();
}

// forward DIT: an odd stage count starts with one length-2 stage; then stages L and 2L run
// as one pass, a thread taking the quartet pos, pos + L/2, pos + L, pos + 3L/2 of a group
// of 2L (both butterflies of both stages in registers)
template <int GX, int NT = kXT>
__device__ __forceinline__ void xf_forward(float2* buf, const float2* w, uint32_t nseq) {
    constexpr int LG = ilog2c(GX);
    if constexpr (LG & 1) xf_single_stage<GX, NT>(buf, nseq);
#pragma unroll
    for (int L = (LG & 1) ? 4 : 2; 2 * L <= GX; L <<= 2) {
        const int qh = L >> 1;
        for (uint32_t q = threadIdx.x; q < nseq * (GX / 4); q += NT) {
            const uint32_t sq = q / (GX / 4), k = q % (GX / 4), pos = k % qh;
            const uint32_t b = (k / qh) * 2 * L + pos;
            const uint32_t i0 = xf_at<GX>(sq, b), i1 = xf_at<GX>(sq, b + qh), i2 = xf_at<GX>(sq, b + L),
                           i3 = xf_at<GX>(sq, b + L + qh);
            const float2 w1 = w[pos * (GX / L)], wa = w[pos * (GX / (2 * L))], wb = w[(pos + qh) * (GX / (2 * L))];
            const float2 a0 = buf[i0], a1 = buf[i1], a2 = buf[i2], a3 = buf[i3];
            float2 t = cmul(w1, a1);
            const float2 b0 = cadd(a0, t), b1 = csub(a0, t);
            t = cmul(w1, a3);
            const float2 b2 = cadd(a2, t), b3 = csub(a2, t);
            t = cmul(wa, b2);
            buf[i0] = cadd(b0, t);
            buf[i2] = csub(b0, t);
            t = cmul(wb, b3);
            buf[i1] = cadd(b1, t);
            buf[i3] = csub(b1, t);
        }
        __syncthreads();
    }
}

// inverse DIF (twiddles conjugated): stages 2L and L as one pass from the top; an odd stage
// count ends with one length-2 stage
template <int GX, int NT = kXT>
__device__ __forceinline__ void xf_inverse(float2* buf, const float2* w, uint32_t nseq) {
    constexpr int LG = ilog2c(GX);
#pragma unroll
    for (int L = GX / 2; L >= ((LG & 1) ? 4 : 2); L >>= 2) {
        const int qh = L >> 1;
        for (uint32_t q = threadIdx.x; q < nseq * (GX / 4); q += NT) {
            const uint32_t sq = q / (GX / 4), k = q % (GX / 4), pos = k % qh;
            const uint32_t b = (k / qh) * 2 * L + pos;
            const uint32_t i0 = xf_at<GX>(sq, b), i1 = xf_at<GX>(sq, b + qh), i2 = xf_at<GX>(sq, b + L),
                           i3 = xf_at<GX>(sq, b + L + qh);
            const float2 w1 = w[pos * (GX / L)], wa = w[pos * (GX / (2 * L))], wb = w[(pos + qh) * (GX / (2 * L))];
            const float2 a0 = buf[i0], a1 = buf[i1], a2 = buf[i2], a3 = buf[i3];
            const float2 b0 = cadd(a0, a2), b2 = cmulc(wa, csub(a0, a2));
            const float2 b1 = cadd(a1, a3), b3 = cmulc(wb, csub(a1, a3));
            buf[i0] = cadd(b0, b1);
            buf[i1] = cmulc(w1, csub(b0, b1));
            buf[i2] = cadd(b2, b3);
            buf[i3] = cmulc(w1, csub(b2, b3));
        }
        __syncthreads();
    }
    if constexpr (LG & 1) xf_single_stage<GX, NT>(buf, nseq);
}

template <int GX, int NS, int COLS>
__device__ __forceinline__ void xf_store(const float2* buf, float2* __restrict__ spec, uint32_t M, uint32_t j0) {
    constexpr int LX = GX / 2, LG = ilog2c(GX);
    for (uint32_t e = threadIdx.x; e < (uint32_t)(NS * COLS * LX); e += kXT) {
        const uint32_t col = e % COLS, x = (e / COLS) % LX, sc = e / (COLS * LX), j = j0 + col;
        if (j < M) spec[((size_t)sc * LX + x) * M + j] = buf[xf_at<GX>(sc * COLS + col, __brev(x) >> (32 - LG))];
    }
}

// (y, z) plane transforms of the half grid, fused with the scatter and the gather.  A block
// takes one plane: sequence sc = species * 3 + component, x < Gx/2.  Forward: the plane's
// occupied points y < Gy/2, z < Gz/2 (mesh point -> device spin from `map`, times the
// moment for the dipolar convolution) into LDS, a direct DFT along z to kz = 0 .. Gz/2 (the
// zero half of z costs nothing), a radix-2 FFT along y over the Gz/2 + 1 kz sequences, and
// the [ky][kz] half spectrum out in hipFFT's R2C layout.  Inverse: the spectrum in, the
// inverse y FFT, and at each occupied point the real inverse DFT along z (Hermitian half
// spectrum: X0 + X_{Gz/2} (-1)^z + 2 Re sum X_kz e^{+i...}) written to the spin's field.
// Dynamic LDS: (Gz/2 + 1) sequences of Gy (+1/32) complex, the plane's real points, the y
// and z twiddles.
struct YZArgs {
    uint32_t n, lx, ly, lz, gz, kz;  // spins; physical extents (G/2); Gz; Gz/2 + 1
    const uint32_t* map;             // [species][x][y][z] over the physical extents: device spin or ~0
    uint32_t dense;                  // the map is the identity (one species, every point, mesh order)
    const float* mu;                 // moments (dipolar) or null (FFT exchange)
    const float2* twy;               // exp(-2 pi i m / Gy), m < Gy / 2
    const float2* twz;               // exp(-2 pi i m / Gz), m < Gz
    float2* spec;                    // [sc][x < Gx/2][Gy (Gz/2 + 1)]
};

// The forward pass splits a plane's kz range over `parts` blocks (kzp sequences each, none
// empty; the plane's spins are read by each part) for parallelism: the plane count alone is
// small. The parts of a plane write interleaved pieces of the same 128-byte lines, so they
// are kept on one XCD (one L2, where the pieces merge before write-back): workgroups are
// dealt to the 8 XCDs round-robin by id, so within each full group of 8 * parts ids, id b
// runs on XCD b % 8 and takes part (b / 8) % parts of plane (b / (8 parts)) * 8 + b % 8;
// ids past the last full group map plainly. The map is a bijection onto (plane, part).
template <int GY>
__global__ __launch_bounds__(kYF) void k_yz_fwd(YZArgs a, const float* __restrict__ S, uint32_t parts, uint32_t kzp) {
    constexpr int LG = ilog2c(GY);
    extern __shared__ float2 yz_lds[];
    const uint32_t b = blockIdx.x, full = gridDim.x - gridDim.x % (8u * parts);
    const uint32_t plane = b < full ? (b / 8u / parts) * 8u + b % 8u : b / parts;
    const uint32_t part = b < full ? (b / 8u) % parts : b % parts;
    const uint32_t kz = a.kz;
    const uint32_t sc = plane / a.lx, x = plane % a.lx, sp = sc / 3, c = sc % 3;
    const uint32_t k0 = part * kzp, nk = (k0 + kzp <= kz ? kzp : kz - k0);  // this block's kz range
    float2* const buf = yz_lds;                                        // kzp (GY + GY/32)
    float2* const wy = buf + (size_t)kzp * (GY + GY / 32);              // GY / 2
    float2* const wz = wy + GY / 2;                                     // gz
    float* const R = reinterpret_cast<float*>(wz + a.gz);               // ly lz
    for (uint32_t t = threadIdx.x; t < (uint32_t)(GY / 2); t += kYF) wy[t] = a.twy[t];
    for (uint32_t t = threadIdx.x; t < a.gz; t += kYF) wz[t] = a.twz[t];
    const uint32_t* const mp = a.map + ((size_t)sp * a.lx + x) * a.ly * a.lz;
    const uint32_t i0 = x * a.ly * a.lz;  // first spin of the plane when the map is the identity
    for (uint32_t e = threadIdx.x; e < a.ly * a.lz; e += kYF) {
        const uint32_t i = a.dense ? i0 + e : mp[e];
        R[e] = i == 0xFFFFFFFFu ? 0.f : (a.mu ? a.mu[i] : 1.f) * S[(size_t)c * a.n + i];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < (uint32_t)GY * nk; e += kYF) {
        const uint32_t y = e / nk, kk = e % nk, k = k0 + kk;
        float2 X = make_float2(0.f, 0.f);
        if (y < a.ly) {
            const float* const r = R + (size_t)y * a.lz;
            uint32_t m = 0;  // k z mod gz
            for (uint32_t z = 0; z < a.lz; ++z) {
                const float2 t = wz[m];
                X.x += r[z] * t.x;
                X.y += r[z] * t.y;
                m += k;
                if (m >= a.gz) m -= a.gz;
            }
        }
        buf[xf_at<GY>(kk, __brev(y) >> (32 - LG))] = X;
    }
    __syncthreads();
    xf_forward<GY, kYF>(buf, wy, nk);
    float2* const out = a.spec + ((size_t)sc * a.lx + x) * GY * kz + k0;
    for (uint32_t e = threadIdx.x; e < (uint32_t)GY * nk; e += kYF)
        out[(size_t)(e / nk) * kz + e % nk] = buf[xf_at<GY>(e % nk, e / nk)];
}

template <int GY>
__global__ __launch_bounds__(kYI) void k_yz_inv(YZArgs a, float invP, float* __restrict__ out) {
    constexpr int LG = ilog2c(GY);
    extern __shared__ float2 yz_lds[];
    const uint32_t kz = a.kz, sc = blockIdx.x / a.lx, x = blockIdx.x % a.lx, sp = sc / 3, c = sc % 3;
    float2* const buf = yz_lds;
    float2* const wy = buf + (size_t)kz * (GY + GY / 32);
    float2* const wz = wy + GY / 2;
    for (uint32_t t = threadIdx.x; t < (uint32_t)(GY / 2); t += kYI) wy[t] = a.twy[t];
    for (uint32_t t = threadIdx.x; t < a.gz; t += kYI) wz[t] = a.twz[t];
    const float2* const in = a.spec + ((size_t)sc * a.lx + x) * GY * kz;
    for (uint32_t e = threadIdx.x; e < (uint32_t)GY * kz; e += kYI) buf[xf_at<GY>(e % kz, e / kz)] = in[e];
    __syncthreads();
    xf_inverse<GY, kYI>(buf, wy, kz);
    const uint32_t* const mp = a.map + ((size_t)sp * a.lx + x) * a.ly * a.lz;
    const uint32_t nyq = a.gz / 2, i0 = x * a.ly * a.lz;
    for (uint32_t e = threadIdx.x; e < a.ly * a.lz; e += kYI) {
        const uint32_t i = a.dense ? i0 + e : mp[e];
        if (i == 0xFFFFFFFFu) continue;
        const uint32_t y = e / a.lz, z = e % a.lz, yb = __brev(y) >> (32 - LG);
        float v = buf[xf_at<GY>(0, yb)].x + ((z & 1u) ? -1.f : 1.f) * buf[xf_at<GY>(nyq, yb)].x;
        float acc = 0.f;
        uint32_t m = z;  // k z mod gz, from k = 1
        for (uint32_t k = 1; k < nyq; ++k) {
            const float2 X = buf[xf_at<GY>(k, yb)], t = wz[m];  // Re(X conj(t))
            acc += X.x * t.x + X.y * t.y;
            m += z;
            if (m >= a.gz) m -= a.gz;
        }
        out[(size_t)c * a.n + i] = (v + 2.f * acc) * invP;
    }
}

// nseq kz sequences (gz / 2 + 1 for the inverse, a forward part's kzp)
static size_t yz_lds_bytes(int gy, uint32_t gz, uint32_t ly, uint32_t lz, uint32_t nseq) {
    return ((size_t)nseq * (gy + gy / 32) + gy / 2 + gz) * sizeof(float2) + (size_t)ly * lz * sizeof(float);
}

static int launch_yz(bool fwd, int gy, uint32_t planes, const YZArgs& a, const float* S, float invP, float* out,
                     hipStream_t s) {
    // forward: two kz parts (kz >= 2, so neither is empty; four parts measured slower, the
    // plane's spins being read once per part)
    const uint32_t parts = fwd ? 2u : 1u;
    const uint32_t kzp = (a.kz + parts - 1) / parts;
    const size_t lds = yz_lds_bytes(gy, a.gz, a.ly, a.lz, kzp);
#define SLLG_YZ(G_)                                                                   \
    case G_:                                                                          \
        if (fwd) k_yz_fwd<G_><<<planes * parts, kYF, lds, s>>>(a, S, parts, kzp);    \
        else k_yz_inv<G_><<<planes, kYI, lds, s>>>(a, invP, out);                     \
        break;
    switch (gy) {
        SLLG_YZ(16) SLLG_YZ(32) SLLG_YZ(64) SLLG_YZ(128) SLLG_YZ(256) SLLG_YZ(512) SLLG_YZ(1024)
        default: SLLG_FAIL(LLG_E_UNSUPPORTED, "y extent %d has no plane transform", gy);
    }
#undef SLLG_YZ
    SLLG_HIP(hipGetLastError());
    return 0;
}

// Dipolar: 3 sequences per column, product with the real symmetric N(k) (6 comps, stored
// per block: [group][kx][comp][COLS]); each thread's kernel values are loaded first so
// their latency overlaps the spectrum loads and the forward stages.
template <int GX>
__global__ __launch_bounds__(kXT) void k_dip_xfft(uint32_t M, const float2* __restrict__ tw,
                                                  const float* __restrict__ NkX, float2* __restrict__ spec) {
    constexpr int COLS = xf_cols(GX, 1), LG = ilog2c(GX), NPT = COLS * GX / kXT;
    static_assert(COLS >= 1 && (1 << LG) == GX, "power-of-two x extent of at most 2048");
    static_assert((COLS * GX) % kXT == 0, "whole (column, kx) pairs per thread");
    __shared__ float2 buf[3 * COLS * (GX + GX / 32)];
    __shared__ float2 w[GX / 2];
    const uint32_t grp = xf_group(), j0 = grp * COLS;
    const float* const ng = NkX + (size_t)grp * GX * 6 * COLS;  // padding columns hold zeros
    float nreg[NPT][6];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const uint32_t e = threadIdx.x + r * kXT, col = e % COLS, kx = e / COLS;
#pragma unroll
        for (int q = 0; q < 6; ++q) nreg[r][q] = ng[(size_t)kx * 6 * COLS + q * COLS + col];
    }
    xf_load<GX, 3, COLS>(buf, w, tw, spec, M, j0);
    xf_forward<GX>(buf, w, 3 * COLS);
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const uint32_t e = threadIdx.x + r * kXT, col = e % COLS, kx = e / COLS, j = j0 + col;
        if (j >= M) continue;
        const float nxx = nreg[r][0], nxy = nreg[r][1], nxz = nreg[r][2];
        const float nyy = nreg[r][3], nyz = nreg[r][4], nzz = nreg[r][5];
        float2* const bx = buf + xf_at<GX>(0 * COLS + col, kx);
        float2* const by = buf + xf_at<GX>(1 * COLS + col, kx);
        float2* const bz = buf + xf_at<GX>(2 * COLS + col, kx);
        const float2 sx = *bx, sy = *by, sz = *bz;
        *bx = make_float2(nxx * sx.x + nxy * sy.x + nxz * sz.x, nxx * sx.y + nxy * sy.y + nxz * sz.y);
        *by = make_float2(nxy * sx.x + nyy * sy.x + nyz * sz.x, nxy * sx.y + nyy * sy.y + nyz * sz.y);
        *bz = make_float2(nxz * sx.x + nyz * sy.x + nzz * sz.x, nxz * sx.y + nyz * sy.y + nzz * sz.y);
    }
    __syncthreads();
    xf_inverse<GX>(buf, w, 3 * COLS);
    xf_store<GX, 3, COLS>(buf, spec, M, j0);
}

// FFT exchange: 3 NMS sequences per column ([species][component]), product with the
// complex kernel, H(s1, a) = sum_{s2, b} N(s1, s2, a, b) S(s2, b), N stored per block:
// [group][kx][(s1 NMS + s2) 9 + 3 a + b][COLS].
template <int GX, int NMS>
__global__ __launch_bounds__(kXT) void k_fx_xfft(uint32_t M, const float2* __restrict__ tw,
                                                 const float2* __restrict__ NkX, float2* __restrict__ spec) {
    constexpr int COLS = xf_cols(GX, NMS), LG = ilog2c(GX), NN = NMS * NMS * 9;
    static_assert(COLS >= 1 && (1 << LG) == GX, "power-of-two x extent, GX NMS at most 2048");
    __shared__ float2 buf[3 * NMS * COLS * (GX + GX / 32)];
    __shared__ float2 w[GX / 2];
    const uint32_t grp = xf_group(), j0 = grp * COLS;
    const float2* const ng = NkX + (size_t)grp * GX * NN * COLS;
    xf_load<GX, 3 * NMS, COLS>(buf, w, tw, spec, M, j0);
    xf_forward<GX>(buf, w, 3 * NMS * COLS);
    for (uint32_t e = threadIdx.x; e < (uint32_t)(COLS * GX); e += kXT) {
        const uint32_t col = e % COLS, kx = e / COLS, j = j0 + col;
        if (j >= M) continue;
        float2 sv[3 * NMS];
#pragma unroll
        for (int q = 0; q < 3 * NMS; ++q) sv[q] = buf[xf_at<GX>(q * COLS + col, kx)];
        const float2* const nv = ng + (size_t)kx * NN * COLS + col;
#pragma unroll
        for (int s1 = 0; s1 < NMS; ++s1)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float2 h = make_float2(0.f, 0.f);
#pragma unroll
                for (int s2 = 0; s2 < NMS; ++s2)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        h = cadd(h, cmul(nv[(size_t)((s1 * NMS + s2) * 9 + a * 3 + b) * COLS], sv[s2 * 3 + b]));
                buf[xf_at<GX>((s1 * 3 + a) * COLS + col, kx)] = h;
            }
    }
    __syncthreads();
    xf_inverse<GX>(buf, w, 3 * NMS * COLS);
    xf_store<GX, 3 * NMS, COLS>(buf, spec, M, j0);
}

// N(k) from [6][kx M + j] to [j / COLS][kx][6][j % COLS] (zero past the last column).
__global__ __launch_bounds__(kTPB) void k_nk_to_x(uint32_t M, uint32_t gx, uint32_t cols, size_t total,
                                                  const float* __restrict__ Nk, float* __restrict__ NkX) {
    const size_t e = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (e >= total) return;
    const uint32_t col = (uint32_t)(e % cols);
    const uint32_t p = (uint32_t)((e / cols) % 6);
    const uint32_t kx = (uint32_t)((e / ((size_t)cols * 6)) % gx);
    const size_t g = e / ((size_t)cols * 6 * gx);
    const size_t j = g * cols + col;
    NkX[e] = j < M ? Nk[(size_t)p * gx * M + (size_t)kx * M + j] : 0.f;
}


// FFT-exchange kernel from [nn][kx M + j] (complex) to [j / COLS][kx][nn][j % COLS].
__global__ __launch_bounds__(kTPB) void k_nkc_to_x(uint32_t M, uint32_t gx, uint32_t cols, uint32_t nn, size_t total,
                                                   const float2* __restrict__ Nk, float2* __restrict__ NkX) {
    const size_t e = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (e >= total) return;
    const uint32_t col = (uint32_t)(e % cols);
    const uint32_t p = (uint32_t)((e / cols) % nn);
    const uint32_t kx = (uint32_t)((e / ((size_t)cols * nn)) % gx);
    const size_t g = e / ((size_t)cols * nn * gx);
    const size_t j = g * cols + col;
    NkX[e] = j < M ? Nk[(size_t)p * gx * M + (size_t)kx * M + j] : make_float2(0.f, 0.f);
}

static int launch_dip_xfft(int gx, uint32_t M, const float2* tw, const float* Nk, float2* spec, hipStream_t s) {
#define SLLG_XF(G_)                                                                          \
    case G_:                                                                                 \
        k_dip_xfft<G_><<<(unsigned)((M + xf_cols(G_, 1) - 1) / xf_cols(G_, 1)), kXT, 0, s>>>(M, tw, Nk, spec); \
        break;
    switch (gx) {
        SLLG_XF(16) SLLG_XF(32) SLLG_XF(64) SLLG_XF(128) SLLG_XF(256) SLLG_XF(512) SLLG_XF(1024) SLLG_XF(2048)
        default: SLLG_FAIL(LLG_E_UNSUPPORTED, "x extent %d has no half-grid transform", gx);
    }
#undef SLLG_XF
    SLLG_HIP(hipGetLastError());
    return 0;
}

static int launch_fx_xfft(int gx, uint32_t nms, uint32_t M, const float2* tw, const float2* Nk, float2* spec,
                          hipStream_t s) {
#define SLLG_XF(G_, S_)                                                                              \
    case G_:                                                                                         \
        k_fx_xfft<G_, S_><<<(unsigned)((M + xf_cols(G_, S_) - 1) / xf_cols(G_, S_)), kXT, 0, s>>>(M, tw, Nk, \
                                                                                                   spec); \
        break;
    if (nms == 1) {
        switch (gx) {
            SLLG_XF(16, 1) SLLG_XF(32, 1) SLLG_XF(64, 1) SLLG_XF(128, 1) SLLG_XF(256, 1) SLLG_XF(512, 1)
            SLLG_XF(1024, 1) SLLG_XF(2048, 1)
            default: SLLG_FAIL(LLG_E_UNSUPPORTED, "x extent %d has no half-grid transform", gx);
        }
    } else if (nms == 2) {
        switch (gx) {
            SLLG_XF(16, 2) SLLG_XF(32, 2) SLLG_XF(64, 2) SLLG_XF(128, 2) SLLG_XF(256, 2) SLLG_XF(512, 2)
            SLLG_XF(1024, 2)
            default: SLLG_FAIL(LLG_E_UNSUPPORTED, "x extent %d has no half-grid transform", gx);
        }
    } else {
        SLLG_FAIL(LLG_E_UNSUPPORTED, "half-grid FFT exchange with %u species", nms);
    }
#undef SLLG_XF
    SLLG_HIP(hipGetLastError());
    return 0;
}

// The half-grid path applies to FP32 plans with a power-of-two x extent (and, for FFT
// exchange, one or two species) unless the caller asks for the whole-grid transforms.
static bool half_grid_ok(const llg_ctx* c, const DipolarPlan* p, uint32_t nms) {
    const int gx = p->G[0];
    return !p->f64 && gx >= 16 && (gx & (gx - 1)) == 0 && nms >= 1 && nms <= 2 && gx * (int)nms <= 2048 &&
           !(c->flags & LLG_FLAG_DIPOLAR_FULL_FFT);
}

// exp(-2 pi i m / Gx), m < Gx / 2, FP32 pairs from FP64
static int upload_twiddles(DipolarPlan* p) {
    const int gx = p->G[0];
    std::vector<float2> tw(gx / 2);
    for (int m = 0; m < gx / 2; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)gx;
        tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    SLLG_TRY(p->tw.alloc(tw.size() * sizeof(float2)));
    SLLG_HIP(hipMemcpy(p->tw.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    return 0;
}

static int upload_tw(DevBuf& b, int g, int count) {
    std::vector<float2> tw(count);
    for (int m = 0; m < count; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)g;
        tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    SLLG_TRY(b.alloc(tw.size() * sizeof(float2)));
    SLLG_HIP(hipMemcpy(b.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    return 0;
}

// The (y, z) plane kernels apply on the half grid when Gy is a power of two up to 1024, Gz
// is at most 64 (direct DFT along z) and a plane fits the LDS.
static bool yz_ok(const DipolarPlan* p) {
    const int gy = p->G[1], gz = p->G[2];
    return p->xh && gy >= 16 && gy <= 1024 && (gy & (gy - 1)) == 0 && gz <= 64 &&
           yz_lds_bytes(gy, (uint32_t)gz, (uint32_t)gy / 2, (uint32_t)gz / 2, (uint32_t)gz / 2 + 1) <= 60 * 1024;
}

// mesh point -> device spin per species, and the y / z twiddles
static int yz_setup(llg_ctx* c, DipolarPlan* p, const llg_dipolar& d, const uint32_t* spin_species, uint32_t nms) {
    const size_t lx = p->G[0] / 2, ly = p->G[1] / 2, lz = p->G[2] / 2;
    std::vector<uint32_t> map((size_t)nms * lx * ly * lz, 0xFFFFFFFFu);
    for (uint32_t i = 0; i < c->n; ++i) {
        const size_t m0 = d.mesh[(size_t)i * 3], m1 = d.mesh[(size_t)i * 3 + 1], m2 = d.mesh[(size_t)i * 3 + 2];
        const uint32_t sp = nms > 1 ? (spin_species ? spin_species[i] : (i % c->ns)) : 0;
        map[(((size_t)sp * lx + m0) * ly + m1) * lz + m2] = c->perm.to_dev(i);
    }
    // identity map: one species, every mesh point holds the spin of its own linear index
    bool dense = nms == 1 && (size_t)c->n == map.size();
    for (size_t e = 0; dense && e < map.size(); ++e) dense = map[e] == (uint32_t)e;
    p->yz_dense = dense;
    SLLG_TRY(p->map.alloc(map.size() * sizeof(uint32_t)));
    SLLG_HIP(hipMemcpy(p->map.p, map.data(), map.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    SLLG_TRY(upload_tw(p->twy, p->G[1], p->G[1] / 2));
    SLLG_TRY(upload_tw(p->twz, p->G[2], p->G[2]));
    p->yz = true;
    return 0;
}

static YZArgs yz_args(const llg_ctx* c, const DipolarPlan* p) {
    YZArgs a;
    a.n = c->n;
    a.lx = (uint32_t)p->G[0] / 2;
    a.ly = (uint32_t)p->G[1] / 2;
    a.lz = (uint32_t)p->G[2] / 2;
    a.gz = (uint32_t)p->G[2];
    a.kz = a.gz / 2 + 1;
    a.map = p->map.as<uint32_t>();
    a.dense = p->yz_dense ? 1u : 0u;
    a.mu = p->general ? nullptr : p->mu.as<float>();
    a.twy = p->twy.as<float2>();
    a.twz = p->twz.as<float2>();
    a.spec = p->spec.as<float2>();
    return a;
}

template <class R>
__global__ __launch_bounds__(kTPB) void k_dip_gather(uint32_t n, size_t P, R invP, const uint32_t* __restrict__ lin,
                                                     const R* __restrict__ hgrid, R* __restrict__ out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    const size_t g = lin[i];
    out[i] = hgrid[g] * invP;
    out[(size_t)n + i] = hgrid[P + g] * invP;
    out[2 * (size_t)n + i] = hgrid[2 * P + g] * invP;
}

// FFT exchange: S of spin i into the grids of its species, [s][c][P].
template <class R>
__global__ __launch_bounds__(kTPB) void k_fx_scatter(uint32_t n, size_t P, const uint32_t* __restrict__ lin,
                                                     const uint32_t* __restrict__ sp, const R* __restrict__ S,
                                                     R* __restrict__ grid) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    R* const gs = grid + (size_t)sp[i] * 3 * P;
    const size_t g = lin[i];
    gs[g] = S[i];
    gs[P + g] = S[(size_t)n + i];
    gs[2 * P + g] = S[2 * (size_t)n + i];
}

// H(s1, a)(k) = sum_{s2, b} N(s1, s2, a, b)(k) S(s2, b)(k), complex.
template <class R, class C>
__global__ __launch_bounds__(kTPB) void k_fx_mult(size_t Ph, uint32_t nms, const C* __restrict__ Nk,
                                                  const C* __restrict__ spec, C* __restrict__ hspec) {
    const size_t k = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (k >= Ph) return;
    for (uint32_t s1 = 0; s1 < nms; ++s1)
        for (int a = 0; a < 3; ++a) {
            R hr = 0, hi = 0;
            for (uint32_t s2 = 0; s2 < nms; ++s2)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const C nv = Nk[(((size_t)s1 * nms + s2) * 9 + a * 3 + b) * Ph + k];
                    const C sv = spec[((size_t)s2 * 3 + b) * Ph + k];
                    hr += nv.x * sv.x - nv.y * sv.y;
                    hi += nv.x * sv.y + nv.y * sv.x;
                }
            C h;
            h.x = hr;
            h.y = hi;
            hspec[((size_t)s1 * 3 + a) * Ph + k] = h;
        }
}

template <class R>
__global__ __launch_bounds__(kTPB) void k_fx_gather(uint32_t n, size_t P, R invP, const uint32_t* __restrict__ lin,
                                                    const uint32_t* __restrict__ sp, const R* __restrict__ hgrid,
                                                    R* __restrict__ out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    const R* const hs = hgrid + (size_t)sp[i] * 3 * P;
    const size_t g = lin[i];
    out[i] = hs[g] * invP;
    out[(size_t)n + i] = hs[P + g] * invP;
    out[2 * (size_t)n + i] = hs[2 * P + g] * invP;
}

template <class C>
__global__ __launch_bounds__(kTPB) void k_take_complex(size_t m, const hipfftDoubleComplex* __restrict__ in,
                                                       C* __restrict__ out) {
    const size_t k = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (k >= m) return;
    C v;
    v.x = in[k].x;
    v.y = in[k].y;
    out[k] = v;
}

template <class R>
__global__ __launch_bounds__(kTPB) void k_take_real(size_t m, const hipfftDoubleComplex* __restrict__ in, R* __restrict__ out) {
    const size_t k = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (k >= m) return;
    out[k] = (R)in[k].x;
}

// FP64 real-space kernel on the padded grid, component-major [6][P].
static void fill_kernel(const int G[3], const uint32_t dim[3], const uint32_t nk[3], const double abc[3],
                        std::vector<double>& N) {
    const size_t P = (size_t)G[0] * G[1] * G[2];
    N.assign(6 * P, 0.0);
    const int D[3] = {G[0] / 2, G[1] / 2, G[2] / 2};
    const int pairs[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int i = 0; i < G[0]; ++i)
        for (int j = 0; j < G[1]; ++j)
            for (int k = 0; k < G[2]; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                const int g[3] = {i, j, k};
                bool skip = false;
                double r[3];
                for (int a = 0; a < 3; ++a) {
                    if (g[a] == D[a] && D[a] > 0) { skip = true; break; }
                    const int d = g[a] < D[a] ? g[a] : g[a] - G[a];
                    r[a] = static_cast<double>(d) * abc[a] / static_cast<double>(nk[a]);
                }
                if (skip) continue;
                const double mr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const double oo = 1. / (mr * mr * mr);
                const double e[3] = {r[0] / mr, r[1] / mr, r[2] / mr};
                const size_t lin = ((size_t)i * G[1] + j) * G[2] + k;
                for (int p = 0; p < 6; ++p) {
                    const int a = pairs[p][0], b = pairs[p][1];
                    N[p * P + lin] = 1e-7 * ((3.0 * e[a] * e[b]) - (a == b ? 1.0 : 0.0)) * oo * LLG_MU_B;
                }
            }
    (void)dim;
}

// FP64 FFT-exchange kernel on the padded grid, [(s1 nms + s2) 9 + 3 a + b][P].
static void fill_general(const int G[3], const uint32_t nk[3], const double abc[3], uint32_t nms,
                         const std::vector<double>& mu_sp, const llg_fftexch& fx, bool dip, std::vector<double>& N) {
    const size_t P = (size_t)G[0] * G[1] * G[2];
    N.assign((size_t)nms * nms * 9 * P, 0.0);
    const int D[3] = {G[0] / 2, G[1] / 2, G[2] / 2};
    auto at = [&](uint32_t s1, uint32_t s2, int ab, size_t lin) -> double& {
        return N[(((size_t)s1 * nms + s2) * 9 + ab) * P + lin];
    };
    // exchange: neighbour at mesh(i) + d, so the convolution index x - y is -d; an offset
    // of D_a or more along an axis never joins two spins of the system
    for (uint32_t q = 0; q < fx.n_int; ++q) {
        const llg_fft_interaction& it = fx.ints[q];
        int g[3];
        bool far = false;
        for (int a = 0; a < 3; ++a) {
            if (std::abs(it.d[a]) >= D[a]) far = true;
            g[a] = ((-it.d[a]) % G[a] + G[a]) % G[a];
        }
        if (far) continue;
        const size_t lin = ((size_t)g[0] * G[1] + g[1]) * G[2] + g[2];
        const double inv = 1.0 / (mu_sp[it.s1] * LLG_MU_B);
        for (int ab = 0; ab < 9; ++ab) at(it.s1, it.s2, ab, lin) += it.J[ab] * inv;
    }
    if (!dip) return;
    for (int i = 0; i < G[0]; ++i)
        for (int j = 0; j < G[1]; ++j)
            for (int k = 0; k < G[2]; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                const int g[3] = {i, j, k};
                bool skip = false;
                double r[3];
                for (int a = 0; a < 3; ++a) {
                    if (g[a] == D[a] && D[a] > 0) { skip = true; break; }
                    const int dd = g[a] < D[a] ? g[a] : g[a] - G[a];
                    r[a] = static_cast<double>(dd) * abc[a] / static_cast<double>(nk[a]);
                }
                if (skip) continue;
                const double mr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                const double oo = 1. / (mr * mr * mr);
                const double e[3] = {r[0] / mr, r[1] / mr, r[2] / mr};
                const size_t lin = ((size_t)i * G[1] + j) * G[2] + k;
                for (uint32_t s2 = 0; s2 < nms; ++s2) {
                    const double m2 = mu_sp[s2] * LLG_MU_B;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            const double v = 1e-7 * ((3.0 * e[a] * e[b]) - (a == b ? 1.0 : 0.0)) * oo * m2;
                            for (uint32_t s1 = 0; s1 < nms; ++s1) at(s1, s2, a * 3 + b, lin) += v;
                        }
                }
            }
}

// FFT exchange plan: species grids, general complex kernel, transforms.
static int fftexch_create(llg_ctx* c, DipolarPlan* p, const llg_dipolar& d, const uint32_t* spin_species,
                          const llg_fftexch& fx, bool dip) {
    p->general = true;
    p->nms = c->ns;
    const uint32_t n = c->n, nms = p->nms;
    std::vector<uint32_t> lin(n), sp(n);
    std::vector<uint8_t> used((size_t)nms * p->P, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t m[3];
        for (int a = 0; a < 3; ++a) {
            m[a] = d.mesh[(size_t)i * 3 + a];
            if (m[a] >= d.dim[a] * d.nk[a]) SLLG_FAIL(LLG_E_INVALID, "mesh coordinate of spin %u outside dim*nk", i);
        }
        const uint32_t s = spin_species ? spin_species[i] : (i % c->ns);
        const size_t g = ((size_t)m[0] * p->G[1] + m[1]) * p->G[2] + m[2];
        if (used[(size_t)s * p->P + g]) SLLG_FAIL(LLG_E_INVALID, "two spins of species %u share a mesh point (spin %u)", s, i);
        used[(size_t)s * p->P + g] = 1;
        lin[c->perm.to_dev(i)] = (uint32_t)g;
        sp[c->perm.to_dev(i)] = s;
    }
    SLLG_TRY(p->lin.alloc(n * sizeof(uint32_t)));
    SLLG_HIP(hipMemcpy(p->lin.p, lin.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    SLLG_TRY(p->sp.alloc(n * sizeof(uint32_t)));
    SLLG_HIP(hipMemcpy(p->sp.p, sp.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    // kernel: FP64 real space -> FP64 D2Z on the device -> path precision complex
    std::vector<double> N;
    fill_general(p->G, d.nk, d.abc, nms, c->mu_sp, fx, dip, N);
    const int howmany = (int)(nms * nms * 9);
    const size_t rs = p->f64 ? sizeof(double) : sizeof(float);
    {
        DevBuf dN, dNk;
        SLLG_TRY(dN.alloc(N.size() * sizeof(double)));
        SLLG_TRY(dNk.alloc((size_t)howmany * p->Ph * sizeof(hipfftDoubleComplex)));
        SLLG_HIP(hipMemcpy(dN.p, N.data(), N.size() * sizeof(double), hipMemcpyHostToDevice));
        hipfftHandle h;
        int dims[3] = {p->G[0], p->G[1], p->G[2]};
        SLLG_FFT(hipfftPlanMany(&h, 3, dims, nullptr, 1, (int)p->P, nullptr, 1, (int)p->Ph, HIPFFT_D2Z, howmany));
        hipfftResult r = hipfftExecD2Z(h, dN.as<double>(), dNk.as<hipfftDoubleComplex>());
        hipfftDestroy(h);
        if (r != HIPFFT_SUCCESS) SLLG_FAIL(LLG_E_DEVICE, "FFT-exchange kernel D2Z failed: %d", (int)r);
        const size_t m = (size_t)howmany * p->Ph;
        SLLG_TRY(p->Nkc.alloc(m * 2 * rs));
        if (p->f64)
            k_take_complex<hipfftDoubleComplex><<<blocks_for(m, kTPB), kTPB>>>(m, dNk.as<hipfftDoubleComplex>(),
                                                                              p->Nkc.as<hipfftDoubleComplex>());
        else
            k_take_complex<hipfftComplex><<<blocks_for(m, kTPB), kTPB>>>(m, dNk.as<hipfftDoubleComplex>(),
                                                                        p->Nkc.as<hipfftComplex>());
        SLLG_HIP(hipGetLastError());
        SLLG_HIP(hipDeviceSynchronize());
    }
    p->xh = half_grid_ok(c, p, nms);
    if (p->xh) {
        // half grid: [nms][3][Gx/2][Gy][Gz] real, [nms][3][Gx/2][Mc] complex (x slowest, so
        // lin is unchanged); the kernel in k_fx_xfft's block order replaces Nkc
        const int gx = p->G[0];
        const uint32_t cols = (uint32_t)xf_cols(gx, (int)nms), nn = nms * nms * 9;
        p->Pc = p->P / 2;
        p->Mc = (size_t)p->G[1] * (p->G[2] / 2 + 1);
        const size_t groups = (p->Mc + cols - 1) / cols, total = groups * (size_t)gx * nn * cols;
        SLLG_TRY(p->NkXc.alloc(total * sizeof(float2)));
        k_nkc_to_x<<<blocks_for(total, kTPB), kTPB>>>((uint32_t)p->Mc, (uint32_t)gx, cols, nn, total,
                                                      p->Nkc.as<float2>(), p->NkXc.as<float2>());
        SLLG_HIP(hipGetLastError());
        SLLG_HIP(hipDeviceSynchronize());
        p->Nkc.release();
        SLLG_TRY(upload_twiddles(p));
        if (yz_ok(p)) {
            SLLG_TRY(p->spec.alloc((size_t)3 * nms * (gx / 2) * p->Mc * 2 * rs));
            return yz_setup(c, p, d, spin_species, nms);
        }
        SLLG_TRY(p->grid.alloc((size_t)3 * nms * p->Pc * rs));
        SLLG_TRY(p->hgrid.alloc((size_t)3 * nms * p->Pc * rs));
        SLLG_TRY(p->spec.alloc((size_t)3 * nms * (gx / 2) * p->Mc * 2 * rs));
        SLLG_HIP(hipMemset(p->grid.p, 0, (size_t)3 * nms * p->Pc * rs));
        int d2[2] = {p->G[1], p->G[2]};
        const int planes = (int)(3 * nms) * (gx / 2);
        SLLG_FFT(hipfftPlanMany(&p->fwd, 2, d2, nullptr, 1, p->G[1] * p->G[2], nullptr, 1, (int)p->Mc, HIPFFT_R2C,
                                planes));
        p->have_fwd = true;
        SLLG_FFT(hipfftPlanMany(&p->inv, 2, d2, nullptr, 1, (int)p->Mc, nullptr, 1, p->G[1] * p->G[2], HIPFFT_C2R,
                                planes));
        p->have_inv = true;
        SLLG_FFT(hipfftSetStream(p->fwd, c->stream));
        SLLG_FFT(hipfftSetStream(p->inv, c->stream));
        return 0;
    }
    SLLG_TRY(p->grid.alloc((size_t)3 * nms * p->P * rs));
    SLLG_TRY(p->hgrid.alloc((size_t)3 * nms * p->P * rs));
    SLLG_TRY(p->spec.alloc((size_t)3 * nms * p->Ph * 2 * rs));
    SLLG_TRY(p->hspec.alloc((size_t)3 * nms * p->Ph * 2 * rs));
    SLLG_HIP(hipMemset(p->grid.p, 0, (size_t)3 * nms * p->P * rs));
    int dims[3] = {p->G[0], p->G[1], p->G[2]};
    const int nb = (int)(3 * nms);
    SLLG_FFT(hipfftPlanMany(&p->fwd, 3, dims, nullptr, 1, (int)p->P, nullptr, 1, (int)p->Ph,
                            p->f64 ? HIPFFT_D2Z : HIPFFT_R2C, nb));
    p->have_fwd = true;
    SLLG_FFT(hipfftPlanMany(&p->inv, 3, dims, nullptr, 1, (int)p->Ph, nullptr, 1, (int)p->P,
                            p->f64 ? HIPFFT_Z2D : HIPFFT_C2R, nb));
    p->have_inv = true;
    SLLG_FFT(hipfftSetStream(p->fwd, c->stream));
    SLLG_FFT(hipfftSetStream(p->inv, c->stream));
    return 0;
}

int dipolar_create(llg_ctx* c, const llg_dipolar& d, const uint32_t* spin_species, const llg_fftexch* fx,
                   bool dip_terms) {
    DipolarPlan* p = new DipolarPlan();
    c->dplan = p;
    p->f64 = (c->prec == LLG_FP64_PLAIN);
    for (int a = 0; a < 3; ++a) {
        if (d.dim[a] == 0 || d.nk[a] == 0 || !(d.abc[a] > 0)) SLLG_FAIL(LLG_E_INVALID, "dipolar dim/nk/abc must be positive");
        p->G[a] = (int)(2 * d.dim[a] * d.nk[a]);
    }
    p->P = (size_t)p->G[0] * p->G[1] * p->G[2];
    p->Ph = (size_t)p->G[0] * p->G[1] * (p->G[2] / 2 + 1);
    if (fx) return fftexch_create(c, p, d, spin_species, *fx, dip_terms);
    const uint32_t n = c->n;
    // grid point of each spin, indexed in the handle's device order; one spin per point (two
    // moments at one point have no finite dipolar coupling)
    std::vector<uint32_t> lin(n);
    std::vector<uint8_t> used(p->P, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t m[3];
        for (int a = 0; a < 3; ++a) {
            m[a] = d.mesh[(size_t)i * 3 + a];
            if (m[a] >= d.dim[a] * d.nk[a]) SLLG_FAIL(LLG_E_INVALID, "mesh coordinate of spin %u outside dim*nk", i);
        }
        const size_t g = ((size_t)m[0] * p->G[1] + m[1]) * p->G[2] + m[2];
        if (used[g]) SLLG_FAIL(LLG_E_INVALID, "two spins share a mesh point (spin %u)", i);
        used[g] = 1;
        lin[c->perm.to_dev(i)] = (uint32_t)g;
    }
    const size_t rs = p->f64 ? sizeof(double) : sizeof(float);
    SLLG_TRY(p->lin.alloc(n * sizeof(uint32_t)));
    SLLG_HIP(hipMemcpy(p->lin.p, lin.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    {
        std::vector<double> mu64(n);
        std::vector<float> mu32(n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t s = spin_species ? spin_species[i] : (i % c->ns);
            const uint32_t dv = c->perm.to_dev(i);
            mu64[dv] = c->mu_sp[s];
            mu32[dv] = (float)c->mu_sp[s];
        }
        SLLG_TRY(p->mu.alloc(n * rs));
        SLLG_HIP(hipMemcpy(p->mu.p, p->f64 ? (void*)mu64.data() : (void*)mu32.data(), n * rs, hipMemcpyHostToDevice));
    }
    // kernel: FP64 real space -> FP64 D2Z on the device -> real part in path precision.
    // Handles with the same device, precision and geometry share one copy (replicas).
    KernelKey key{};
    key.device = c->device;
    key.f64 = p->f64;
    for (int a = 0; a < 3; ++a) {
        key.G[a] = p->G[a];
        key.nk[a] = d.nk[a];
        key.abc[a] = d.abc[a];
    }
    const int gx = p->G[0];
    p->xh = half_grid_ok(c, p, 1);
    {
        std::lock_guard<std::mutex> lk(g_nk_mu);
        if (p->xh) {
            auto it = g_nkx_cache.find(key);
            if (it != g_nkx_cache.end()) p->NkX = it->second.lock();
        }
        if (!p->NkX) {
            auto it = g_nk_cache.find(key);
            if (it != g_nk_cache.end()) p->Nk = it->second.lock();
        }
    }
    if (!p->Nk && !p->NkX) {
        p->Nk = std::make_shared<DevBuf>();
        std::vector<double> N;
        fill_kernel(p->G, d.dim, d.nk, d.abc, N);
        DevBuf dN, dNk;
        SLLG_TRY(dN.alloc(6 * p->P * sizeof(double)));
        SLLG_TRY(dNk.alloc(6 * p->Ph * sizeof(hipfftDoubleComplex)));
        SLLG_HIP(hipMemcpy(dN.p, N.data(), 6 * p->P * sizeof(double), hipMemcpyHostToDevice));
        hipfftHandle h;
        int dims[3] = {p->G[0], p->G[1], p->G[2]};
        SLLG_FFT(hipfftPlanMany(&h, 3, dims, nullptr, 1, (int)p->P, nullptr, 1, (int)p->Ph, HIPFFT_D2Z, 6));
        hipfftResult r = hipfftExecD2Z(h, dN.as<double>(), dNk.as<hipfftDoubleComplex>());
        hipfftDestroy(h);
        if (r != HIPFFT_SUCCESS) SLLG_FAIL(LLG_E_DEVICE, "kernel D2Z failed: %d", (int)r);
        SLLG_TRY(p->Nk->alloc(6 * p->Ph * rs));
        const size_t m = 6 * p->Ph;
        if (p->f64)
            k_take_real<double><<<blocks_for(m, kTPB), kTPB>>>(m, dNk.as<hipfftDoubleComplex>(), p->Nk->as<double>());
        else
            k_take_real<float><<<blocks_for(m, kTPB), kTPB>>>(m, dNk.as<hipfftDoubleComplex>(), p->Nk->as<float>());
        SLLG_HIP(hipGetLastError());
        SLLG_HIP(hipDeviceSynchronize());
        std::lock_guard<std::mutex> lk(g_nk_mu);
        g_nk_cache[key] = p->Nk;
    }
    if (p->xh) {
        // half grid: [3][Gx/2][Gy][Gz] real, [3][Gx/2][Mc] complex; lin is unchanged (x slowest)
        p->Pc = p->P / 2;
        p->Mc = (size_t)p->G[1] * (p->G[2] / 2 + 1);
        if (!p->NkX) {  // N(k) in the x kernel's block order, shared like N(k) itself
            const uint32_t cols = (uint32_t)xf_cols(gx, 1);
            const size_t groups = (p->Mc + cols - 1) / cols, total = groups * (size_t)gx * 6 * cols;
            p->NkX = std::make_shared<DevBuf>();
            SLLG_TRY(p->NkX->alloc(total * sizeof(float)));
            k_nk_to_x<<<blocks_for(total, kTPB), kTPB>>>((uint32_t)p->Mc, (uint32_t)gx, cols, total,
                                                         p->Nk->as<float>(), p->NkX->as<float>());
            SLLG_HIP(hipGetLastError());
            SLLG_HIP(hipDeviceSynchronize());
            std::lock_guard<std::mutex> lk(g_nk_mu);
            g_nkx_cache[key] = p->NkX;
        }
        p->Nk.reset();  // the [6][Ph] layout is not read on this path
        SLLG_TRY(upload_twiddles(p));
        if (yz_ok(p)) {
            SLLG_TRY(p->spec.alloc(3 * (size_t)(gx / 2) * p->Mc * 2 * rs));
            return yz_setup(c, p, d, spin_species, 1);
        }
        SLLG_TRY(p->grid.alloc(3 * p->Pc * rs));
        SLLG_TRY(p->hgrid.alloc(3 * p->Pc * rs));
        SLLG_TRY(p->spec.alloc(3 * (size_t)(gx / 2) * p->Mc * 2 * rs));
        SLLG_HIP(hipMemset(p->grid.p, 0, 3 * p->Pc * rs));
        int d2[2] = {p->G[1], p->G[2]};
        const int planes = 3 * (gx / 2);
        SLLG_FFT(hipfftPlanMany(&p->fwd, 2, d2, nullptr, 1, p->G[1] * p->G[2], nullptr, 1, (int)p->Mc, HIPFFT_R2C,
                                planes));
        p->have_fwd = true;
        SLLG_FFT(hipfftPlanMany(&p->inv, 2, d2, nullptr, 1, (int)p->Mc, nullptr, 1, p->G[1] * p->G[2], HIPFFT_C2R,
                                planes));
        p->have_inv = true;
        SLLG_FFT(hipfftSetStream(p->fwd, c->stream));
        SLLG_FFT(hipfftSetStream(p->inv, c->stream));
        return 0;
    }
    SLLG_TRY(p->grid.alloc(3 * p->P * rs));
    SLLG_TRY(p->hgrid.alloc(3 * p->P * rs));
    SLLG_TRY(p->spec.alloc(3 * p->Ph * 2 * rs));
    SLLG_HIP(hipMemset(p->grid.p, 0, 3 * p->P * rs));
    int dims[3] = {p->G[0], p->G[1], p->G[2]};
    SLLG_FFT(hipfftPlanMany(&p->fwd, 3, dims, nullptr, 1, (int)p->P, nullptr, 1, (int)p->Ph,
                            p->f64 ? HIPFFT_D2Z : HIPFFT_R2C, 3));
    p->have_fwd = true;
    SLLG_FFT(hipfftPlanMany(&p->inv, 3, dims, nullptr, 1, (int)p->Ph, nullptr, 1, (int)p->P,
                            p->f64 ? HIPFFT_Z2D : HIPFFT_C2R, 3));
    p->have_inv = true;
    SLLG_FFT(hipfftSetStream(p->fwd, c->stream));
    SLLG_FFT(hipfftSetStream(p->inv, c->stream));
    return 0;
}

int dipolar_set_stream(llg_ctx* c) {
    if (!c->dplan) return 0;
    if (c->dplan->have_fwd) SLLG_FFT(hipfftSetStream(c->dplan->fwd, c->stream));
    if (c->dplan->have_inv) SLLG_FFT(hipfftSetStream(c->dplan->inv, c->stream));
    return 0;
}

bool dipolar_is_exchange(const llg_ctx* c) { return c->dplan && c->dplan->general; }

// The FP32 half-grid plane path runs only kernels of this file (no hipFFT plans, no memsets),
// so a stream capture of a refresh replays exactly.
bool dipolar_capturable(const llg_ctx* c) { return c->dplan && !c->dplan->f64 && c->dplan->yz; }

void dipolar_destroy(llg_ctx* c) {
    delete c->dplan;
    c->dplan = nullptr;
}

uint64_t dipolar_bytes(const llg_ctx* c) {
    const DipolarPlan* p = c->dplan;
    if (!p) return 0;
    uint64_t b = p->lin.bytes + p->mu.bytes + p->grid.bytes + p->spec.bytes + p->hgrid.bytes + p->sp.bytes +
                 p->Nkc.bytes + p->hspec.bytes + p->tw.bytes + p->map.bytes + p->twy.bytes + p->twz.bytes +
                 (p->Nk ? p->Nk->bytes : 0) +
                 (p->NkX ? p->NkX->bytes : 0) + p->NkXc.bytes;
    size_t w = 0;
    if (p->have_fwd && hipfftGetSize(p->fwd, &w) == HIPFFT_SUCCESS) b += w;
    if (p->have_inv && hipfftGetSize(p->inv, &w) == HIPFFT_SUCCESS) b += w;
    return b;
}

int dipolar_refresh(llg_ctx* c, const void* S, void* out) {
    DipolarPlan* p = c->dplan;
    if (!p) SLLG_FAIL(LLG_E_STATE, "dipolar plan missing");
    const uint32_t n = c->n;
    const dim3 gn(blocks_for(n, kTPB)), gk(blocks_for(p->Ph, kTPB)), b(kTPB);
    if (p->general) {
        const uint32_t* lin = p->lin.as<uint32_t>();
        const uint32_t* sp = p->sp.as<uint32_t>();
        if (p->f64) {
            k_fx_scatter<double><<<gn, b, 0, c->stream>>>(n, p->P, lin, sp, (const double*)S, p->grid.as<double>());
            SLLG_HIP(hipGetLastError());
            SLLG_FFT(hipfftExecD2Z(p->fwd, p->grid.as<double>(), p->spec.as<hipfftDoubleComplex>()));
            k_fx_mult<double, hipfftDoubleComplex><<<gk, b, 0, c->stream>>>(
                p->Ph, p->nms, p->Nkc.as<hipfftDoubleComplex>(), p->spec.as<hipfftDoubleComplex>(),
                p->hspec.as<hipfftDoubleComplex>());
            SLLG_HIP(hipGetLastError());
            SLLG_FFT(hipfftExecZ2D(p->inv, p->hspec.as<hipfftDoubleComplex>(), p->hgrid.as<double>()));
            k_fx_gather<double><<<gn, b, 0, c->stream>>>(n, p->P, 1.0 / (double)p->P, lin, sp, p->hgrid.as<double>(),
                                                         (double*)out);
        } else if (p->yz) {
            const YZArgs ya = yz_args(c, p);
            const uint32_t planes = 3 * p->nms * ya.lx;
            SLLG_TRY(launch_yz(true, p->G[1], planes, ya, (const float*)S, 0.f, nullptr, c->stream));
            SLLG_TRY(launch_fx_xfft(p->G[0], p->nms, (uint32_t)p->Mc, p->tw.as<float2>(), p->NkXc.as<float2>(),
                                    p->spec.as<float2>(), c->stream));
            SLLG_TRY(launch_yz(false, p->G[1], planes, ya, nullptr, 1.0f / (float)p->P, (float*)out, c->stream));
        } else if (p->xh) {
            k_fx_scatter<float><<<gn, b, 0, c->stream>>>(n, p->Pc, lin, sp, (const float*)S, p->grid.as<float>());
            SLLG_HIP(hipGetLastError());
            SLLG_FFT(hipfftExecR2C(p->fwd, p->grid.as<float>(), p->spec.as<hipfftComplex>()));
            SLLG_TRY(launch_fx_xfft(p->G[0], p->nms, (uint32_t)p->Mc, p->tw.as<float2>(), p->NkXc.as<float2>(),
                                    p->spec.as<float2>(), c->stream));
            SLLG_FFT(hipfftExecC2R(p->inv, p->spec.as<hipfftComplex>(), p->hgrid.as<float>()));
            k_fx_gather<float><<<gn, b, 0, c->stream>>>(n, p->Pc, 1.0f / (float)p->P, lin, sp, p->hgrid.as<float>(),
                                                        (float*)out);
        } else {
            k_fx_scatter<float><<<gn, b, 0, c->stream>>>(n, p->P, lin, sp, (const float*)S, p->grid.as<float>());
            SLLG_HIP(hipGetLastError());
            SLLG_FFT(hipfftExecR2C(p->fwd, p->grid.as<float>(), p->spec.as<hipfftComplex>()));
            k_fx_mult<float, hipfftComplex><<<gk, b, 0, c->stream>>>(p->Ph, p->nms, p->Nkc.as<hipfftComplex>(),
                                                                     p->spec.as<hipfftComplex>(),
                                                                     p->hspec.as<hipfftComplex>());
            SLLG_HIP(hipGetLastError());
            SLLG_FFT(hipfftExecC2R(p->inv, p->hspec.as<hipfftComplex>(), p->hgrid.as<float>()));
            k_fx_gather<float><<<gn, b, 0, c->stream>>>(n, p->P, 1.0f / (float)p->P, lin, sp, p->hgrid.as<float>(),
                                                        (float*)out);
        }
        SLLG_HIP(hipGetLastError());
        return 0;
    }
    if (p->f64) {
        k_dip_scatter<double><<<gn, b, 0, c->stream>>>(n, p->P, p->lin.as<uint32_t>(), p->mu.as<double>(),
                                                       (const double*)S, p->grid.as<double>());
        SLLG_HIP(hipGetLastError());
        SLLG_FFT(hipfftExecD2Z(p->fwd, p->grid.as<double>(), p->spec.as<hipfftDoubleComplex>()));
        k_dip_mult<double, hipfftDoubleComplex><<<gk, b, 0, c->stream>>>(p->Ph, p->Nk->as<double>(),
                                                                         p->spec.as<hipfftDoubleComplex>());
        SLLG_HIP(hipGetLastError());
        SLLG_FFT(hipfftExecZ2D(p->inv, p->spec.as<hipfftDoubleComplex>(), p->hgrid.as<double>()));
        k_dip_gather<double><<<gn, b, 0, c->stream>>>(n, p->P, 1.0 / (double)p->P, p->lin.as<uint32_t>(),
                                                      p->hgrid.as<double>(), (double*)out);
    } else if (p->yz) {
        const YZArgs ya = yz_args(c, p);
        const uint32_t planes = 3 * ya.lx;
        SLLG_TRY(launch_yz(true, p->G[1], planes, ya, (const float*)S, 0.f, nullptr, c->stream));
        SLLG_TRY(launch_dip_xfft(p->G[0], (uint32_t)p->Mc, p->tw.as<float2>(), p->NkX->as<float>(),
                                 p->spec.as<float2>(), c->stream));
        SLLG_TRY(launch_yz(false, p->G[1], planes, ya, nullptr, 1.0f / (float)p->P, (float*)out, c->stream));
    } else if (p->xh) {
        k_dip_scatter<float><<<gn, b, 0, c->stream>>>(n, p->Pc, p->lin.as<uint32_t>(), p->mu.as<float>(),
                                                      (const float*)S, p->grid.as<float>());
        SLLG_HIP(hipGetLastError());
        SLLG_FFT(hipfftExecR2C(p->fwd, p->grid.as<float>(), p->spec.as<hipfftComplex>()));
        SLLG_TRY(launch_dip_xfft(p->G[0], (uint32_t)p->Mc, p->tw.as<float2>(), p->NkX->as<float>(),
                                 p->spec.as<float2>(), c->stream));
        SLLG_FFT(hipfftExecC2R(p->inv, p->spec.as<hipfftComplex>(), p->hgrid.as<float>()));
        k_dip_gather<float><<<gn, b, 0, c->stream>>>(n, p->Pc, 1.0f / (float)p->P, p->lin.as<uint32_t>(),
                                                     p->hgrid.as<float>(), (float*)out);
    } else {
        k_dip_scatter<float><<<gn, b, 0, c->stream>>>(n, p->P, p->lin.as<uint32_t>(), p->mu.as<float>(),
                                                      (const float*)S, p->grid.as<float>());
        SLLG_HIP(hipGetLastError());
        SLLG_FFT(hipfftExecR2C(p->fwd, p->grid.as<float>(), p->spec.as<hipfftComplex>()));
        k_dip_mult<float, hipfftComplex><<<gk, b, 0, c->stream>>>(p->Ph, p->Nk->as<float>(), p->spec.as<hipfftComplex>());
        SLLG_HIP(hipGetLastError());
        SLLG_FFT(hipfftExecC2R(p->inv, p->spec.as<hipfftComplex>(), p->hgrid.as<float>()));
        k_dip_gather<float><<<gn, b, 0, c->stream>>>(n, p->P, 1.0f / (float)p->P, p->lin.as<uint32_t>(),
                                                     p->hgrid.as<float>(), (float*)out);
    }
    SLLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace sllg
