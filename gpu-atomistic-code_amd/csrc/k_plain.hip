// k_plain.hip - FP64 plain path: the reference GPU step structure (separate SpMV and
// integrator passes, FP64 couplings per nnz, per-spin parameter arrays) written for
// gfx950 with SoA state.  Built with -ffp-contract=off; every operation follows the
// order of the FP64 oracle so that both give the same bits.
//
//   k_plain_spmv     <- cufields::CSpMV_CSR / _OffDiag / CSpMV_DIA / _OffDiag (src/cufields.cu:97-246)
//   k_plain_noise    <- curandGenerateNormal (src/cuda.cu:148), replaced by Philox4x32-10 + Box-Muller
//   k_plain_heun1    <- cuint::CHeun1 (src/cuint.cu:60-163)
//   k_plain_heun2    <- cuint::CHeun2 (src/cuint.cu:165-242)
//   k_plain_rk_stage <- cuintRK4::CdetRK4k1..k4 (src/cuintRK4.cu:62-434), cuintRK5::CdetRK5k1..k6
//                       (src/cuintRK5.cu:79-628), with textbook stage inputs
//   k_plain_rk_comb  <- CdetSnp1 (src/cuintRK4.cu:436-456, src/cuintRK5.cu:630-651)
#include "field.hpp"
#include "kernels.hpp"
#include "philox.hpp"

namespace sllg {

struct PlainArgs {
    uint32_t n;
    double rdt, T;
    double B[3];
    const double* sigma;
    const double* llgpf;
    const double* lambda;
    const double* k1u;
    const double* k1udir;  // SoA 3n
    const double* kc;      // per spin or null
    const double* stg;     // 3n or null
    const double* xi;      // 3n or null (T == 0)
    AnisG<double> an;
};

template <int NC>
__global__ __launch_bounds__(kTPB) void k_plain_spmv(Exch<double> e, const double* __restrict__ S,
                                                     const double* __restrict__ dem, double* __restrict__ H) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    const uint32_t n = e.n;
    if (i >= n) return;
    double dot[3];
    exch_row_ref<NC, double, double, double>(e, i, S, dot);
    double hx[3];
    for (int a = 0; a < 3; ++a) hx[a] = (dem ? dem[(size_t)a * n + i] : 0.0) + dot[a];
    if (e.fs_xadj) {  // CH += four-spin field (src/cuda.cu:143-146)
        double h4[3];
        fourspin_row<double, double, double>(e, i, S, h4);
        for (int a = 0; a < 3; ++a) hx[a] = hx[a] + h4[a];
    }
    for (int a = 0; a < 3; ++a) H[(size_t)a * n + i] = hx[a];
}

__global__ __launch_bounds__(kTPB) void k_plain_noise(uint32_t n, uint64_t seed, uint64_t stream, uint64_t t,
                                                      double* __restrict__ xi) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    double z[3];
    normals3_f64(seed, stream, i, t, z);
    xi[i] = z[0];
    xi[(size_t)n + i] = z[1];
    xi[2 * (size_t)n + i] = z[2];
}

// Field at spin i (own spin s, exchange+demag hex): h = ((hex + stg) + th) + B, then local terms
// (th = 0 when noise is false: the deterministic field).
__device__ __forceinline__ void plain_field(const PlainArgs& a, uint32_t i, const double s[3], const double hex[3],
                                            double th[3], double h[3], bool noise = true) {
    const uint32_t n = a.n;
    if (noise && a.xi) {
        const double TP = sqrt(a.T) * a.sigma[i];
        for (int c = 0; c < 3; ++c) th[c] = a.xi[(size_t)c * n + i] * TP;
    } else {
        th[0] = th[1] = th[2] = 0.0;
    }
    for (int c = 0; c < 3; ++c) {
        const double st = a.stg ? a.stg[(size_t)c * n + i] : 0.0;
        h[c] = ((hex[c] + st) + th[c]) + a.B[c];
    }
    const double d[3] = {a.k1udir[i], a.k1udir[(size_t)n + i], a.k1udir[2 * (size_t)n + i]};
    local_terms<double>(a.k1u[i], d, a.kc ? a.kc[i] : 0.0, a.an, s, h);
}

__global__ __launch_bounds__(kTPB) void k_plain_heun1(PlainArgs a, const double* __restrict__ H,
                                                      const double* __restrict__ S, double* __restrict__ Y,
                                                      double* __restrict__ FN) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    const uint32_t n = a.n;
    if (i >= n) return;
    const double s[3] = {S[i], S[(size_t)n + i], S[2 * (size_t)n + i]};
    const double hex[3] = {H[i], H[(size_t)n + i], H[2 * (size_t)n + i]};
    double th[3], h[3], f[3];
    plain_field(a, i, s, hex, th, h);
    llg_rhs<double>(a.llgpf[i], a.lambda[i], s, h, f);
    double es[3], mods = 0.0;
    for (int c = 0; c < 3; ++c) {
        FN[(size_t)c * n + i] = f[c];
        es[c] = s[c] + f[c] * a.rdt;
        mods += es[c] * es[c];
    }
    const double nrm = sqrt(mods);
    for (int c = 0; c < 3; ++c) Y[(size_t)c * n + i] = es[c] / nrm;
}

__global__ __launch_bounds__(kTPB) void k_plain_heun2(PlainArgs a, const double* __restrict__ H,
                                                      const double* __restrict__ Y, double* __restrict__ S,
                                                      const double* __restrict__ FN, double* __restrict__ DET) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    const uint32_t n = a.n;
    if (i >= n) return;
    const double s[3] = {Y[i], Y[(size_t)n + i], Y[2 * (size_t)n + i]};
    const double hex[3] = {H[i], H[(size_t)n + i], H[2 * (size_t)n + i]};
    double th[3], h[3], f[3];
    plain_field(a, i, s, hex, th, h);
    llg_rhs<double>(a.llgpf[i], a.lambda[i], s, h, f);
    double ps[3] = {S[i], S[(size_t)n + i], S[2 * (size_t)n + i]};
    double mods = 0.0;
    for (int c = 0; c < 3; ++c) {
        ps[c] += (0.5 * (FN[(size_t)c * n + i] + f[c]) * a.rdt);
        mods += ps[c] * ps[c];
    }
    const double nrm = sqrt(mods);
    for (int c = 0; c < 3; ++c) {
        S[(size_t)c * n + i] = ps[c] / nrm;
        DET[(size_t)c * n + i] = h[c] - th[c];
    }
}

// RK stage m: k_m = rdt f(Y_m) into K[m]; then the next stage input in place of Y:
// RK4: Y = S + a_m k_m (a = 1/2, 1/2, 1); Cash-Karp: Y = (((S + b_{m+1,0} k_0) + ...) + b_{m+1,m} k_m).
struct RkTab {
    double b[6][5];
    int nb[6];
};
__constant__ RkTab c_rk5;

template <int SCHEME, int M>
__global__ __launch_bounds__(kTPB) void k_plain_rk_stage(PlainArgs a, const double* __restrict__ H,
                                                         double* __restrict__ Y, const double* __restrict__ S,
                                                         double* __restrict__ K0, double* __restrict__ K1,
                                                         double* __restrict__ K2, double* __restrict__ K3,
                                                         double* __restrict__ K4, double* __restrict__ K5) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    const uint32_t n = a.n;
    if (i >= n) return;
    const double* src = (M == 0) ? S : Y;
    const double s[3] = {src[i], src[(size_t)n + i], src[2 * (size_t)n + i]};
    const double hex[3] = {H[i], H[(size_t)n + i], H[2 * (size_t)n + i]};
    double th[3], h[3], f[3];
    plain_field(a, i, s, hex, th, h);
    llg_rhs<double>(a.llgpf[i], a.lambda[i], s, h, f);
    double* Ks[6] = {K0, K1, K2, K3, K4, K5};
    double k[3];
    for (int c = 0; c < 3; ++c) {
        k[c] = a.rdt * f[c];
        Ks[M][(size_t)c * n + i] = k[c];
    }
    if constexpr (SCHEME == LLG_RK4) {
        if constexpr (M < 3) {
            const double am = (M == 2) ? 1.0 : 0.5;
            for (int c = 0; c < 3; ++c) Y[(size_t)c * n + i] = S[(size_t)c * n + i] + am * k[c];
        }
    } else {
        if constexpr (M < 5) {
            for (int c = 0; c < 3; ++c) {
                double v = S[(size_t)c * n + i];
                for (int j = 0; j <= M; ++j) {
                    const double kj = (j == M) ? k[c] : Ks[j][(size_t)c * n + i];
                    v = v + c_rk5.b[M + 1][j] * kj;
                }
                Y[(size_t)c * n + i] = v;
            }
        }
    }
}

template <int SCHEME>
__global__ __launch_bounds__(kTPB) void k_plain_rk_comb(uint32_t n, double* __restrict__ S,
                                                        const double* __restrict__ K0, const double* __restrict__ K1,
                                                        const double* __restrict__ K2, const double* __restrict__ K3,
                                                        const double* __restrict__ K4, const double* __restrict__ K5) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    double ls[3], mods = 0.0;
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        if constexpr (SCHEME == LLG_RK4) {
            const double onesixth = 1.0 / 6.0, onethird = 1.0 / 3.0;
            ls[c] = S[o] + K0[o] * onesixth + K1[o] * onethird + K2[o] * onethird + K3[o] * onesixth;
        } else {
            double v = S[o];
            v = v + K0[o] * (37.0 / 378.0);
            v = v + K2[o] * (250.0 / 621.0);
            v = v + K3[o] * (125.0 / 594.0);
            v = v + K5[o] * (512.0 / 1771.0);
            ls[c] = v;
        }
        mods += ls[c] * ls[c];
    }
    const double nrm = sqrt(mods);
    for (int c = 0; c < 3; ++c) S[(size_t)c * n + i] = ls[c] / nrm;
}

// Deterministic field (no thermal part) at spins Sp, FP64 plain layout.
__global__ __launch_bounds__(kTPB) void k_plain_detfield(PlainArgs a, const double* __restrict__ H,
                                                         const double* __restrict__ Sp, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    const uint32_t n = a.n;
    if (i >= n) return;
    const double s[3] = {Sp[i], Sp[(size_t)n + i], Sp[2 * (size_t)n + i]};
    const double hex[3] = {H[i], H[(size_t)n + i], H[2 * (size_t)n + i]};
    double th[3], h[3];
    plain_field(a, i, s, hex, th, h, false);
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = h[c];
}

// ---------------------------------------------------------------- host launchers

static PlainArgs plain_args(const llg_ctx* c, bool with_noise) {
    PlainArgs a;
    a.n = c->n;
    a.rdt = c->rdt;
    a.T = c->T;
    for (int k = 0; k < 3; ++k) a.B[k] = c->B[k];
    a.sigma = c->p_sigma.as<double>();
    a.llgpf = c->p_llgpf.as<double>();
    a.lambda = c->p_lambda.as<double>();
    a.k1u = c->p_k1u.as<double>();
    a.k1udir = c->p_k1udir.as<double>();
    a.kc = c->p_kc.as<double>();
    a.stg = c->stag_set ? c->STG.as<double>() : nullptr;
    a.xi = (with_noise && c->T > 0.0) ? c->XI.as<double>() : nullptr;
    a.an = c->an64;
    return a;
}

int plain_spmv(llg_ctx* c, const double* S, double* H) {
    const dim3 g(blocks_for(c->n, kTPB)), b(kTPB);
    const double* dem = c->dip ? c->DEM.as<double>() : nullptr;
    switch (c->ex64.ncomp) {
        case 1: k_plain_spmv<1><<<g, b, 0, c->stream>>>(c->ex64, S, dem, H); break;
        case 3: k_plain_spmv<3><<<g, b, 0, c->stream>>>(c->ex64, S, dem, H); break;
        default: k_plain_spmv<9><<<g, b, 0, c->stream>>>(c->ex64, S, dem, H); break;
    }
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_noise(llg_ctx* c, uint64_t t) {
    k_plain_noise<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, c->seed, c->stream_id, t, c->XI.as<double>());
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_heun1(llg_ctx* c) {
    PlainArgs a = plain_args(c, true);
    k_plain_heun1<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(a, c->H.as<double>(), c->S.as<double>(),
                                                                  c->Y.as<double>(), c->FN.as<double>());
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_heun2(llg_ctx* c) {
    PlainArgs a = plain_args(c, true);
    k_plain_heun2<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(a, c->H.as<double>(), c->Y.as<double>(),
                                                                  c->S.as<double>(), c->FN.as<double>(),
                                                                  c->DET.as<double>());
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_rk_init_tables() {
    static bool done = false;
    if (done) return 0;
    RkTab t{};
    const double b2[1] = {0.2};
    const double b3[2] = {3.0 / 40.0, 9.0 / 40.0};
    const double b4[3] = {0.3, -0.9, 1.2};
    const double b5[4] = {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0};
    const double b6[5] = {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0};
    const double* rows[6] = {nullptr, b2, b3, b4, b5, b6};
    for (int m = 1; m < 6; ++m)
        for (int j = 0; j < m; ++j) t.b[m][j] = rows[m][j];
    SLLG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rk5), &t, sizeof(t)));
    done = true;
    return 0;
}

int plain_rk_stage(llg_ctx* c, int m) {
    PlainArgs a = plain_args(c, true);
    const dim3 g(blocks_for(c->n, kTPB)), b(kTPB);
    double* K[6];
    for (int k = 0; k < 6; ++k) K[k] = c->K[k].as<double>();
    const double* H = c->H.as<double>();
    double* Y = c->Y.as<double>();
    const double* S = c->S.as<double>();
#define SLLG_RK_CASE(SCH, MM) \
    k_plain_rk_stage<SCH, MM><<<g, b, 0, c->stream>>>(a, H, Y, S, K[0], K[1], K[2], K[3], K[4], K[5])
    if (c->scheme == LLG_RK4) {
        switch (m) {
            case 0: SLLG_RK_CASE(LLG_RK4, 0); break;
            case 1: SLLG_RK_CASE(LLG_RK4, 1); break;
            case 2: SLLG_RK_CASE(LLG_RK4, 2); break;
            default: SLLG_RK_CASE(LLG_RK4, 3); break;
        }
    } else {
        switch (m) {
            case 0: SLLG_RK_CASE(LLG_RK5, 0); break;
            case 1: SLLG_RK_CASE(LLG_RK5, 1); break;
            case 2: SLLG_RK_CASE(LLG_RK5, 2); break;
            case 3: SLLG_RK_CASE(LLG_RK5, 3); break;
            case 4: SLLG_RK_CASE(LLG_RK5, 4); break;
            default: SLLG_RK_CASE(LLG_RK5, 5); break;
        }
    }
#undef SLLG_RK_CASE
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_rk_comb(llg_ctx* c) {
    const dim3 g(blocks_for(c->n, kTPB)), b(kTPB);
    double* K[6];
    for (int k = 0; k < 6; ++k) K[k] = c->K[k].as<double>();
    if (c->scheme == LLG_RK4)
        k_plain_rk_comb<LLG_RK4><<<g, b, 0, c->stream>>>(c->n, c->S.as<double>(), K[0], K[1], K[2], K[3], K[4], K[5]);
    else
        k_plain_rk_comb<LLG_RK5><<<g, b, 0, c->stream>>>(c->n, c->S.as<double>(), K[0], K[1], K[2], K[3], K[4], K[5]);
    SLLG_HIP(hipGetLastError());
    return 0;
}

int plain_detfield(llg_ctx* c, const double* Sp, double* out) {
    SLLG_TRY(plain_spmv(c, Sp, c->H.as<double>()));
    PlainArgs a = plain_args(c, false);
    k_plain_detfield<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(a, c->H.as<double>(), Sp, out);
    SLLG_HIP(hipGetLastError());
    return 0;
}

// Staggered-field ramp on the device, cufields::CincStagFields (src/cufields.cu:46-56):
// H_stag = r * H0 with H0 the field last given, kept on the device in FP64 (device order,
// SoA).  The product is taken in FP64 and rounded once to the path's precision.
template <class P>
__global__ __launch_bounds__(kTPB) void k_stag_scale(size_t n3, double r, const double* __restrict__ h0,
                                                     P* __restrict__ h) {
    const size_t k = (size_t)blockIdx.x * kTPB + threadIdx.x;
    if (k < n3) h[k] = (P)(r * h0[k]);
}

int stag_scale(llg_ctx* c, double r) {
    const size_t n3 = 3 * (size_t)c->n;
    if (c->prec == LLG_FP64_PLAIN)
        k_stag_scale<double><<<blocks_for(n3, kTPB), kTPB, 0, c->stream>>>(n3, r, c->STG0.as<double>(),
                                                                           c->STG.as<double>());
    else
        k_stag_scale<float><<<blocks_for(n3, kTPB), kTPB, 0, c->stream>>>(n3, r, c->STG0.as<double>(),
                                                                          c->STG.as<float>());
    SLLG_HIP(hipGetLastError());
    return 0;
}

}  // namespace sllg
