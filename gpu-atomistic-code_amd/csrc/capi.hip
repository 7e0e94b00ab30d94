// capi.hip - the C ABI declared in include/llg_capi.h.
//
// llg_step reproduces llg::integrate(t) (src/llg.cpp:639-664) for t = t0 .. t0+nsteps-1:
//   Heun: [dipolar refresh if t % refresh == 0] -> noise -> predictor -> corrector
//         (cullg::llgGPU, src/cuda.cu:43-258)
//   RK4 / RK5: [dipolar refresh] -> noise -> stages -> combine
//         (cullg::llgGPURK4 / llgGPURK5, src/cudaRK4.cu:38-233, src/cudaRK5.cu:38-273)
// The FP64 plain path launches the reference's kernel sequence (SpMV, stage, SpMV,
// stage); the FP32 path launches one fused kernel per field evaluation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace sllg {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
const char* last_error() { return g_err.c_str(); }

// Events around every launch of a step, for llg_profile_kernels.
struct KernelTimer {
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    ~KernelTimer() {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
};

namespace {

struct Scope {
    llg_ctx* c;
    const char* name;
    hipEvent_t a{}, b{};
    Scope(llg_ctx* c_, const char* n) : c(c_), name(n) {
        if (c->timer) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Scope() {
        if (c->timer) {
            (void)hipEventRecord(b, c->stream);
            c->timer->recs.push_back({name, a, b});
        }
    }
};

#define SLLG_TIMED(name, expr)      \
    do {                            \
        Scope sc_(c, name);         \
        SLLG_TRY(expr);             \
    } while (0)

size_t prec_bytes(const llg_ctx* c) { return c->prec == LLG_FP64_PLAIN ? sizeof(double) : sizeof(float); }

// Upload a host SoA double[3n] (reference spin order) into a state buffer of the path's
// precision, in the device order.
int upload_state(llg_ctx* c, DevBuf& dst, const double* x, const double* y, const double* z) {
    const uint32_t n = c->n;
    const double* src[3] = {x, y, z};
    if (&dst == &c->S) {
        c->g3_valid = false;  // the AoS3 gather copy no longer matches
        c->s_soa_valid = true;
    }
    if (&dst == &c->Y) c->y_soa_valid = true;
    if (c->prec == LLG_FP64_PLAIN) {
        for (int a = 0; a < 3; ++a)
            SLLG_HIP(hipMemcpyAsync(dst.as<double>() + (size_t)a * n, src[a], n * 8, hipMemcpyHostToDevice, c->stream));
    } else {
        std::vector<float> t(3 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t d = c->perm.to_dev(i);
            for (int a = 0; a < 3; ++a) t[(size_t)a * n + d] = (float)src[a][i];
        }
        SLLG_HIP(hipMemcpyAsync(dst.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, c->stream));
    }
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Download a device SoA vector (FP64 or FP32, device order) into host SoA doubles in the
// reference spin order.
int download_state(llg_ctx* c, const void* src_dev, bool src_f64, double* x, double* y, double* z) {
    const uint32_t n = c->n;
    double* dst[3] = {x, y, z};
    if (src_f64 && !c->permuted()) {
        const double* s = (const double*)src_dev;
        for (int a = 0; a < 3; ++a)
            SLLG_HIP(hipMemcpyAsync(dst[a], s + (size_t)a * n, n * 8, hipMemcpyDeviceToHost, c->stream));
        SLLG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    }
    const size_t eb = src_f64 ? 8 : 4;
    std::vector<unsigned char> raw(3 * (size_t)n * eb);
    SLLG_HIP(hipMemcpyAsync(raw.data(), src_dev, raw.size(), hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    const double* r64 = reinterpret_cast<const double*>(raw.data());
    const float* r32 = reinterpret_cast<const float*>(raw.data());
    for (uint32_t i = 0; i < n; ++i) {
        const size_t d = c->perm.to_dev(i);
        for (int a = 0; a < 3; ++a) dst[a][i] = src_f64 ? r64[(size_t)a * n + d] : (double)r32[(size_t)a * n + d];
    }
    return 0;
}

template <class R>
void make_anis(const llg_anisotropy& an, double basemm, AnisG<R>& g) {
    // prefactor 2K/(mu_0 mu_B) as in src/cuint.cu:98-137, terms in the reference's order
    g = AnisG<R>{};
    auto add = [&](double K, const double* d) {
        g.lin_pf[g.nlin] = (R)(2.0 * K / basemm);
        for (int a = 0; a < 3; ++a) g.lin_dir[g.nlin][a] = (R)d[a];
        g.nlin++;
    };
    if (an.k2perp != 0.0) add(an.k2perp, an.k2perp_dir);
    if (an.k2par != 0.0) add(an.k2par, an.k2par_dir);
    for (int dir = 0; dir < 3; ++dir) {
        if (an.k4perp != 0.0) add(an.k4perp, an.k4perp_dirs + 3 * dir);
        if (an.k4par != 0.0) add(an.k4par, an.k4par_dirs + 3 * dir);
    }
}

int take_noise(llg_ctx* c, const double** dev_xi) {
    *dev_xi = nullptr;
    if (c->noise_mode != LLG_NOISE_INJECTED) return 0;
    if (c->noise_q.empty()) SLLG_FAIL(LLG_E_STATE, "injected noise exhausted: push more with llg_push_noise");
    std::vector<double> v = std::move(c->noise_q.front());
    c->noise_q.pop_front();
    if (c->permuted()) {  // [3][n] reference order -> device order
        const uint32_t n = c->n;
        std::vector<double> w(v.size());
        for (uint32_t i = 0; i < n; ++i) {
            const size_t d = c->perm.to_dev(i);
            for (int a = 0; a < 3; ++a) w[(size_t)a * n + d] = v[(size_t)a * n + i];
        }
        v.swap(w);
    }
    SLLG_HIP(hipMemcpyAsync(c->XI.p, v.data(), v.size() * 8, hipMemcpyHostToDevice, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    *dev_xi = c->XI.as<double>();
    return 0;
}

int step_plain(llg_ctx* c, uint64_t t) {
    double* S = c->S.as<double>();
    double* Y = c->Y.as<double>();
    double* H = c->H.as<double>();
    if (c->dip && t % c->dip_refresh == 0) SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, S, c->DEM.p));
    const double* xi = nullptr;
    SLLG_TRY(take_noise(c, &xi));
    if (c->noise_mode == LLG_NOISE_COUNTER && c->T > 0.0) SLLG_TIMED("plain_noise", plain_noise(c, t));
    if (c->scheme == LLG_HEUN) {
        SLLG_TIMED("plain_spmv", plain_spmv(c, S, H));
        SLLG_TIMED("plain_heun1", plain_heun1(c));
        if (c->dip && c->dip_every_stage) SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, Y, c->DEM.p));
        SLLG_TIMED("plain_spmv", plain_spmv(c, Y, H));
        SLLG_TIMED("plain_heun2", plain_heun2(c));
        c->det_valid = true;
    } else {
        const int ns = c->scheme == LLG_RK4 ? 4 : 6;
        for (int m = 0; m < ns; ++m) {
            if (m > 0 && c->dip && c->dip_every_stage) SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, Y, c->DEM.p));
            SLLG_TIMED("plain_spmv", plain_spmv(c, m == 0 ? S : Y, H));
            SLLG_TIMED("plain_rk_stage", plain_rk_stage(c, m));
        }
        SLLG_TIMED("plain_rk_combine", plain_rk_comb(c));
        c->det_valid = false;
    }
    return 0;
}

int step_fused(llg_ctx* c, uint64_t t) {
    if (c->dip && t % c->dip_refresh == 0) {
        SLLG_TRY(sync_spins(c));
        SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, c->S.p, c->DEM.p));
    }
    StepSched s{t, nullptr};
    SLLG_TRY(take_noise(c, &s.xi));
    if (c->scheme == LLG_HEUN) {
        SLLG_TIMED("heun_predictor", fused_heun_pred(c, s));
        if (c->dip && c->dip_every_stage) SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, c->Y.p, c->DEM.p));
        SLLG_TIMED("heun_corrector", fused_heun_corr(c, s));
        c->det_valid = true;
    } else {
        const int ns = c->scheme == LLG_RK4 ? 4 : 6;
        void* P[2] = {c->Y.p, c->A.p};
        for (int m = 0; m < ns; ++m) {
            if (m > 0 && c->dip && c->dip_every_stage)
                SLLG_TIMED("dipolar_refresh", dipolar_refresh(c, P[m & 1], c->DEM.p));
            SLLG_TIMED(c->scheme == LLG_RK4 ? "rk4_stage" : "rk5_stage", fused_rk_stage(c, s, m));
        }
        c->det_valid = false;
    }
    return 0;
}

int do_steps(llg_ctx* c, uint64_t t0, uint32_t nsteps) {
    if (c->scheme != LLG_HEUN) {
        if (c->prec != LLG_FP64_PLAIN) SLLG_TRY(sync_spins(c));  // RK stages read and write S
        c->g3_valid = false;  // RK stages write S without its gather copy
        c->y_soa_valid = true;
    }
    for (uint32_t k = 0; k < nsteps; ++k) {
        if (c->prec == LLG_FP64_PLAIN)
            SLLG_TRY(step_plain(c, t0 + k));
        else
            SLLG_TRY(step_fused(c, t0 + k));
    }
    return 0;
}

}  // namespace
}  // namespace sllg

using namespace sllg;

extern "C" {

uint32_t llg_abi_version(void) { return LLG_ABI_VERSION; }
const char* llg_last_error(void) { return sllg::last_error(); }

int llg_device_count(int* count) {
    if (!count) SLLG_FAIL(LLG_E_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return 0;
}

void llg_desc_init(llg_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->abi_version = LLG_ABI_VERSION;
    d->update = 1;
    d->scheme = LLG_HEUN;
    d->precision = LLG_FP32;
    d->noise = LLG_NOISE_COUNTER;
    d->n_sublattice = 1;
    d->exch.ncomp = 1;
}

static int create_impl(const llg_desc* d, const std::shared_ptr<ExchStore>& share, llg_ctx** out) {
    if (!d || !out) SLLG_FAIL(LLG_E_INVALID, "descriptor or output pointer is NULL");
    *out = nullptr;
    if (share && (share->n != d->n || share->prec != d->precision || share->device != d->device))
        SLLG_FAIL(LLG_E_INVALID, "shared couplings differ in n, precision or device");
    const bool proto = (bool)share;
    if (d->abi_version != LLG_ABI_VERSION) SLLG_FAIL(LLG_E_INVALID, "abi_version %u != %u", d->abi_version, LLG_ABI_VERSION);
    if (d->n == 0) SLLG_FAIL(LLG_E_INVALID, "n must be positive");
    if (d->n >= 0x7FFFFFFFu) SLLG_FAIL(LLG_E_UNSUPPORTED, "n above 2^31-1");
    if (d->n_species == 0 || d->n_species > LLG_MAX_SPECIES || !d->species)
        SLLG_FAIL(LLG_E_INVALID, "n_species must be 1..%u with a species table", LLG_MAX_SPECIES);
    const uint32_t nsub = d->n_sublattice ? d->n_sublattice : 1;
    if (nsub > kMaxSub) SLLG_FAIL(LLG_E_UNSUPPORTED, "more than %d sublattices", kMaxSub);
    if (!(d->dt > 0) || !std::isfinite(d->dt)) SLLG_FAIL(LLG_E_INVALID, "dt must be positive");
    if (d->update == 0) SLLG_FAIL(LLG_E_INVALID, "update must be >= 1");
    if (d->scheme < LLG_HEUN || d->scheme > LLG_RK5) SLLG_FAIL(LLG_E_INVALID, "unknown scheme %d", d->scheme);
    if (d->precision != LLG_FP64_PLAIN && d->precision != LLG_FP32) SLLG_FAIL(LLG_E_INVALID, "unknown precision %d", d->precision);
    if (d->noise != LLG_NOISE_COUNTER && d->noise != LLG_NOISE_INJECTED) SLLG_FAIL(LLG_E_INVALID, "unknown noise mode %d", d->noise);
    for (uint32_t s = 0; s < d->n_species; ++s) {
        const llg_species& p = d->species[s];
        if (!(p.mu > 0) || !(p.lambda >= 0) || !(p.gamma_rel > 0))
            SLLG_FAIL(LLG_E_INVALID, "species %u: mu and gamma_rel must be positive, lambda >= 0", s);
        if (p.sublattice >= nsub) SLLG_FAIL(LLG_E_INVALID, "species %u: sublattice %u >= n_sublattice", s, p.sublattice);
        const double nn = p.k1u_dir[0] * p.k1u_dir[0] + p.k1u_dir[1] * p.k1u_dir[1] + p.k1u_dir[2] * p.k1u_dir[2];
        if (p.k1u != 0.0 && std::fabs(nn - 1.0) > 1e-6) SLLG_FAIL(LLG_E_INVALID, "species %u: k1u axis is not a unit vector", s);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) SLLG_FAIL(LLG_E_NODEVICE, "no HIP device visible");
    if (d->device < 0 || d->device >= ndev) SLLG_FAIL(LLG_E_INVALID, "device %d not in [0, %d)", d->device, ndev);
    SLLG_HIP(hipSetDevice(d->device));

    llg_ctx* c = new llg_ctx();
    struct Guard {
        llg_ctx*& c;
        ~Guard() { if (c) llg_destroy(c); }
    } guard{c};
    c->device = d->device;
    SLLG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->n = d->n;
    c->ns = d->n_species;
    c->nsub = nsub;
    c->scheme = d->scheme;
    c->prec = d->precision;
    c->noise_mode = d->noise;
    c->seed = d->seed;
    c->stream_id = d->stream;
    c->update = d->update;
    c->flags = d->flags;
    c->dt = d->dt;
    c->rdt = d->dt * LLG_GAMMA_E;  // src/llg.cpp:306
    c->anis_in = d->anis;
    c->basemm = d->species[0].mu * LLG_MU_B;  // src/cuint.cu:56
    const uint32_t n = c->n;
    c->perm.na = 1;
    c->perm.ncell = n;
    if (proto) {
        c->xs = share;  // shared, immutable couplings
    } else {
        c->xs = std::make_shared<ExchStore>();
        c->xs->device = d->device;
        c->xs->prec = d->precision;
        c->xs->n = n;
        c->xs->perm = c->perm;
    }

    // per-species constants: src/llg.cpp:591-592, src/geom.cpp:244
    c->sp64.resize(c->ns);
    c->mu_sp.resize(c->ns);
    c->sub_sp.resize(c->ns);
    std::vector<Species<float>> sp32(c->ns);
    for (uint32_t s = 0; s < c->ns; ++s) {
        const llg_species& p = d->species[s];
        Species<double>& q = c->sp64[s];
        q.sigma = std::sqrt(2.0 * LLG_K_B * p.lambda / (p.mu * LLG_MU_B * d->dt * p.gamma_rel * LLG_GAMMA_E));
        q.llgpf = -1. / (1.0 + p.lambda * p.lambda);
        q.lambda = p.lambda;
        q.k1u = 2.0 * p.k1u / (LLG_MU_B * p.mu);
        q.kc = 2.0 * d->anis.kc / (LLG_MU_B * p.mu);
        for (int a = 0; a < 3; ++a) q.k1dir[a] = p.k1u_dir[a];
        c->mu_sp[s] = p.mu;
        c->sub_sp[s] = p.sublattice;
        Species<float>& f = sp32[s];
        f.sigma = (float)q.sigma; f.llgpf = (float)q.llgpf; f.lambda = (float)q.lambda;
        f.k1u = (float)q.k1u; f.kc = (float)q.kc;
        for (int a = 0; a < 3; ++a) f.k1dir[a] = (float)q.k1dir[a];
    }
    make_anis<double>(d->anis, c->basemm, c->an64);
    make_anis<float>(d->anis, c->basemm, c->an32);
    std::vector<uint32_t> spec(n);
    std::vector<uint8_t> spec8(n), sub8(n);
    c->sub_count.assign(nsub, 0);
    bool multi = false;
    for (uint32_t i = 0; i < n; ++i) {
        spec[i] = d->spin_species ? d->spin_species[i] : (i % c->ns);
        if (spec[i] >= c->ns) SLLG_FAIL(LLG_E_INVALID, "spin %u: species %u >= n_species", i, spec[i]);
        if (spec[i] != 0) multi = true;
        spec8[i] = (uint8_t)spec[i];
        sub8[i] = (uint8_t)c->sub_sp[spec[i]];
        c->sub_count[sub8[i]]++;
    }
    c->multi_species = multi;
    SLLG_TRY(c->sp32.alloc(sp32.size() * sizeof(Species<float>)));
    SLLG_HIP(hipMemcpy(c->sp32.p, sp32.data(), sp32.size() * sizeof(Species<float>), hipMemcpyHostToDevice));
    SLLG_TRY(c->sp64d.alloc(c->sp64.size() * sizeof(Species<double>)));
    SLLG_HIP(hipMemcpy(c->sp64d.p, c->sp64.data(), c->sp64.size() * sizeof(Species<double>), hipMemcpyHostToDevice));

    const size_t pb = prec_bytes(c);
    const size_t v3 = 3 * (size_t)n * pb;
    SLLG_TRY(c->S.alloc(v3));
    SLLG_TRY(c->Y.alloc(v3));
    SLLG_TRY(c->STG.alloc(v3));
    if (c->prec == LLG_FP64_PLAIN || c->noise_mode == LLG_NOISE_INJECTED)
        SLLG_TRY(c->XI.alloc(3 * (size_t)n * sizeof(double)));
    if (c->prec == LLG_FP64_PLAIN) {
        SLLG_TRY(c->H.alloc(v3));
        SLLG_TRY(c->FN.alloc(v3));
        SLLG_TRY(c->DET.alloc(v3));
        // per-spin parameter arrays, as the reference uploads them (src/cuda_util.cu:282-556)
        std::vector<double> sg(n), lp(n), lm(n), k1(n), kd(3 * (size_t)n), kc(n);
        bool any_kc = false;
        for (uint32_t i = 0; i < n; ++i) {
            const Species<double>& q = c->sp64[spec[i]];
            sg[i] = q.sigma; lp[i] = q.llgpf; lm[i] = q.lambda; k1[i] = q.k1u; kc[i] = q.kc;
            any_kc = any_kc || q.kc != 0.0;
            for (int a = 0; a < 3; ++a) kd[(size_t)a * n + i] = q.k1dir[a];
        }
        DevBuf* bufs[5] = {&c->p_sigma, &c->p_llgpf, &c->p_lambda, &c->p_k1u, &c->p_kc};
        std::vector<double>* srcs[5] = {&sg, &lp, &lm, &k1, &kc};
        for (int k = 0; k < 5; ++k) {
            if (k == 4 && !any_kc) break;
            SLLG_TRY(bufs[k]->alloc(n * 8));
            SLLG_HIP(hipMemcpy(bufs[k]->p, srcs[k]->data(), n * 8, hipMemcpyHostToDevice));
        }
        SLLG_TRY(c->p_k1udir.alloc(3 * (size_t)n * 8));
        SLLG_HIP(hipMemcpy(c->p_k1udir.p, kd.data(), 3 * (size_t)n * 8, hipMemcpyHostToDevice));
        if (c->scheme != LLG_HEUN) {
            SLLG_TRY(plain_rk_init_tables());
            for (int m = 0; m < 6; ++m) SLLG_TRY(c->K[m].alloc(v3));
        }
        if (!proto) SLLG_TRY(build_exchange_fp64(c, d->exch));
    } else {
        SLLG_TRY(c->A.alloc(v3));
        if (c->scheme == LLG_RK4) SLLG_TRY(c->K[0].alloc(v3));
        if (c->scheme == LLG_RK5) {
            SLLG_TRY(fused_rk_init_tables());
            for (int m = 0; m < 5; ++m) SLLG_TRY(c->K[m].alloc(v3));
        }
        if (!proto) SLLG_TRY(build_exchange_fp32(c, d->exch));
    }
    if (!proto && d->four.nq) SLLG_TRY(build_fourspin(c, d->four));
    c->ex32 = c->xs->ex32;
    c->ex64 = c->xs->ex64;
    c->perm = c->xs->perm;
    // per-spin species / sublattice bytes in the device order chosen by the encoder
    {
        std::vector<uint8_t> s8(n), u8(n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t dv = c->perm.to_dev(i);
            s8[dv] = spec8[i];
            u8[dv] = sub8[i];
        }
        SLLG_TRY(c->spec.alloc(n));
        SLLG_HIP(hipMemcpy(c->spec.p, s8.data(), n, hipMemcpyHostToDevice));
        SLLG_TRY(c->sub.alloc(n));
        SLLG_HIP(hipMemcpy(c->sub.p, u8.data(), n, hipMemcpyHostToDevice));
    }
    const bool fftx = d->fftx.n_int > 0;
    if (fftx) {
        if (d->exch.format != LLG_EXCH_NONE)
            SLLG_FAIL(LLG_E_INVALID, "FFT exchange replaces the exchange SpMV: exch.format must be LLG_EXCH_NONE");
        if (!d->fftx.ints) SLLG_FAIL(LLG_E_INVALID, "FFT exchange without interactions");
        if (!d->dip.mesh) SLLG_FAIL(LLG_E_INVALID, "FFT exchange needs the mesh geometry of desc.dip (dim, nk, abc, mesh)");
        for (uint32_t q = 0; q < d->fftx.n_int; ++q)
            if (d->fftx.ints[q].s1 >= c->ns || d->fftx.ints[q].s2 >= c->ns)
                SLLG_FAIL(LLG_E_INVALID, "FFT exchange interaction %u: species out of range", q);
    }
    if (d->dip.enabled || fftx) {
        if (!d->dip.mesh) SLLG_FAIL(LLG_E_INVALID, "dipolar enabled without mesh coordinates");
        c->dip = true;
        // FFT exchange: the convolution is the exchange field, evaluated at every stage
        c->dip_every_stage = (fftx || d->dip.every_stage) ? 1 : 0;
        c->dip_refresh = fftx ? 1 : (d->dip.refresh ? d->dip.refresh : d->update);
        SLLG_TRY(c->DEM.alloc(v3));
        SLLG_HIP(hipMemset(c->DEM.p, 0, v3));
        SLLG_TRY(dipolar_create(c, d->dip, d->spin_species, fftx ? &d->fftx : nullptr, d->dip.enabled != 0));
    }
    SLLG_HIP(hipMemset(c->STG.p, 0, v3));
    // initial state: all spins along +z
    {
        std::vector<double> zx(n, 0.0), zz(n, 1.0);
        SLLG_TRY(upload_state(c, c->S, zx.data(), zx.data(), zz.data()));
        SLLG_TRY(upload_state(c, c->Y, zx.data(), zx.data(), zz.data()));
    }
    SLLG_HIP(hipDeviceSynchronize());
    *out = c;
    c = nullptr;  // disarm guard
    return 0;
}

int llg_create(const llg_desc* d, llg_ctx** out) { return create_impl(d, nullptr, out); }

int llg_create_like(const llg_desc* d, llg_ctx* proto, llg_ctx** out) {
    if (!proto || !proto->xs) SLLG_FAIL(LLG_E_INVALID, "NULL prototype");
    return create_impl(d, proto->xs, out);
}

void llg_destroy(llg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dipolar_destroy(c);
    delete c->timer;
    c->timer = nullptr;
    hipStream_t s = c->own_stream ? c->stream : nullptr;
    delete c;
    if (s) (void)hipStreamDestroy(s);
}

int llg_set_spins(llg_ctx* c, const double* sx, const double* sy, const double* sz) {
    if (!c || !sx || !sy || !sz) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    for (uint32_t i = 0; i < c->n; ++i) {
        const double m = sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i];
        if (!(m > 0.0) || !std::isfinite(m)) SLLG_FAIL(LLG_E_INVALID, "spin %u has zero or non-finite length", i);
    }
    SLLG_HIP(hipSetDevice(c->device));
    SLLG_TRY(upload_state(c, c->S, sx, sy, sz));
    c->det_valid = false;
    return 0;
}

int llg_get_spins(llg_ctx* c, double* sx, double* sy, double* sz) {
    if (!c || !sx || !sy || !sz) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    if (c->prec != LLG_FP64_PLAIN) SLLG_TRY(sync_spins(c));
    return download_state(c, c->S.p, c->prec == LLG_FP64_PLAIN, sx, sy, sz);
}

int llg_set_temperature(llg_ctx* c, double T) {
    if (!c || !(T >= 0.0) || !std::isfinite(T)) SLLG_FAIL(LLG_E_INVALID, "temperature must be finite and >= 0");
    c->T = T;
    return 0;
}

int llg_set_applied(llg_ctx* c, const double B[3]) {
    if (!c || !B) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(B[a])) SLLG_FAIL(LLG_E_INVALID, "applied field not finite");
        c->B[a] = B[a];
    }
    return 0;
}

int llg_set_staggered(llg_ctx* c, const double* hx, const double* hy, const double* hz) {
    if (!c) SLLG_FAIL(LLG_E_INVALID, "NULL handle");
    SLLG_HIP(hipSetDevice(c->device));
    const uint32_t n = c->n;
    if (!hx && !hy && !hz) {
        c->stag_set = false;
        c->STG0.release();
        SLLG_HIP(hipMemsetAsync(c->STG.p, 0, c->STG.bytes, c->stream));
        SLLG_HIP(hipStreamSynchronize(c->stream));
        return 0;
    }
    if (!hx || !hy || !hz) SLLG_FAIL(LLG_E_INVALID, "staggered field: give all three components or none");
    // H0 stays on the device in FP64 (device order) for llg_scale_staggered; H = 1 * H0
    const double* src[3] = {hx, hy, hz};
    std::vector<double> h0(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const size_t d = c->perm.to_dev(i);
        for (int a = 0; a < 3; ++a) h0[(size_t)a * n + d] = src[a][i];
    }
    if (c->STG0.bytes != h0.size() * 8) SLLG_TRY(c->STG0.alloc(h0.size() * 8));
    SLLG_HIP(hipMemcpyAsync(c->STG0.p, h0.data(), h0.size() * 8, hipMemcpyHostToDevice, c->stream));
    SLLG_TRY(stag_scale(c, 1.0));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    c->stag_set = true;
    return 0;
}

// cullg::rampStag / cufields::CincStagFields (src/cuda.cu:38-41, src/cufields.cu:46-56):
// H = r H0, one kernel on the handle's stream (asynchronous, like the reference's launch).
int llg_scale_staggered(llg_ctx* c, double r) {
    if (!c || !std::isfinite(r)) SLLG_FAIL(LLG_E_INVALID, "bad argument");
    if (!c->stag_set || !c->STG0.p) SLLG_FAIL(LLG_E_STATE, "no staggered field set");
    SLLG_HIP(hipSetDevice(c->device));
    return stag_scale(c, r);
}

int llg_push_noise(llg_ctx* c, uint32_t nsteps, const double* xi) {
    if (!c || (nsteps && !xi)) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (c->noise_mode != LLG_NOISE_INJECTED) SLLG_FAIL(LLG_E_STATE, "handle is not in injected-noise mode");
    const size_t m = 3 * (size_t)c->n;
    for (uint32_t s = 0; s < nsteps; ++s) c->noise_q.emplace_back(xi + s * m, xi + (s + 1) * m);
    return 0;
}

int llg_step(llg_ctx* c, uint64_t t0, uint32_t nsteps) {
    if (!c) SLLG_FAIL(LLG_E_INVALID, "NULL handle");
    SLLG_HIP(hipSetDevice(c->device));
    SLLG_TRY(do_steps(c, t0, nsteps));
    if (c->flags & LLG_FLAG_CHECK_FINITE) {
        int bad = 0;
        SLLG_TRY(check_finite(c, &bad));
        if (bad) SLLG_FAIL(LLG_E_NONFINITE, "non-finite spin after steps %llu..%llu", (unsigned long long)t0,
                           (unsigned long long)(t0 + nsteps - 1));
    }
    return 0;
}

int llg_sync(llg_ctx* c) {
    if (!c) SLLG_FAIL(LLG_E_INVALID, "NULL handle");
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

static int field_at(llg_ctx* c, const void* Sp, double* hx, double* hy, double* hz) {
    DevBuf out;
    SLLG_TRY(out.alloc(3 * (size_t)c->n * 8));
    // FFT exchange is the exchange field: evaluate it at these spins (every step refreshes it
    // before use, so the buffer may be overwritten); a dipolar field keeps its refresh schedule
    if (c->dip && dipolar_is_exchange(c)) SLLG_TRY(dipolar_refresh(c, Sp, c->DEM.p));
    if (c->prec == LLG_FP64_PLAIN)
        SLLG_TRY(plain_detfield(c, (const double*)Sp, out.as<double>()));
    else
        SLLG_TRY(fused_detfield(c, (const float*)Sp, out.as<double>()));
    return download_state(c, out.p, true, hx, hy, hz);
}

int llg_get_det_field(llg_ctx* c, double* hx, double* hy, double* hz) {
    if (!c || !hx || !hy || !hz) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    if (c->det_valid && c->prec == LLG_FP64_PLAIN) return download_state(c, c->DET.p, true, hx, hy, hz);
    if (c->det_valid) {
        SLLG_TRY(fused_y_soa(c));
        return field_at(c, c->Y.p, hx, hy, hz);
    }
    if (c->prec != LLG_FP64_PLAIN) SLLG_TRY(sync_spins(c));
    return field_at(c, c->S.p, hx, hy, hz);
}

int llg_get_field(llg_ctx* c, double* hx, double* hy, double* hz) {
    if (!c || !hx || !hy || !hz) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    if (c->prec != LLG_FP64_PLAIN) SLLG_TRY(sync_spins(c));
    return field_at(c, c->S.p, hx, hy, hz);
}

int llg_reduce(llg_ctx* c, double* m, double* energy) {
    if (!c || !m) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    DevBuf out;
    SLLG_TRY(out.alloc((c->nsub * 3 + 1) * sizeof(double)));
    SLLG_TRY(reduce_mag(c, out.as<double>()));
    if (energy) {
        DevBuf dem;
        if (c->dip) {
            SLLG_TRY(dem.alloc(3 * (size_t)c->n * prec_bytes(c)));
            SLLG_TRY(dipolar_refresh(c, c->S.p, dem.p));
        }
        SLLG_TRY(reduce_energy(c, c->dip ? dem.p : nullptr, out.as<double>() + c->nsub * 3));
        SLLG_HIP(hipStreamSynchronize(c->stream));
    }
    std::vector<double> h(c->nsub * 3 + 1);
    SLLG_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(m, h.data(), c->nsub * 3 * 8);
    if (energy) *energy = h[c->nsub * 3];
    return 0;
}

int llg_reduce_mag_noalloc(llg_ctx* c, double* m) {
    if (!c || !m) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    return reduce_mag_noalloc(c, m);
}

int llg_set_bins(llg_ctx* c, uint32_t nbins, const uint32_t* bin, const double* weight) {
    if (!c || (!bin && nbins)) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    SLLG_HIP(hipSetDevice(c->device));
    return set_bins(c, nbins, bin, weight);
}

int llg_reduce_bins(llg_ctx* c, double* out) {
    if (!c || !out) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (c->nbins == 0) SLLG_FAIL(LLG_E_STATE, "no bins set (llg_set_bins)");
    SLLG_HIP(hipSetDevice(c->device));
    DevBuf d;
    SLLG_TRY(d.alloc((size_t)c->nbins * 3 * sizeof(double)));
    SLLG_TRY(reduce_bins(c, d.as<double>()));
    SLLG_HIP(hipMemcpyAsync(out, d.p, (size_t)c->nbins * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int llg_set_sq(llg_ctx* c, const uint32_t grid[3], const uint32_t* mesh, uint32_t nk, const uint32_t* kpts, int axis,
               const double* ua, const double* ub, uint32_t ngroups, const uint32_t* group) {
    if (!c || !grid || !mesh || (nk && !kpts) || ngroups == 0) SLLG_FAIL(LLG_E_INVALID, "bad argument");
    if (axis < 0 || axis > 2) SLLG_FAIL(LLG_E_INVALID, "quantisation axis must be 0, 1 or 2");
    for (int d = 0; d < 3; ++d)
        if (grid[d] == 0) SLLG_FAIL(LLG_E_INVALID, "empty grid");
    for (uint64_t q = 0; q < (uint64_t)c->n * 3; ++q)
        if (mesh[q] >= grid[q % 3]) SLLG_FAIL(LLG_E_INVALID, "mesh coordinate of spin %llu outside the grid",
                                              (unsigned long long)(q / 3));
    for (uint64_t q = 0; q < (uint64_t)nk * 3; ++q)
        if (kpts[q] >= grid[q % 3]) SLLG_FAIL(LLG_E_INVALID, "k-point %llu outside the grid", (unsigned long long)(q / 3));
    SLLG_HIP(hipSetDevice(c->device));
    return set_sq(c, grid, mesh, nk, kpts, axis, ua, ub, ngroups, group);
}

int llg_sample_sq(llg_ctx* c, double* out) {
    if (!c || !out) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (c->sq_nk == 0) SLLG_FAIL(LLG_E_STATE, "no k-points set (llg_set_sq)");
    SLLG_HIP(hipSetDevice(c->device));
    const size_t cnt = (size_t)c->sq_ngroups * c->sq_nk * 2;
    DevBuf d;
    SLLG_TRY(d.alloc(cnt * sizeof(double)));
    SLLG_TRY(sample_sq(c, d.as<double>()));
    SLLG_HIP(hipMemcpyAsync(out, d.p, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int llg_run_sampled(llg_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t every, double* m_series) {
    if (!c || !m_series || every == 0) SLLG_FAIL(LLG_E_INVALID, "bad argument");
    SLLG_HIP(hipSetDevice(c->device));
    const uint32_t nrec = (nsteps + every - 1) / every;
    const size_t rec = (size_t)c->nsub * 3;
    DevBuf ser;
    SLLG_TRY(ser.alloc(std::max<size_t>(1, nrec) * rec * sizeof(double)));
    uint32_t r = 0;
    for (uint32_t s = 0; s < nsteps; s += every) {
        SLLG_TRY(reduce_mag(c, ser.as<double>() + (size_t)r * rec));
        ++r;
        SLLG_TRY(do_steps(c, t0 + s, std::min(every, nsteps - s)));
    }
    SLLG_HIP(hipMemcpyAsync(m_series, ser.p, (size_t)nrec * rec * 8, hipMemcpyDeviceToHost, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int llg_noise_sample(llg_ctx* c, uint64_t t, double* xi) {
    if (!c || !xi) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (c->noise_mode != LLG_NOISE_COUNTER) SLLG_FAIL(LLG_E_STATE, "handle is not in counter-noise mode");
    SLLG_HIP(hipSetDevice(c->device));
    DevBuf out;
    SLLG_TRY(out.alloc(3 * (size_t)c->n * 8));
    SLLG_TRY(noise_sample(c, t, out.as<double>()));
    return download_state(c, out.p, true, xi, xi + c->n, xi + 2 * (size_t)c->n);
}

void* llg_stream(llg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int llg_profile_kernels(llg_ctx* c, uint64_t t0, uint32_t nsteps, int max_kinds, char* names, double* mean_ms,
                        int* n_kinds) {
    if (!c || !names || !mean_ms || !n_kinds || max_kinds <= 0) SLLG_FAIL(LLG_E_INVALID, "bad argument");
    SLLG_HIP(hipSetDevice(c->device));
    delete c->timer;
    c->timer = new KernelTimer();
    int rc = do_steps(c, t0, nsteps);
    (void)hipStreamSynchronize(c->stream);
    KernelTimer* tm = c->timer;
    c->timer = nullptr;
    if (rc) { delete tm; return rc; }
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
    for (auto& r : tm->recs) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        if (!acc.count(r.name)) order.push_back(r.name);
        acc[r.name].first += ms;
        acc[r.name].second += 1;
    }
    delete tm;
    int k = 0;
    for (auto& nm : order) {
        if (k >= max_kinds) break;
        std::snprintf(names + 48 * k, 48, "%s", nm.c_str());
        mean_ms[k] = acc[nm].first / acc[nm].second;
        ++k;
    }
    *n_kinds = k;
    return 0;
}

int llg_get_info(llg_ctx* c, llg_info* info) {
    if (!c || !info) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    size_t dev = 0;
    DevBuf* all[] = {&c->S, &c->Y, &c->A, &c->H, &c->FN, &c->DET, &c->XI, &c->STG, &c->STG0, &c->DEM,
                     &c->K[0], &c->K[1], &c->K[2], &c->K[3], &c->K[4], &c->K[5], &c->spec, &c->sub,
                     &c->sp32, &c->sp64d, &c->p_sigma, &c->p_llgpf, &c->p_lambda, &c->p_k1u, &c->p_k1udir,
                     &c->p_kc, &c->red_part, &c->red_out};
    for (DevBuf* b : all) dev += b->bytes;
    info->device_bytes = dev + dipolar_bytes(c) + (c->xs ? c->xs->bytes() : 0);
    std::snprintf(info->exch_encoding, sizeof info->exch_encoding, "%s", c->xs ? c->xs->encoding.c_str() : "none");
    const double n = (double)c->n;
    static const HostExch empty;
    const HostExch& h = c->xs ? c->xs->h : empty;
    const int nc = h.ncomp;
    // coupling bytes per row per field evaluation (M)
    double M = 0.0;
    if (c->prec == LLG_FP64_PLAIN) {
        if (h.fmt == LLG_EXCH_CSR) M = 4.0 + (double)h.nnz / n * (4.0 + 8.0 * nc);
        else if (h.fmt == LLG_EXCH_DIA) M = (double)h.ndiag * 8.0 * nc;
        else if (h.fmt == LLG_EXCH_STENCIL && h.per_row) M = (double)h.sk * 8.0 * nc;
    } else {
        const Exch<float>& x = c->ex32;
        if (x.fmt == FMT_SELL_PACK || is_line_fmt(x.fmt)) M = (double)x.nent / n * 4.0;
        else if (x.fmt == FMT_SELL_RAW) M = (double)x.nent / n * (4.0 + 4.0 * nc);
        else if (x.fmt == FMT_DIA) M = (double)h.ndiag * 4.0 * nc;
        else if (x.fmt == FMT_STENCIL && h.per_row) M = (double)h.sk * 4.0 * nc;
    }
    const double P = (double)prec_bytes(c);
    const double extra = (c->dip ? 3 * P : 0) + (c->stag_set ? 3 * P : 0) + (c->multi_species && c->prec != LLG_FP64_PLAIN ? 1 : 0) +
                         (c->noise_mode == LLG_NOISE_INJECTED && c->T > 0 ? 24 : 0);
    if (c->prec == LLG_FP64_PLAIN) {
        const double params = 5 * 8.0 + 3 * 8.0;  // sigma, llgpf, lambda, k1u (+kc) and k1udir
        const double noise = c->T > 0 ? 3 * 8.0 : 0.0;
        const double spmv = M + 3 * P + (c->dip ? 3 * P : 0);
        if (c->scheme == LLG_HEUN) {
            const double h1 = 3 * P + 3 * P + params + noise + 6 * P + (c->stag_set ? 3 * P : 0);
            const double h2 = 3 * P * 3 + params + noise + 6 * P + (c->stag_set ? 3 * P : 0);
            info->bytes_per_spin_update = 2 * spmv + h1 + h2 + (c->T > 0 && c->noise_mode == LLG_NOISE_COUNTER ? noise : 0);
            info->dominant_kernel_bytes_per_spin = spmv;
            std::snprintf(info->dominant_kernel, sizeof info->dominant_kernel, "k_plain_spmv");
        } else {
            const int ns = c->scheme == LLG_RK4 ? 4 : 6;
            info->bytes_per_spin_update = ns * (spmv + 6 * P + params + noise + 3 * P) + 6 * P + ns * 3 * P;
            info->dominant_kernel_bytes_per_spin = spmv;
            std::snprintf(info->dominant_kernel, sizeof info->dominant_kernel, "k_plain_spmv");
        }
    } else {
        if (c->scheme == LLG_HEUN) {
            info->dominant_kernel_bytes_per_spin = 9 * P + M + extra;
            info->bytes_per_spin_update = 2 * info->dominant_kernel_bytes_per_spin;
            std::snprintf(info->dominant_kernel, sizeof info->dominant_kernel, "k_heun_corr");
        } else if (c->scheme == LLG_RK4) {
            info->bytes_per_spin_update = 51 * P + 4 * (M + extra);
            info->dominant_kernel_bytes_per_spin = 15 * P + M + extra;
            std::snprintf(info->dominant_kernel, sizeof info->dominant_kernel, "k_rk4");
        } else {
            // stage m: read Yin, S (m>0), K_0..K_{m-1} (m>0), write K_m, Yout; last: read S,Yin,K0,K2,K3, write S
            double b = 0;
            for (int m = 0; m < 5; ++m) b += 3 * P * (1 + (m > 0) + m + 2);
            b += 3 * P * 6;
            info->bytes_per_spin_update = b + 6 * (M + extra);
            info->dominant_kernel_bytes_per_spin = 3 * P * 7 + M + extra;
            std::snprintf(info->dominant_kernel, sizeof info->dominant_kernel, "k_rk5");
        }
    }
    return 0;
}

}  // extern "C"

namespace sllg {
// Internal: create a handle over an existing coupling store (batch.hip, serving).
int create_with_store(const llg_desc* d, const std::shared_ptr<ExchStore>& store, llg_ctx** out) {
    return create_impl(d, store, out);
}
}  // namespace sllg

// ---------------------------------------------------------------- launch-shape autotuner

namespace {
std::mutex g_tune_mu;
struct TuneChoice {
    int tpb, xcd, unr, thc, g3, line;
};
std::map<std::string, TuneChoice> g_tune_cache;

std::string tune_key(const llg_ctx* c) {
    hipDeviceProp_t p{};
    (void)hipGetDeviceProperties(&p, c->device);
    char buf[256];
    // what changes the kernels that run: the thermal field (drawn only at T > 0 with counter
    // noise), the dipolar field operand, the four-spin term, a per-spin species index, and a
    // reader pinned by the caller
    const int thermal = c->T > 0.0 && c->noise_mode == LLG_NOISE_COUNTER;
    std::snprintf(buf, sizeof buf, "%s|%d|%d|%u|%llu|%u|T%d|D%d|F%d|M%d|R%d", p.gcnArchName, c->ex32.fmt,
                  c->ex32.ncomp, c->n, (unsigned long long)c->ex32.nent, c->perm.na, thermal, c->dip ? 1 : 0,
                  c->ex32.fs_xadj ? 1 : 0, c->multi_species ? 1 : 0,
                  (c->flags & LLG_FLAG_TUNE_FIXED_READER) ? 1 + c->tune_line : 0);
    return buf;
}

void tune_file_load() {
    const char* f = std::getenv("SLLG_TUNE_FILE");
    if (!f) return;
    FILE* fp = std::fopen(f, "r");
    if (!fp) return;
    char key[256];
    TuneChoice t{};
    while (std::fscanf(fp, "%255s %d %d %d %d %d %d", key, &t.tpb, &t.xcd, &t.unr, &t.thc, &t.g3, &t.line) == 7)
        g_tune_cache[key] = t;
    std::fclose(fp);
}

void tune_file_append(const std::string& key, const TuneChoice& t) {
    const char* f = std::getenv("SLLG_TUNE_FILE");
    if (!f) return;
    FILE* fp = std::fopen(f, "a");
    if (!fp) return;
    std::fprintf(fp, "%s %d %d %d %d %d %d\n", key.c_str(), t.tpb, t.xcd, t.unr, t.thc, t.g3, t.line);
    std::fclose(fp);
}
}  // namespace

extern "C" int llg_set_launch(llg_ctx* c, int tpb, int xcd, int unroll, int thermal_cache, int gather_copy,
                              int line_mode) {
    if (!c || (tpb != 256 && tpb != 512 && tpb != 1024)) SLLG_FAIL(LLG_E_INVALID, "block size must be 256, 512 or 1024");
    if (unroll != 2 && unroll != 4) SLLG_FAIL(LLG_E_INVALID, "unroll must be 2 or 4");
    c->tune_tpb = tpb;
    c->tune_xcd = xcd ? 1 : 0;
    c->tune_unr = unroll;
    c->tune_thc = thermal_cache ? 1 : 0;
    c->tune_g3 = gather_copy ? 1 : 0;
    c->tune_line = line_mode == 2 ? 2 : (line_mode ? 1 : 0);
    return 0;
}

extern "C" int llg_autotune(llg_ctx* c, uint32_t nsteps, llg_tune_result* out) {
    if (!c || !out || nsteps == 0) SLLG_FAIL(LLG_E_INVALID, "bad argument");
    std::memset(out, 0, sizeof(*out));
    out->tpb = c->tune_tpb;
    out->xcd = c->tune_xcd;
    out->unroll = c->tune_unr;
    out->thermal_cache = c->tune_thc;
    out->gather_copy = c->tune_g3;
    out->line_mode = c->tune_line;
    if (c->prec != LLG_FP32 || c->scheme != LLG_HEUN) return 0;
    SLLG_HIP(hipSetDevice(c->device));
    const std::string key = tune_key(c);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        static bool loaded = false;
        if (!loaded) {
            tune_file_load();
            loaded = true;
        }
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            c->tune_tpb = out->tpb = it->second.tpb;
            c->tune_xcd = out->xcd = it->second.xcd;
            c->tune_unr = out->unroll = it->second.unr;
            c->tune_thc = out->thermal_cache = it->second.thc;
            c->tune_g3 = out->gather_copy = it->second.g3;
            c->tune_line = out->line_mode = it->second.line;
            out->from_cache = 1;
            return 0;
        }
    }
    // keep the state: time on a copy of the spins (and of the dipolar field cached between
    // refreshes, which the timed steps overwrite), restore afterwards
    DevBuf save, save_dem;
    SLLG_TRY(sync_spins(c));
    SLLG_TRY(save.alloc(c->S.bytes));
    SLLG_HIP(hipMemcpyAsync(save.p, c->S.p, c->S.bytes, hipMemcpyDeviceToDevice, c->stream));
    if (c->dip) {
        SLLG_TRY(save_dem.alloc(c->DEM.bytes));
        SLLG_HIP(hipMemcpyAsync(save_dem.p, c->DEM.p, c->DEM.bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    const TuneChoice caller{c->tune_tpb, c->tune_xcd, c->tune_unr, c->tune_thc, c->tune_g3, c->tune_line};
    const int tpbs[3] = {256, 512, 1024};
    // line encodings: tile reader (when the encoding has tile streams), line reader, or the
    // packed entries kept beside them (unless the caller pinned the reader)
    std::vector<int> readers;
    const bool lfmt = is_line_fmt(c->ex32.fmt);
    if (lfmt && !(c->flags & LLG_FLAG_TUNE_FIXED_READER)) {
        if (c->ex32.tnw) readers.push_back(2);
        readers.push_back(1);
        readers.push_back(0);
    } else {
        readers.push_back(c->tune_line);
    }
    const int unrs[2] = {2, 4};
    // handing the thermal field over only matters at T > 0 with counter noise
    const int nthc = (c->T > 0.0 && c->noise_mode == LLG_NOISE_COUNTER) ? 2 : 1;
    // the AoS3 gather copy applies to the packed and line encodings
    const int ng3 = (c->ex32.fmt == FMT_SELL_PACK || lfmt) ? 2 : 1;
    const int rd_default = lfmt ? 1 : c->tune_line;
    double best = 1e300;
    TuneChoice b{256, 1, 2, c->tune_thc, c->tune_g3, c->tune_line};
    std::vector<std::pair<double, TuneChoice>> timed;  // every candidate's first timing
    hipEvent_t e0, e1;
    SLLG_HIP(hipEventCreate(&e0));
    SLLG_HIP(hipEventCreate(&e1));
    const int noise = c->noise_mode;
    c->noise_mode = LLG_NOISE_COUNTER;  // no host queue traffic while timing
    int rc = 0;
    for (size_t li = 0; li < readers.size() && !rc; ++li) {
    const int rd = readers[li];
    // the tile reader has one block size (its workgroup's slices) and stages from the AoS3
    // copy; the gather depth applies to the packed reader only
    const bool packed = lfmt ? rd == 0 : c->ex32.fmt == FMT_SELL_PACK;
    for (int gi = 0; gi < (rd == 2 ? 1 : ng3) && !rc; ++gi)
    for (int hi = 0; hi < nthc && !rc; ++hi)
    for (int ui = 0; ui < (packed ? 2 : 1) && !rc; ++ui)
        for (int ti = 0; ti < (rd == 2 ? 1 : 3) && !rc; ++ti)
            for (int xcd = 1; xcd >= 0 && !rc; --xcd) {
                const int tpb = rd == 2 ? (int)(64 * c->ex32.tnw) : tpbs[ti];
                c->tune_tpb = tpb;
                c->tune_xcd = xcd;
                c->tune_unr = unrs[ui];
                if (nthc == 2) c->tune_thc = 1 - hi;
                if (ng3 == 2) c->tune_g3 = rd == 2 ? 1 : 1 - gi;
                c->tune_line = rd;
                rc = do_steps(c, 0, 2);
                if (rc) break;
                (void)hipEventRecord(e0, c->stream);
                rc = do_steps(c, 2, nsteps);
                (void)hipEventRecord(e1, c->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double per = ms / nsteps;
                if (tpb == 256 && xcd == 1 && unrs[ui] == 2 && hi == 0 && gi == 0 && rd == rd_default)
                    out->ms_per_step_default = per;
                out->candidates++;
                timed.push_back({per, TuneChoice{tpb, xcd, unrs[ui], c->tune_thc, c->tune_g3, c->tune_line}});
                if (per < best) {
                    best = per;
                    b = timed.back().second;
                }
            }
    }
    // The minimum of many short noisy timings favours a lucky candidate: the fastest few are
    // timed again, in turn, over four times as many steps, and the best of those is kept.
    if (!rc && timed.size() > 1) {
        std::sort(timed.begin(), timed.end(),
                  [](const std::pair<double, TuneChoice>& x, const std::pair<double, TuneChoice>& y) {
                      return x.first < y.first;
                  });
        const size_t top = std::min<size_t>(4, timed.size());
        std::vector<double> acc(top, 0.0);
        for (int pass = 0; pass < 2 && !rc; ++pass)
            for (size_t k = 0; k < top && !rc; ++k) {
                const TuneChoice& t = timed[k].second;
                c->tune_tpb = t.tpb;
                c->tune_xcd = t.xcd;
                c->tune_unr = t.unr;
                c->tune_thc = t.thc;
                c->tune_g3 = t.g3;
                c->tune_line = t.line;
                rc = do_steps(c, 0, 2);
                if (rc) break;
                (void)hipEventRecord(e0, c->stream);
                rc = do_steps(c, 2, 2 * nsteps);
                (void)hipEventRecord(e1, c->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                acc[k] += ms / (2.0 * nsteps);
            }
        if (!rc) {
            size_t kb = 0;
            for (size_t k = 1; k < top; ++k)
                if (acc[k] < acc[kb]) kb = k;
            b = timed[kb].second;
            best = acc[kb] / 2.0;
        }
    }
    c->noise_mode = noise;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SLLG_HIP(hipMemcpyAsync(c->S.p, save.p, c->S.bytes, hipMemcpyDeviceToDevice, c->stream));
    if (c->dip) SLLG_HIP(hipMemcpyAsync(c->DEM.p, save_dem.p, c->DEM.bytes, hipMemcpyDeviceToDevice, c->stream));
    SLLG_HIP(hipStreamSynchronize(c->stream));
    c->det_valid = false;
    c->g3_valid = false;
    c->s_soa_valid = true;
    if (rc) {  // the caller's launch shape stays in force
        c->tune_tpb = caller.tpb;
        c->tune_xcd = caller.xcd;
        c->tune_unr = caller.unr;
        c->tune_thc = caller.thc;
        c->tune_g3 = caller.g3;
        c->tune_line = caller.line;
        return rc;
    }
    c->tune_tpb = out->tpb = b.tpb;
    c->tune_xcd = out->xcd = b.xcd;
    c->tune_unr = out->unroll = b.unr;
    c->tune_thc = out->thermal_cache = b.thc;
    c->tune_g3 = out->gather_copy = b.g3;
    c->tune_line = out->line_mode = b.line;
    out->ms_per_step = best;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = b;
    tune_file_append(key, b);
    return 0;
}
