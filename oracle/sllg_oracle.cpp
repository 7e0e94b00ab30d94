// sllg_oracle.cpp - FP64 CPU oracle of the stochastic-LLG time step.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker / the timed CPU baseline.  The
// product library never links or calls it.
//
// Parity contract with the FP64 GPU plain path: every floating-point operation
// below is performed in the same order as in the HIP kernels, both sides are
// compiled with -ffp-contract=off, and only correctly rounded operations
// (+ - * / sqrt) touch the state, so at T = 0 (and at T > 0 with injected noise)
// the two produce the same bits.  The order itself follows the reference GPU
// kernels: field sum CH + CHstg + thermal + applied (src/cuint.cu:72), uniaxial
// (:77-83), K2 perp/par (:98-110), K4 per axis perp then par (:117-137), torque
// (:142-150), Euler + renormalise (:152-161); corrector (:216-236).
//
// Parity status per function (what pins it) is listed in the tests that call it;
// the functions themselves cite the reference lines they restate.
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

namespace {

const double kGammaE = LLG_GAMMA_E;  // src/llg.cpp:25
const double kMuB = LLG_MU_B;
const double kKB = LLG_K_B;

struct Ctx {
    llg_desc d;
    uint32_t n = 0, opts = 0, nsub = 0;
    std::vector<uint32_t> spec, sub;        // per spin
    std::vector<double> mu_sp, lam_sp, sigma_sp, llgpf_sp, k1u_sp, kc_sp, k1dir_sp;
    std::vector<uint32_t> sub_sp;
    double basemm = 0, rdt = 0;
    llg_anisotropy an;
    // exchange
    int fmt = 0, ncomp = 1;
    std::vector<uint32_t> xadj, adj;
    std::vector<double> cvals;  // [ncomp][nnz]
    uint64_t nnz = 0;
    std::vector<int32_t> doff;
    std::vector<double> dvals;  // [ncomp][ndiag][n]
    uint32_t ndiag = 0;
    uint32_t dims[3] = {0, 0, 0}, na = 1, pbc[3] = {0, 0, 0}, sk = 0, per_row = 0;
    std::vector<uint32_t> scount;
    std::vector<int32_t> sent;
    std::vector<double> svals;
    // four-spin quartets (j, k, l) per row, jq in tesla
    std::vector<uint32_t> fs_xadj, fs_jkl;
    double fs_jq = 0.0;
    int fs_cpu = 0;
    // dipolar
    bool dip = false;
    int every_stage = 0;
    uint32_t refresh = 1;
    std::vector<double> pos;  // [n][3] metres
    std::vector<double> mu_spin;
    // state (SoA)
    std::vector<double> S[3], Y[3], Hx[3], DET[3], DEM[3], STG[3], STG0[3], FN[3], XI[3];
    std::vector<double> K[6][3];
    double T = 0, B[3] = {0, 0, 0};
    bool det_valid = false;
    std::deque<std::vector<double>> noise;
};

// Philox4x32-10, Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11.
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void counter_normals(uint64_t seed, uint64_t stream, uint32_t spin, uint64_t t, double z[3]) {
    uint32_t ctr[4] = {spin, (uint32_t)t, (uint32_t)(t >> 32), (uint32_t)stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32_10(ctr, key, o);
    // Box-Muller on u in (0,1): u = (x + 0.5) 2^-32.
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double u0 = ((double)o[0] + 0.5) * s, u1 = ((double)o[1] + 0.5) * s;
    const double u2 = ((double)o[2] + 0.5) * s, u3 = ((double)o[3] + 0.5) * s;
    const double r0 = std::sqrt(-2.0 * std::log(u0));
    const double r1 = std::sqrt(-2.0 * std::log(u2));
    const double twopi = 6.283185307179586;
    z[0] = r0 * std::cos(twopi * u1);
    z[1] = r0 * std::sin(twopi * u1);
    z[2] = r1 * std::cos(twopi * u3);
}

// Exchange sum for row i from spin arrays (sx, sy, sz): cufields::CSpMV_CSR (src/cufields.cu:180-206),
// CSpMV_CSR_OffDiag (:208-246), CSpMV_DIA (:97-134), CSpMV_DIA_OffDiag (:136-178).
// dot starts at 0 and each neighbour adds val[co]*S[co] (diagonal) or, for the full
// tensor, val[co][ro]*S[ro] for ro = 0,1,2 in turn.
inline void accum(int ncomp, const double* v, const double s[3], double dot[3]) {
    if (ncomp == 1) {
        dot[0] += v[0] * s[0];
        dot[1] += v[0] * s[1];
        dot[2] += v[0] * s[2];
    } else if (ncomp == 3) {
        dot[0] += v[0] * s[0];
        dot[1] += v[1] * s[1];
        dot[2] += v[2] * s[2];
    } else {
        for (int co = 0; co < 3; ++co)
            for (int ro = 0; ro < 3; ++ro) dot[co] += v[co * 3 + ro] * s[ro];
    }
}

void exch_row(const Ctx& c, uint32_t i, const std::vector<double>* sp, double dot[3]) {
    dot[0] = dot[1] = dot[2] = 0.0;
    double v[9];
    if (c.fmt == LLG_EXCH_CSR) {
        for (uint32_t q = c.xadj[i]; q < c.xadj[i + 1]; ++q) {
            const uint32_t j = c.adj[q];
            for (int k = 0; k < c.ncomp; ++k) v[k] = c.cvals[(size_t)k * c.nnz + q];
            const double s[3] = {sp[0][j], sp[1][j], sp[2][j]};
            accum(c.ncomp, v, s, dot);
        }
    } else if (c.fmt == LLG_EXCH_DIA) {
        for (uint32_t q = 0; q < c.ndiag; ++q) {
            const int64_t col = (int64_t)i + c.doff[q];
            if (col >= 0 && col < (int64_t)c.n) {
                for (int k = 0; k < c.ncomp; ++k)
                    v[k] = c.dvals[((size_t)k * c.ndiag + q) * c.n + i];
                const double s[3] = {sp[0][col], sp[1][col], sp[2][col]};
                accum(c.ncomp, v, s, dot);
            }
        }
    } else if (c.fmt == LLG_EXCH_STENCIL) {
        const uint32_t t = i % c.na;
        uint32_t rest = i / c.na;
        const int32_t z = rest % c.dims[2];
        rest /= c.dims[2];
        const int32_t y = rest % c.dims[1];
        const int32_t x = rest / c.dims[1];
        const int32_t cell[3] = {x, y, z};
        for (uint32_t e = 0; e < c.scount[t]; ++e) {
            const int32_t* en = &c.sent[((size_t)t * c.sk + e) * 4];
            int32_t nc[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                int32_t p = cell[a] + en[a];
                const int32_t D = (int32_t)c.dims[a];
                if (p < 0 || p >= D) {
                    if (!c.pbc[a]) { ok = false; break; }
                    p = ((p % D) + D) % D;
                }
                nc[a] = p;
            }
            if (!ok) continue;
            const uint32_t j = (((uint32_t)nc[0] * c.dims[1] + nc[1]) * c.dims[2] + nc[2]) * c.na + en[3];
            for (int k = 0; k < c.ncomp; ++k) {
                if (c.per_row)
                    v[k] = c.svals[((size_t)k * c.sk + e) * c.n + i];
                else
                    v[k] = c.svals[((size_t)k * c.na + t) * c.sk + e];
            }
            const double s[3] = {sp[0][j], sp[1][j], sp[2][j]};
            accum(c.ncomp, v, s, dot);
        }
    }
}

// Field at spin i with own spin s: src/cuint.cu:72-137 (+ cubic, no reference counterpart).
// hex = demag + exchange sum (CH), stg = staggered field, th = thermal field.
inline void local_field(const Ctx& c, uint32_t i, const double s[3], const double hex[3],
                        const double stg[3], const double th[3], double h[3]) {
    for (int a = 0; a < 3; ++a) h[a] = ((hex[a] + stg[a]) + th[a]) + c.B[a];
    const uint32_t sp = c.spec[i];
    const double* d = &c.k1dir_sp[sp * 3];
    const double k1u = c.k1u_sp[sp];
    const double sdn = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
    for (int a = 0; a < 3; ++a) h[a] += (k1u * sdn * d[a]);
    const llg_anisotropy& an = c.an;
    {
        const double pf = 2.0 * an.k2perp / c.basemm;
        const double sd = s[0] * an.k2perp_dir[0] + s[1] * an.k2perp_dir[1] + s[2] * an.k2perp_dir[2];
        for (int a = 0; a < 3; ++a) h[a] += pf * sd * an.k2perp_dir[a];
    }
    {
        const double pf = 2.0 * an.k2par / c.basemm;
        const double sd = s[0] * an.k2par_dir[0] + s[1] * an.k2par_dir[1] + s[2] * an.k2par_dir[2];
        for (int a = 0; a < 3; ++a) h[a] += pf * sd * an.k2par_dir[a];
    }
    {
        const double pfp = 2.0 * an.k4perp / c.basemm;
        const double pfa = 2.0 * an.k4par / c.basemm;
        for (int dir = 0; dir < 3; ++dir) {
            const double* dp = &an.k4perp_dirs[dir * 3];
            const double* da = &an.k4par_dirs[dir * 3];
            const double sdp = s[0] * dp[0] + s[1] * dp[1] + s[2] * dp[2];
            const double sda = s[0] * da[0] + s[1] * da[1] + s[2] * da[2];
            for (int a = 0; a < 3; ++a) h[a] += pfp * sdp * dp[a];
            for (int a = 0; a < 3; ++a) h[a] += pfa * sda * da[a];
        }
    }
    const double kc = c.kc_sp[sp];
    for (int a = 0; a < 3; ++a) h[a] += kc * (s[a] * s[a] * s[a]);
}

// LLG right-hand side f = llgpf (S x h + lambda S x (S x h)) (src/cuint.cu:142-150).
inline void torque(double llgpf, double lambda, const double s[3], const double h[3], double f[3]) {
    const double sxh[3] = {s[1] * h[2] - s[2] * h[1], s[2] * h[0] - s[0] * h[2], s[0] * h[1] - s[1] * h[0]};
    const double sxsxh[3] = {s[1] * sxh[2] - s[2] * sxh[1], s[2] * sxh[0] - s[0] * sxh[2],
                             s[0] * sxh[1] - s[1] * sxh[0]};
    for (int a = 0; a < 3; ++a) f[a] = llgpf * (sxh[a] + lambda * sxsxh[a]);
}

// Brute-force dipolar field, fields::bfdip (src/fields.cpp:199-251), including its
// 1e-7 mu_B prefactor (src/fields.cpp:243) and mu_j weighting (:238).
void dipolar_bf(const Ctx& c, const std::vector<double>* sp, std::vector<double>* out) {
    for (uint32_t i = 0; i < c.n; ++i) {
        double h[3] = {0, 0, 0};
        const double* ri = &c.pos[i * 3];
        for (uint32_t j = 0; j < c.n; ++j) {
            if (j == i) continue;
            const double* rj = &c.pos[j * 3];
            const double r[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]};
            const double mr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            const double oo = 1.0 / (mr * mr * mr);
            const double e[3] = {r[0] / mr, r[1] / mr, r[2] / mr};
            const double sj[3] = {sp[0][j], sp[1][j], sp[2][j]};
            const double sde = sj[0] * e[0] + sj[1] * e[1] + sj[2] * e[2];
            for (int a = 0; a < 3; ++a)
                h[a] += (3.0 * sde * e[a] * c.mu_spin[j] - sj[a] * c.mu_spin[j]) * oo;
        }
        for (int a = 0; a < 3; ++a) out[a][i] = h[a] * 1e-7 * kMuB;
    }
}

void refresh_dip(Ctx& c, const std::vector<double>* sp) {
    if (c.dip) dipolar_bf(c, sp, c.DEM);
}

// Thermal field for step t into XI (xi only; scaled by sqrt(T) sigma at use).
int fill_noise(Ctx& c, uint64_t t) {
    if (c.d.noise == LLG_NOISE_INJECTED) {
        if (c.noise.empty()) return LLG_E_STATE;
        const std::vector<double>& v = c.noise.front();
        for (int a = 0; a < 3; ++a) std::memcpy(c.XI[a].data(), v.data() + (size_t)a * c.n, c.n * sizeof(double));
        c.noise.pop_front();
    } else {
        for (uint32_t i = 0; i < c.n; ++i) {
            double z[3];
            counter_normals(c.d.seed, c.d.stream, i, t, z);
            c.XI[0][i] = z[0]; c.XI[1][i] = z[1]; c.XI[2][i] = z[2];
        }
    }
    return 0;
}

inline void thermal(const Ctx& c, uint32_t i, double th[3]) {
    // TP = sqrt(T) sigma; lrn = xi * TP (src/cuint.cu:66-69)
    const double TP = std::sqrt(c.T) * c.sigma_sp[c.spec[i]];
    th[0] = c.XI[0][i] * TP;
    th[1] = c.XI[1][i] * TP;
    th[2] = ((c.opts & OR_COMPAT_LLGCPU) ? c.XI[0][i] : c.XI[2][i]) * TP;
}

// Four-spin field of row i, cufields::CSpMV_CSR_FourSpin (src/cufields.cu:59-90): from zero,
// each quartet subtracts (JQ (S_j (S_k.S_l) + S_k (S_j.S_l) + S_l (S_k.S_j))) / 3; the
// reference CPU form (src/llgCPU.cpp:163-165) has no division.
void fourspin_row(const Ctx& c, uint32_t i, const std::vector<double>* sp, double h[3]) {
    h[0] = h[1] = h[2] = 0.0;
    for (uint32_t q = c.fs_xadj[i]; q < c.fs_xadj[i + 1]; ++q) {
        const uint32_t j = c.fs_jkl[3 * (size_t)q], k = c.fs_jkl[3 * (size_t)q + 1], l = c.fs_jkl[3 * (size_t)q + 2];
        const double sj[3] = {sp[0][j], sp[1][j], sp[2][j]};
        const double sk[3] = {sp[0][k], sp[1][k], sp[2][k]};
        const double sl[3] = {sp[0][l], sp[1][l], sp[2][l]};
        const double skdotsl = sk[0] * sl[0] + sk[1] * sl[1] + sk[2] * sl[2];
        const double sjdotsl = sj[0] * sl[0] + sj[1] * sl[1] + sj[2] * sl[2];
        const double skdotsj = sk[0] * sj[0] + sk[1] * sj[1] + sk[2] * sj[2];
        for (int a = 0; a < 3; ++a) {
            double v = c.fs_jq * (sj[a] * skdotsl + sk[a] * sjdotsl + sl[a] * skdotsj);
            if (!c.fs_cpu) v = v / 3.0;
            h[a] -= v;
        }
    }
}

// CH = HDemag + exchange (src/cufields.cu:201-203), then CH += four-spin (src/cuda.cu:143-146).
inline void exch_plus_demag(const Ctx& c, uint32_t i, const std::vector<double>* sp, double hex[3]) {
    double dot[3];
    exch_row(c, i, sp, dot);
    for (int a = 0; a < 3; ++a) hex[a] = (c.dip ? c.DEM[a][i] : 0.0) + dot[a];
    if (!c.fs_xadj.empty()) {
        double h4[3];
        fourspin_row(c, i, sp, h4);
        for (int a = 0; a < 3; ++a) hex[a] = hex[a] + h4[a];
    }
}

// One Heun step at the GPU semantics: llgGPU (src/cuda.cu:43-258), CHeun1/CHeun2.
int heun_step(Ctx& c, uint64_t t) {
    const uint32_t n = c.n;
    if (c.dip && t % c.refresh == 0) refresh_dip(c, c.S);
    int rc = fill_noise(c, t);
    if (rc) return rc;
    // predictor
    for (uint32_t i = 0; i < n; ++i) {
        double hex[3], th[3], h[3], f[3];
        exch_plus_demag(c, i, c.S, hex);
        thermal(c, i, th);
        const double stg[3] = {c.STG[0][i], c.STG[1][i], c.STG[2][i]};
        const double s[3] = {c.S[0][i], c.S[1][i], c.S[2][i]};
        local_field(c, i, s, hex, stg, th, h);
        const uint32_t sp = c.spec[i];
        torque(c.llgpf_sp[sp], c.lam_sp[sp], s, h, f);
        double es[3], mods = 0.0;
        for (int a = 0; a < 3; ++a) {
            c.FN[a][i] = f[a];
            es[a] = s[a] + f[a] * c.rdt;
            mods += es[a] * es[a];
        }
        const double nrm = std::sqrt(mods);
        for (int a = 0; a < 3; ++a) c.Y[a][i] = es[a] / nrm;
        if (c.opts & OR_COMPAT_LLGCPU)
            for (int a = 0; a < 3; ++a) c.Hx[a][i] = hex[a];
    }
    if (c.dip && c.every_stage) refresh_dip(c, c.Y);
    // corrector
    std::vector<double> nS[3];
    for (int a = 0; a < 3; ++a) nS[a].resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        double hex[3], th[3], h[3], f[3];
        exch_plus_demag(c, i, c.Y, hex);
        if (c.opts & OR_COMPAT_LLGCPU) {
            // A1: the CPU corrector SpMV adds onto the predictor field (src/llgCPU.cpp:253-274)
            double dot[3];
            exch_row(c, i, c.Y, dot);
            for (int a = 0; a < 3; ++a) hex[a] = c.Hx[a][i] + dot[a];
        }
        thermal(c, i, th);
        const double stg[3] = {c.STG[0][i], c.STG[1][i], c.STG[2][i]};
        const double s[3] = {c.Y[0][i], c.Y[1][i], c.Y[2][i]};
        local_field(c, i, s, hex, stg, th, h);
        const uint32_t sp = c.spec[i];
        torque(c.llgpf_sp[sp], c.lam_sp[sp], s, h, f);
        double ps[3] = {c.S[0][i], c.S[1][i], c.S[2][i]};
        double mods = 0.0;
        for (int a = 0; a < 3; ++a) {
            ps[a] += (0.5 * (c.FN[a][i] + f[a]) * c.rdt);
            mods += ps[a] * ps[a];
        }
        const double nrm = std::sqrt(mods);
        for (int a = 0; a < 3; ++a) {
            nS[a][i] = ps[a] / nrm;
            c.DET[a][i] = h[a] - th[a];
        }
    }
    for (int a = 0; a < 3; ++a) c.S[a].swap(nS[a]);
    c.det_valid = true;
    return 0;
}

// Stage increment k = rdt (llgpf (S x h + lambda S x (S x h))) at stage vector y.
void rk_stage(Ctx& c, const std::vector<double>* y, std::vector<double>* k) {
    for (uint32_t i = 0; i < c.n; ++i) {
        double hex[3], th[3], h[3], f[3];
        exch_plus_demag(c, i, y, hex);
        thermal(c, i, th);
        const double stg[3] = {c.STG[0][i], c.STG[1][i], c.STG[2][i]};
        const double s[3] = {y[0][i], y[1][i], y[2][i]};
        local_field(c, i, s, hex, stg, th, h);
        const uint32_t sp = c.spec[i];
        torque(c.llgpf_sp[sp], c.lam_sp[sp], s, h, f);
        for (int a = 0; a < 3; ++a) k[a][i] = c.rdt * f[a];
    }
}

// Textbook RK4 with the reference's combine weights (src/cuintRK4.cu:436-456); stage
// inputs are S + a k (not the defective forms of src/cuintRK4.cu:141-142,172).
int rk4_step(Ctx& c, uint64_t t) {
    const uint32_t n = c.n;
    if (c.dip && t % c.refresh == 0) refresh_dip(c, c.S);
    int rc = fill_noise(c, t);
    if (rc) return rc;
    const double a[3] = {0.5, 0.5, 1.0};
    const std::vector<double>* y = c.S;
    for (int m = 0; m < 4; ++m) {
        if (m > 0 && c.dip && c.every_stage) refresh_dip(c, y);
        rk_stage(c, y, c.K[m]);
        if (m < 3) {
            for (int q = 0; q < 3; ++q)
                for (uint32_t i = 0; i < n; ++i) c.Y[q][i] = c.S[q][i] + a[m] * c.K[m][q][i];
            y = c.Y;
        }
    }
    const double onesixth = 1.0 / 6.0, onethird = 1.0 / 3.0;
    for (uint32_t i = 0; i < n; ++i) {
        double ls[3], mods = 0.0;
        for (int q = 0; q < 3; ++q) {
            ls[q] = c.S[q][i] + c.K[0][q][i] * onesixth + c.K[1][q][i] * onethird +
                    c.K[2][q][i] * onethird + c.K[3][q][i] * onesixth;
            mods += ls[q] * ls[q];
        }
        const double nrm = std::sqrt(mods);
        for (int q = 0; q < 3; ++q) c.S[q][i] = ls[q] / nrm;
    }
    c.det_valid = false;
    return 0;
}

// Cash-Karp RK5 (tableau src/cuintRK5.cu:44-54), 5th-order weights Ci, no error control.
const double kB21 = 0.2;
const double kB3[2] = {3.0 / 40.0, 9.0 / 40.0};
const double kB4[3] = {0.3, -0.9, 1.2};
const double kB5[4] = {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0};
const double kB6[5] = {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0};
const double kC5[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};

int rk5_step(Ctx& c, uint64_t t) {
    const uint32_t n = c.n;
    if (c.dip && t % c.refresh == 0) refresh_dip(c, c.S);
    int rc = fill_noise(c, t);
    if (rc) return rc;
    const double* B[6] = {nullptr, &kB21, kB3, kB4, kB5, kB6};
    for (int m = 0; m < 6; ++m) {
        const std::vector<double>* y = c.S;
        if (m > 0) {
            for (int q = 0; q < 3; ++q)
                for (uint32_t i = 0; i < n; ++i) {
                    double v = c.S[q][i];
                    for (int j = 0; j < m; ++j) v = v + B[m][j] * c.K[j][q][i];
                    c.Y[q][i] = v;
                }
            y = c.Y;
            if (c.dip && c.every_stage) refresh_dip(c, y);
        }
        rk_stage(c, y, c.K[m]);
    }
    for (uint32_t i = 0; i < n; ++i) {
        double ls[3], mods = 0.0;
        for (int q = 0; q < 3; ++q) {
            double v = c.S[q][i];
            for (int j = 0; j < 6; ++j)
                if (kC5[j] != 0.0) v = v + c.K[j][q][i] * kC5[j];
            ls[q] = v;
            mods += ls[q] * ls[q];
        }
        const double nrm = std::sqrt(mods);
        for (int q = 0; q < 3; ++q) c.S[q][i] = ls[q] / nrm;
    }
    c.det_valid = false;
    return 0;
}

void det_field_at(Ctx& c, const std::vector<double>* sp, double* hx, double* hy, double* hz) {
    for (uint32_t i = 0; i < c.n; ++i) {
        double hex[3], h[3];
        const double th[3] = {0.0, 0.0, 0.0};
        exch_plus_demag(c, i, sp, hex);
        const double stg[3] = {c.STG[0][i], c.STG[1][i], c.STG[2][i]};
        const double s[3] = {sp[0][i], sp[1][i], sp[2][i]};
        local_field(c, i, s, hex, stg, th, h);
        hx[i] = h[0]; hy[i] = h[1]; hz[i] = h[2];
    }
}

}  // namespace

struct or_ctx : Ctx {};

extern "C" {

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        philox_round(c, k);
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

void or_counter_normals_f64(uint64_t seed, uint64_t stream, uint32_t spin, uint64_t t, double z[3]) {
    counter_normals(seed, stream, spin, t, z);
}

// MWC1616 (src/random.cpp:17-23)
void or_mwc_seed(or_mwc* r, uint32_t a, uint32_t b) { r->x = a; r->y = b; r->have = 0; r->n2 = 0; }
uint32_t or_mwc_next(or_mwc* r) {
    r->x = 18000u * (r->x & 0xFFFFu) + (r->x >> 16);
    r->y = 30903u * (r->y & 0xFFFFu) + (r->y >> 16);
    return (r->x << 16) + (r->y & 0xFFFFu);
}
// Random::rand (src/random.cpp:37-40)
double or_mwc_uniform(or_mwc* r) { return or_mwc_next(r) * (1.0 / 4294967295.0); }
// Random::normal, Marsaglia polar method (src/random.cpp:42-64)
double or_mwc_normal(or_mwc* r) {
    if (r->have) { r->have = 0; return r->n2; }
    double w, x1, x2;
    const double rmax = 1.0 / 4294967295.0;
    for (;;) {
        const double u = 1.0 - or_mwc_next(r) * rmax;
        const double v = 1.0 - or_mwc_next(r) * rmax;
        x1 = 2.0 * u - 1.0;
        x2 = 2.0 * v - 1.0;
        w = x1 * x1 + x2 * x2;
        if (w < 1.0) break;
    }
    w = std::sqrt((-2.0 * std::log(w)) / w);
    r->n2 = w * x2;
    r->have = 1;
    return w * x1;
}
// spins::setSpinsRandom (src/spins.cpp:265-281)
void or_random_spins(uint32_t a, uint32_t b, uint32_t n, double* sx, double* sy, double* sz) {
    or_mwc r;
    or_mwc_seed(&r, a, b);
    for (uint32_t i = 0; i < n; ++i) {
        double v1 = 0, v2 = 0, s = 2.0;
        while (s > 1.0) {
            v1 = 2.0 * or_mwc_uniform(&r) - 1.0;
            v2 = 2.0 * or_mwc_uniform(&r) - 1.0;
            s = v1 * v1 + v2 * v2;
        }
        const double ss = std::sqrt(1.0 - s);
        sx[i] = 2.0 * v1 * ss;
        sy[i] = 2.0 * v2 * ss;
        sz[i] = 1.0 - 2.0 * s;
    }
}

// exch::exchunitcell (src/exch_unitcell.cpp:292-420): for each spin in (x,y,z,t) order,
// for each neighbour species s1, for each interaction of (t, s1) in list order: neighbour
// cell = cell + offset, wrapped on periodic axes and dropped on open ones; value
// J/(mu_t mu_B) (:399).
int or_build_csr(const llg_unitcell* uc, const uint32_t dims[3], const uint32_t pbc[3],
                 uint32_t n_int, const llg_interaction* ints, const double* mu_atom, int ncomp,
                 uint32_t* xadj, uint32_t* adjncy, double* vals, uint64_t* nnz) {
    if (!uc || !dims || !pbc || !nnz || (n_int && !ints) || !mu_atom) return LLG_E_INVALID;
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) return LLG_E_INVALID;
    const uint32_t na = uc->n_atoms;
    const int32_t D[3] = {(int32_t)dims[0], (int32_t)dims[1], (int32_t)dims[2]};
    // pass 1: count; pass 2 (adjncy != NULL): fill
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && !adjncy) return 0;
        uint64_t q = 0;
        const uint64_t total = *nnz;
        uint32_t idx = 0;
        for (int32_t x = 0; x < D[0]; ++x)
            for (int32_t y = 0; y < D[1]; ++y)
                for (int32_t z = 0; z < D[2]; ++z)
                    for (uint32_t t = 0; t < na; ++t, ++idx) {
                        if (pass == 1 && xadj) xadj[idx] = (uint32_t)q;
                        const int32_t cell[3] = {x, y, z};
                        for (uint32_t s1 = 0; s1 < na; ++s1)
                            for (uint32_t k = 0; k < n_int; ++k) {
                                const llg_interaction& it = ints[k];
                                if (it.atom_i != t || it.atom_j != s1) continue;
                                int32_t lv[3];
                                int ok = 0;
                                for (int a = 0; a < 3; ++a) {
                                    lv[a] = cell[a] + it.dc[a];
                                    if (lv[a] < 0 && pbc[a]) { lv[a] += D[a]; ++ok; }
                                    else if (lv[a] >= D[a] && pbc[a]) { lv[a] -= D[a]; ++ok; }
                                    else if (lv[a] >= 0 && lv[a] < D[a]) ++ok;
                                }
                                if (ok != 3) continue;
                                if (pass == 1) {
                                    adjncy[q] = (((uint32_t)lv[0] * dims[1] + lv[1]) * dims[2] + lv[2]) * na + s1;
                                    const double den = mu_atom[t] * kMuB;
                                    if (ncomp == 1) vals[q] = it.J[0] / den;
                                    else if (ncomp == 3)
                                        for (int a = 0; a < 3; ++a) vals[(size_t)a * total + q] = it.J[a * 4] / den;
                                    else
                                        for (int a = 0; a < 9; ++a) vals[(size_t)a * total + q] = it.J[a] / den;
                                }
                                ++q;
                            }
                    }
        if (pass == 0) {
            *nnz = q;
        } else if (xadj) {
            xadj[idx] = (uint32_t)q;
        }
    }
    return 0;
}

int or_create(const llg_desc* d, uint32_t options, or_ctx** out) {
    if (!d || !out || d->n == 0 || d->n_species == 0 || d->n_species > LLG_MAX_SPECIES || !d->species)
        return LLG_E_INVALID;
    if (d->dt <= 0 || d->update == 0) return LLG_E_INVALID;
    or_ctx* c = new (std::nothrow) or_ctx();
    if (!c) return LLG_E_NOMEM;
    c->d = *d;
    c->n = d->n;
    c->opts = options;
    const uint32_t n = d->n, ns = d->n_species;
    c->rdt = d->dt * kGammaE;  // src/llg.cpp:306
    c->an = d->anis;
    c->nsub = d->n_sublattice ? d->n_sublattice : 1;
    c->mu_sp.resize(ns); c->lam_sp.resize(ns); c->sigma_sp.resize(ns); c->llgpf_sp.resize(ns);
    c->k1u_sp.resize(ns); c->kc_sp.resize(ns); c->k1dir_sp.resize(ns * 3); c->sub_sp.resize(ns);
    for (uint32_t s = 0; s < ns; ++s) {
        const llg_species& p = d->species[s];
        c->mu_sp[s] = p.mu;
        c->lam_sp[s] = p.lambda;
        // src/llg.cpp:591-592; k1u = 2K/(mu_B mu) src/geom.cpp:244
        c->sigma_sp[s] = std::sqrt(2.0 * kKB * p.lambda / (p.mu * kMuB * d->dt * p.gamma_rel * kGammaE));
        c->llgpf_sp[s] = -1. / (1.0 + p.lambda * p.lambda);
        c->k1u_sp[s] = 2.0 * p.k1u / (kMuB * p.mu);
        c->kc_sp[s] = 2.0 * d->anis.kc / (kMuB * p.mu);
        for (int a = 0; a < 3; ++a) c->k1dir_sp[s * 3 + a] = p.k1u_dir[a];
        c->sub_sp[s] = p.sublattice;
    }
    c->basemm = d->species[0].mu * kMuB;  // src/cuint.cu:56
    c->spec.resize(n);
    c->sub.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        c->spec[i] = d->spin_species ? d->spin_species[i] : (i % ns);
        if (c->spec[i] >= ns) { delete c; return LLG_E_INVALID; }
        c->sub[i] = c->sub_sp[c->spec[i]];
        if (c->sub[i] >= c->nsub) { delete c; return LLG_E_INVALID; }
    }
    const llg_exchange& e = d->exch;
    c->fmt = e.format;
    c->ncomp = e.ncomp ? e.ncomp : 1;
    if (c->fmt != LLG_EXCH_NONE && c->ncomp != 1 && c->ncomp != 3 && c->ncomp != 9) { delete c; return LLG_E_INVALID; }
    if (c->fmt == LLG_EXCH_CSR) {
        if (!e.xadj || (e.nnz && (!e.adjncy || !e.csr_vals))) { delete c; return LLG_E_INVALID; }
        c->nnz = e.nnz;
        c->xadj.assign(e.xadj, e.xadj + n + 1);
        c->adj.assign(e.adjncy, e.adjncy + e.nnz);
        c->cvals.assign(e.csr_vals, e.csr_vals + (size_t)c->ncomp * e.nnz);
        if (c->xadj[0] != 0 || c->xadj[n] != e.nnz) { delete c; return LLG_E_INVALID; }
        for (uint32_t i = 0; i < n; ++i)
            if (c->xadj[i + 1] < c->xadj[i]) { delete c; return LLG_E_INVALID; }
        for (uint64_t q = 0; q < e.nnz; ++q)
            if (c->adj[q] >= n) { delete c; return LLG_E_INVALID; }
    } else if (c->fmt == LLG_EXCH_DIA) {
        c->ndiag = e.ndiag;
        c->doff.assign(e.dia_offsets, e.dia_offsets + e.ndiag);
        c->dvals.assign(e.dia_vals, e.dia_vals + (size_t)c->ncomp * e.ndiag * n);
    } else if (c->fmt == LLG_EXCH_STENCIL) {
        for (int a = 0; a < 3; ++a) { c->dims[a] = e.dims[a]; c->pbc[a] = e.pbc[a]; }
        c->na = e.n_atoms;
        c->sk = e.stencil_k;
        c->per_row = e.stencil_vals_per_row;
        if ((uint64_t)e.dims[0] * e.dims[1] * e.dims[2] * e.n_atoms != n) { delete c; return LLG_E_INVALID; }
        c->scount.assign(e.stencil_count, e.stencil_count + c->na);
        c->sent.assign(e.stencil_entries, e.stencil_entries + (size_t)c->na * c->sk * 4);
        const size_t nv = c->per_row ? (size_t)c->ncomp * c->sk * n : (size_t)c->ncomp * c->na * c->sk;
        c->svals.assign(e.stencil_vals, e.stencil_vals + nv);
    } else if (c->fmt != LLG_EXCH_NONE) {
        delete c;
        return LLG_E_INVALID;
    }
    if (d->four.nq) {
        const llg_fourspin& f = d->four;
        if (!f.xadj || !f.jkl) { delete c; return LLG_E_INVALID; }
        c->fs_xadj.assign(f.xadj, f.xadj + n + 1);
        c->fs_jkl.assign(f.jkl, f.jkl + 3 * (size_t)f.nq);
        c->fs_jq = f.jq;
        c->fs_cpu = f.cpu_form;
        if (c->fs_xadj[0] != 0 || c->fs_xadj[n] != f.nq) { delete c; return LLG_E_INVALID; }
        for (uint32_t i = 0; i < n; ++i)
            if (c->fs_xadj[i + 1] < c->fs_xadj[i]) { delete c; return LLG_E_INVALID; }
        for (uint32_t v : c->fs_jkl)
            if (v >= n) { delete c; return LLG_E_INVALID; }
    }
    if (d->dip.enabled) {
        if (!d->dip.mesh) { delete c; return LLG_E_INVALID; }
        c->dip = true;
        c->every_stage = d->dip.every_stage;
        c->refresh = d->dip.refresh ? d->dip.refresh : d->update;
        c->pos.resize((size_t)n * 3);
        c->mu_spin.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a)
                c->pos[i * 3 + a] = d->dip.abc[a] * (double)d->dip.mesh[i * 3 + a] / (double)d->dip.nk[a];
            c->mu_spin[i] = c->mu_sp[c->spec[i]];
        }
    }
    for (int a = 0; a < 3; ++a) {
        c->S[a].assign(n, 0.0); c->Y[a].assign(n, 0.0); c->Hx[a].assign(n, 0.0);
        c->DET[a].assign(n, 0.0); c->DEM[a].assign(n, 0.0); c->STG[a].assign(n, 0.0);
        c->STG0[a].assign(n, 0.0); c->FN[a].assign(n, 0.0); c->XI[a].assign(n, 0.0);
        for (int m = 0; m < 6; ++m) c->K[m][a].assign(n, 0.0);
        c->S[2].assign(n, 1.0);
    }
    *out = c;
    return 0;
}

void or_destroy(or_ctx* c) { delete c; }

int or_set_spins(or_ctx* c, const double* sx, const double* sy, const double* sz) {
    if (!c || !sx || !sy || !sz) return LLG_E_INVALID;
    c->S[0].assign(sx, sx + c->n); c->S[1].assign(sy, sy + c->n); c->S[2].assign(sz, sz + c->n);
    c->det_valid = false;
    return 0;
}
int or_get_spins(or_ctx* c, double* sx, double* sy, double* sz) {
    if (!c || !sx || !sy || !sz) return LLG_E_INVALID;
    std::memcpy(sx, c->S[0].data(), c->n * 8); std::memcpy(sy, c->S[1].data(), c->n * 8);
    std::memcpy(sz, c->S[2].data(), c->n * 8);
    return 0;
}
int or_set_temperature(or_ctx* c, double T) {
    if (!c || !(T >= 0)) return LLG_E_INVALID;
    c->T = T;
    return 0;
}
int or_set_applied(or_ctx* c, const double B[3]) {
    if (!c || !B) return LLG_E_INVALID;
    for (int a = 0; a < 3; ++a) c->B[a] = B[a];
    return 0;
}
int or_set_staggered(or_ctx* c, const double* hx, const double* hy, const double* hz) {
    if (!c) return LLG_E_INVALID;
    const double* h[3] = {hx, hy, hz};
    for (int a = 0; a < 3; ++a) {
        if (h[a]) c->STG0[a].assign(h[a], h[a] + c->n);
        else c->STG0[a].assign(c->n, 0.0);
        c->STG[a] = c->STG0[a];
    }
    return 0;
}
// cufields::CincStagFields (src/cufields.cu:46-56): Hstg = r * Hstg0
int or_scale_staggered(or_ctx* c, double r) {
    if (!c) return LLG_E_INVALID;
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < c->n; ++i) c->STG[a][i] = r * c->STG0[a][i];
    return 0;
}
int or_push_noise(or_ctx* c, uint32_t nsteps, const double* xi) {
    if (!c || (nsteps && !xi)) return LLG_E_INVALID;
    for (uint32_t s = 0; s < nsteps; ++s)
        c->noise.emplace_back(xi + (size_t)s * 3 * c->n, xi + (size_t)(s + 1) * 3 * c->n);
    return 0;
}
int or_step(or_ctx* c, uint64_t t0, uint32_t nsteps) {
    if (!c) return LLG_E_INVALID;
    for (uint32_t s = 0; s < nsteps; ++s) {
        int rc;
        if (c->d.scheme == LLG_RK4) rc = rk4_step(*c, t0 + s);
        else if (c->d.scheme == LLG_RK5) rc = rk5_step(*c, t0 + s);
        else rc = heun_step(*c, t0 + s);
        if (rc) return rc;
    }
    return 0;
}
int or_get_det_field(or_ctx* c, double* hx, double* hy, double* hz) {
    if (!c || !hx || !hy || !hz) return LLG_E_INVALID;
    if (c->det_valid) {
        std::memcpy(hx, c->DET[0].data(), c->n * 8); std::memcpy(hy, c->DET[1].data(), c->n * 8);
        std::memcpy(hz, c->DET[2].data(), c->n * 8);
        return 0;
    }
    det_field_at(*c, c->S, hx, hy, hz);
    return 0;
}
int or_get_field(or_ctx* c, double* hx, double* hy, double* hz) {
    if (!c || !hx || !hy || !hz) return LLG_E_INVALID;
    det_field_at(*c, c->S, hx, hy, hz);
    return 0;
}
int or_get_dipolar(or_ctx* c, double* hx, double* hy, double* hz) {
    if (!c || !c->dip || !hx || !hy || !hz) return LLG_E_INVALID;
    std::vector<double> o[3];
    for (int a = 0; a < 3; ++a) o[a].resize(c->n);
    dipolar_bf(*c, c->S, o);
    std::memcpy(hx, o[0].data(), c->n * 8); std::memcpy(hy, o[1].data(), c->n * 8);
    std::memcpy(hz, o[2].data(), c->n * 8);
    return 0;
}
// util::calc_mag method 0 (src/util_mag.cpp:44-63) and the energy
// E = sum_i mu_i mu_B [ -1/2 S.(H_ex + H_dip + H_quadratic-anisotropy) - 1/4 S.(H_cubic + H_4spin)
//     - S.(B + H_stag) ].
int or_reduce(or_ctx* c, double* m, double* energy) {
    if (!c || !m) return LLG_E_INVALID;
    std::vector<double> acc((size_t)c->nsub * 3, 0.0);
    std::vector<uint64_t> cnt(c->nsub, 0);
    for (uint32_t i = 0; i < c->n; ++i) {
        const uint32_t s = c->sub[i];
        for (int a = 0; a < 3; ++a) acc[s * 3 + a] += c->S[a][i];
        cnt[s]++;
    }
    for (uint32_t s = 0; s < c->nsub; ++s) {
        const double oo = cnt[s] ? 1. / (double)cnt[s] : 0.0;
        for (int a = 0; a < 3; ++a) m[s * 3 + a] = acc[s * 3 + a] * oo;
    }
    if (energy) {
        std::vector<double> dem[3];
        if (c->dip) {
            for (int a = 0; a < 3; ++a) dem[a].resize(c->n);
            dipolar_bf(*c, c->S, dem);
        }
        double E = 0.0;
        const llg_anisotropy& an = c->an;
        for (uint32_t i = 0; i < c->n; ++i) {
            double dot[3];
            exch_row(*c, i, c->S, dot);
            const double s[3] = {c->S[0][i], c->S[1][i], c->S[2][i]};
            const uint32_t sp = c->spec[i];
            const double mm = c->mu_sp[sp] * kMuB;
            double e_pair = s[0] * dot[0] + s[1] * dot[1] + s[2] * dot[2];
            if (c->dip) e_pair += s[0] * dem[0][i] + s[1] * dem[1][i] + s[2] * dem[2][i];
            // quadratic anisotropy: field linear in S
            double hq[3] = {0, 0, 0};
            const double* d = &c->k1dir_sp[sp * 3];
            const double sdn = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
            for (int a = 0; a < 3; ++a) hq[a] += c->k1u_sp[sp] * sdn * d[a];
            const double* dirs[2] = {an.k2perp_dir, an.k2par_dir};
            const double ks[2] = {an.k2perp, an.k2par};
            for (int q = 0; q < 2; ++q) {
                const double sd = s[0] * dirs[q][0] + s[1] * dirs[q][1] + s[2] * dirs[q][2];
                for (int a = 0; a < 3; ++a) hq[a] += 2.0 * ks[q] / c->basemm * sd * dirs[q][a];
            }
            for (int dir = 0; dir < 3; ++dir) {
                const double* dp = &an.k4perp_dirs[dir * 3];
                const double* da = &an.k4par_dirs[dir * 3];
                const double sdp = s[0] * dp[0] + s[1] * dp[1] + s[2] * dp[2];
                const double sda = s[0] * da[0] + s[1] * da[1] + s[2] * da[2];
                for (int a = 0; a < 3; ++a) hq[a] += 2.0 * an.k4perp / c->basemm * sdp * dp[a];
                for (int a = 0; a < 3; ++a) hq[a] += 2.0 * an.k4par / c->basemm * sda * da[a];
            }
            const double e_q = s[0] * hq[0] + s[1] * hq[1] + s[2] * hq[2];
            double e_4 = 0.0;  // quartic in the spins: -1/4 S.H4
            if (!c->fs_xadj.empty()) {
                double h4[3];
                fourspin_row(*c, i, c->S, h4);
                e_4 = s[0] * h4[0] + s[1] * h4[1] + s[2] * h4[2];
            }
            const double kc = c->kc_sp[sp];
            const double e_c = kc * (s[0] * s[0] * s[0] * s[0] + s[1] * s[1] * s[1] * s[1] + s[2] * s[2] * s[2] * s[2]);
            double e_z = 0.0;
            for (int a = 0; a < 3; ++a) e_z += s[a] * (c->B[a] + c->STG[a][i]);
            E += mm * (-0.5 * e_pair - 0.5 * e_q - 0.25 * e_c - 0.25 * e_4 - e_z);
        }
        *energy = E;
    }
    return 0;
}
int or_noise_sample(or_ctx* c, uint64_t t, double* xi) {
    if (!c || !xi) return LLG_E_INVALID;
    for (uint32_t i = 0; i < c->n; ++i) {
        double z[3];
        counter_normals(c->d.seed, c->d.stream, i, t, z);
        xi[i] = z[0]; xi[c->n + i] = z[1]; xi[2 * (size_t)c->n + i] = z[2];
    }
    return 0;
}
int or_run_sampled(or_ctx* c, uint64_t t0, uint32_t nsteps, uint32_t every, double* m_series) {
    if (!c || !m_series || every == 0) return LLG_E_INVALID;
    uint32_t rec = 0;
    for (uint32_t s = 0; s < nsteps; ++s) {
        if (s % every == 0) {
            int rc = or_reduce(c, m_series + (size_t)rec * c->nsub * 3, nullptr);
            if (rc) return rc;
            ++rec;
        }
        int rc = or_step(c, t0 + s, 1);
        if (rc) return rc;
    }
    return 0;
}

}  // extern "C"
