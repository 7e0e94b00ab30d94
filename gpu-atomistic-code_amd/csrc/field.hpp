// field.hpp - effective-field assembly shared by the FP64 plain kernels and the
// FP32 fused kernels.  The order of every floating-point operation is the order of
// the reference GPU kernels (src/cufields.cu:180-246 exchange, src/cuint.cu:72-150
// local terms and torque), so that the FP64 path built with -ffp-contract=off
// reproduces the FP64 oracle bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace sllg {

// A global-memory view of p.  Pointers the kernels read from memory (a batch member's
// buffers and coupling arrays, read from the member table) are generic to the compiler, and
// loads and stores through generic pointers become flat instructions: VGPR addresses only,
// no scalar loads for uniform data, and counted against the LDS wait counter as well.  All
// these buffers are device memory.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gview(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gview(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
// A constant-memory view of p, for device memory that no kernel writes while it runs (slice
// tables, member tables).  A uniform load through a global pointer becomes a scalar load only
// when the compiler can show that nothing in the kernel has written memory before it; a kernel
// that stores, fences or takes atomics in a loop (the persistent batched kernel) loses that, and
// such loads turn into vector loads plus readfirstlane.  Loads from the constant address space
// are scalar whenever the address is uniform.  (The record streams keep the global view: on
// the per-launch kernels it measured slightly faster.)
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cview(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// dot += J s for one neighbour: isotropic (NC 1), diagonal (NC 3), full tensor (NC 9,
// val[co][ro] * s[ro] for ro = 0,1,2 in turn as in src/cufields.cu:233-238).
template <int NC, class R>
__device__ __forceinline__ void accum(const R v[NC], const R s[3], R dot[3]) {
    if constexpr (NC == 1) {
        dot[0] += v[0] * s[0];
        dot[1] += v[0] * s[1];
        dot[2] += v[0] * s[2];
    } else if constexpr (NC == 3) {
        dot[0] += v[0] * s[0];
        dot[1] += v[1] * s[1];
        dot[2] += v[2] * s[2];
    } else {
#pragma unroll
        for (int co = 0; co < 3; ++co)
#pragma unroll
            for (int ro = 0; ro < 3; ++ro) dot[co] += v[co * 3 + ro] * s[ro];
    }
}

// Neighbour index of stencil entry en for spin i (spin order ((x*d1+y)*d2+z)*na+t);
// returns false across a non-periodic face.
template <class V>
__device__ __forceinline__ bool stencil_neighbour(const Exch<V>& e, uint32_t i, int4 en, uint32_t& j) {
    const uint32_t t = i % e.na;
    uint32_t r = i / e.na;
    const int32_t z = (int32_t)(r % e.dims[2]);
    r /= e.dims[2];
    const int32_t y = (int32_t)(r % e.dims[1]);
    const int32_t x = (int32_t)(r / e.dims[1]);
    int32_t p[3] = {x + en.x, y + en.y, z + en.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t D = (int32_t)e.dims[a];
        if (p[a] < 0) {
            if (!e.pbc[a]) return false;
            p[a] += D;
        } else if (p[a] >= D) {
            if (!e.pbc[a]) return false;
            p[a] -= D;
        }
    }
    (void)t;
    j = (((uint32_t)p[0] * e.dims[1] + (uint32_t)p[1]) * e.dims[2] + (uint32_t)p[2]) * e.na + (uint32_t)en.w;
    return true;
}

// Exchange sum of row i in the reference formats (CSR, DIA) and the periodic stencil,
// reading spins from S (SoA, precision SpinT) and couplings of precision V,
// accumulating in R.
// Four-spin field of row i (cufields::CSpMV_CSR_FourSpin, src/cufields.cu:59-90), from zero:
// each quartet subtracts (jq (S_j (S_k.S_l) + S_k (S_j.S_l) + S_l (S_k.S_j))) / 3 (no division
// in the reference CPU form).  Same operation order as the oracle.
template <class R, class V, class SpinT>
__device__ __forceinline__ void fourspin_row(const Exch<V>& e, uint32_t i, const SpinT* __restrict__ S, R h[3]) {
    const uint32_t n = e.n;
    h[0] = h[1] = h[2] = R(0);
    const R jq = (R)e.fs_jq;
    const auto* __restrict__ Sg = gview(S);
    const auto* __restrict__ xq = gview(e.fs_xadj);
    for (uint32_t q = xq[i]; q < xq[i + 1]; ++q) {
        const u32x4 t = gview(reinterpret_cast<const u32x4*>(e.fs_jkl))[q];
        const R sj[3] = {(R)Sg[t.x], (R)Sg[n + t.x], (R)Sg[2 * (size_t)n + t.x]};
        const R sk[3] = {(R)Sg[t.y], (R)Sg[n + t.y], (R)Sg[2 * (size_t)n + t.y]};
        const R sl[3] = {(R)Sg[t.z], (R)Sg[n + t.z], (R)Sg[2 * (size_t)n + t.z]};
        const R skdotsl = sk[0] * sl[0] + sk[1] * sl[1] + sk[2] * sl[2];
        const R sjdotsl = sj[0] * sl[0] + sj[1] * sl[1] + sj[2] * sl[2];
        const R skdotsj = sk[0] * sj[0] + sk[1] * sj[1] + sk[2] * sj[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            R v = jq * (sj[a] * skdotsl + sk[a] * sjdotsl + sl[a] * skdotsj);
            if (e.fs_div3) v = v / R(3);
            h[a] -= v;
        }
    }
}

template <int NC, class R, class V, class SpinT>
__device__ __forceinline__ void exch_row_ref(const Exch<V>& e, uint32_t i, const SpinT* __restrict__ S,
                                             R dot[3]) {
    const uint32_t n = e.n;
    dot[0] = dot[1] = dot[2] = R(0);
    const auto* __restrict__ Sg = gview(S);
    const auto* __restrict__ vals = gview(e.vals);
    R v[NC];
    if (e.fmt == FMT_CSR) {
        // CSR rows stored by 64-row slices, entry q of lane l at [q][l] (encode.hip): the
        // row's entries in CSR order, its own count from the row pointers
        const auto* __restrict__ xadj = gview(e.xadj);
        const uint32_t rl = xadj[i + 1] - xadj[i];
        const size_t o0 = (size_t)gview(reinterpret_cast<const u32x2*>(e.slice))[i >> 6].x * 64u + (i & 63u);
        for (uint32_t q = 0; q < rl; ++q) {
            const size_t o = o0 + (size_t)q * 64u;
            const uint32_t j = gview(e.ent)[o];
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] = (R)vals[(size_t)k * e.nent + o];
            const R s[3] = {(R)Sg[j], (R)Sg[n + j], (R)Sg[2 * (size_t)n + j]};
            accum<NC, R>(v, s, dot);
        }
    } else if (e.fmt == FMT_DIA) {
        for (uint32_t q = 0; q < e.ndiag; ++q) {
            const int64_t col = (int64_t)i + gview(e.doff)[q];
            if (col >= 0 && col < (int64_t)n) {
#pragma unroll
                for (int k = 0; k < NC; ++k) v[k] = (R)vals[((size_t)k * e.ndiag + q) * n + i];
                const R s[3] = {(R)Sg[col], (R)Sg[n + col], (R)Sg[2 * (size_t)n + col]};
                accum<NC, R>(v, s, dot);
            }
        }
    } else if (e.fmt == FMT_STENCIL) {
        const uint32_t t = i % e.na;
        const uint32_t cnt = gview(e.scount)[t];
        for (uint32_t q = 0; q < cnt; ++q) {
            const auto* en4 = gview(e.sent) + (t * e.sk + q);
            const int4 en = make_int4(en4->x, en4->y, en4->z, en4->w);
            uint32_t j;
            if (!stencil_neighbour(e, i, en, j)) continue;
            if (e.per_row) {
#pragma unroll
                for (int k = 0; k < NC; ++k) v[k] = (R)vals[((size_t)k * e.sk + q) * n + i];
            } else {
#pragma unroll
                for (int k = 0; k < NC; ++k) v[k] = (R)vals[((size_t)k * e.na + t) * e.sk + q];
            }
            const R s[3] = {(R)Sg[j], (R)Sg[n + j], (R)Sg[2 * (size_t)n + j]};
            accum<NC, R>(v, s, dot);
        }
    }
}

// Local terms: uniaxial (src/cuint.cu:77-83), K2 perp/par (:98-110), K4 per axis perp then
// par (:117-137), cubic (new).  k1u is in tesla (2K/(mu mu_B)), pf = 2K/(mu_0 mu_B).
template <class R>
__device__ __forceinline__ void local_terms(R k1u, const R d[3], R kc, const AnisG<R>& an, const R s[3], R h[3]) {
    if (k1u != R(0)) {
        const R sdn = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) h[a] += (k1u * sdn * d[a]);
    }
#pragma unroll 1
    for (int q = 0; q < an.nlin; ++q) {  // a uniform loop: no work for absent terms
        const R* dq = an.lin_dir[q];
        const R sd = s[0] * dq[0] + s[1] * dq[1] + s[2] * dq[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) h[a] += an.lin_pf[q] * sd * dq[a];
    }
    if (kc != R(0)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) h[a] += kc * (s[a] * s[a] * s[a]);
    }
}

// f = llgpf (S x h + lambda S x (S x h))  (src/cuint.cu:142-150)
template <class R>
__device__ __forceinline__ void llg_rhs(R llgpf, R lambda, const R s[3], const R h[3], R f[3]) {
    const R sxh[3] = {s[1] * h[2] - s[2] * h[1], s[2] * h[0] - s[0] * h[2], s[0] * h[1] - s[1] * h[0]};
    const R sxsxh[3] = {s[1] * sxh[2] - s[2] * sxh[1], s[2] * sxh[0] - s[0] * sxh[2], s[0] * sxh[1] - s[1] * sxh[0]};
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = llgpf * (sxh[a] + lambda * sxsxh[a]);
}

}  // namespace sllg
