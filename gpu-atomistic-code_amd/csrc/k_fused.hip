// k_fused.hip - FP32 fused path: one kernel per field evaluation that gathers the
// exchange field, adds dipolar/staggered/thermal/applied/anisotropy terms, forms the
// LLG right-hand side and advances the integrator, with |S| = 1 renormalisation.
//
// Heun (the reference CHeun1 + SpMV + CHeun2 sequence, src/cuda.cu:92-233) becomes two
// passes:  predictor  reads S, writes Y = N(S + rdt f(S)) and A = S + rdt/2 f(S);
//          corrector  reads Y, A, writes S = N(A + rdt/2 f(Y)).
// Bytes per spin-update: 18*4 (spins) + 2*M (couplings), see the per-format M in encode.cpp.
// Thermal noise is regenerated from Philox in both passes (0 bytes), so the corrector
// sees the same xi as the predictor (src/cuint.cu:69,174).
//
// RK4 / Cash-Karp RK5 stages (src/cuintRK4.cu, src/cuintRK5.cu) are one pass per stage
// with ping-pong stage inputs, textbook stage formation.
#include <mutex>

#include "field.hpp"
#include "kernels.hpp"
#include "philox.hpp"

namespace sllg {

// Sliced ELL, scalar reader (any precision; dictionary read from global memory).  Slice
// = 64 consecutive device rows (one wavefront), located by the slice table {first, len}.
// Packed entry (pack_word): neighbour row offset and coupling-dictionary index
// [ndict][NC]; a slice's entries are stored [chunk of 4][lane][4] + [lane][2] tail and slices with
// identical entries share one copy.  Raw entry: u32 column, values [NC][nent], both
// [entry][lane].
template <int NC, class R, class V, class SpinT>
__device__ __forceinline__ void exch_row_sell(const Exch<V>& e, uint32_t i, const SpinT* __restrict__ S, R dot[3]) {
    const uint32_t n = e.n;
    dot[0] = dot[1] = dot[2] = R(0);
    const uint32_t sl = i >> 6, lane = i & 63u;
    const u32x2 sd = cview(reinterpret_cast<const u32x2*>(e.slice))[sl];
    const uint32_t q0 = sd.x, len = sd.y;
    const auto* __restrict__ Sg = gview(S);
    R v[NC];
    if (e.fmt == FMT_SELL_PACK) {
        const auto* __restrict__ ent = gview(e.ent) + (size_t)q0 * kSlice;
        for (uint32_t q = 0; q < len; ++q) {
            const uint32_t w = ent[pack_pos(q, lane, len)];
            const uint32_t j = pack_nb(w, i, e.dbits);
            const auto* dv = gview(e.dict) + (size_t)pack_dict(w, e.dbits) * NC;
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] = (R)dv[k];
            const R s[3] = {(R)Sg[j], (R)Sg[n + j], (R)Sg[2 * (size_t)n + j]};
            accum<NC, R>(v, s, dot);
        }
    } else {
        const auto* __restrict__ ent = gview(e.ent) + (size_t)q0 * kSlice + lane;
        const auto* __restrict__ val = gview(e.dict) + (size_t)q0 * kSlice + lane;
#pragma unroll 4
        for (uint32_t q = 0; q < len; ++q) {
            const uint32_t j = ent[(size_t)q * kSlice];
#pragma unroll
            for (int k = 0; k < NC; ++k) v[k] = (R)val[(size_t)k * e.nent + (size_t)q * kSlice];
            const R s[3] = {(R)Sg[j], (R)Sg[n + j], (R)Sg[2 * (size_t)n + j]};
            accum<NC, R>(v, s, dot);
        }
    }
}

// Spins as one raw buffer resource: a neighbour's three components are three buffer loads
// sharing one 32-bit byte offset (scalar component offsets 4n, 8n), so no 64-bit address
// arithmetic per entry; a load outside the 12n bytes returns 0 instead of faulting.
struct Spin3 {  // one spin of an AoS3 copy
    float x, y, z;
};

struct SpinBuf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t o1, o2;
};
__device__ __forceinline__ SpinBuf spin_buf(const float* S, uint32_t n) {
    SpinBuf b;
    b.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(S), 0, (int)(12u * n), 0x00020000);
    b.o1 = 4u * n;
    b.o2 = 8u * n;
    return b;
}

// Neighbour gathers from either copy of the spins: the SoA arrays (three dword loads) or
// the AoS3 gather copy [3 j + c] the Heun kernels keep beside them (one dwordx3 load).
template <bool G3>
__device__ __forceinline__ void gather3(const SpinBuf& sb, uint32_t j, float s[3]) {
    if constexpr (G3) {
        const auto q = __builtin_amdgcn_raw_buffer_load_b96(sb.r, j * 12u, 0, 0);
        s[0] = __uint_as_float(q[0]);
        s[1] = __uint_as_float(q[1]);
        s[2] = __uint_as_float(q[2]);
    } else {
        const uint32_t off = j << 2;
        s[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, off, 0, 0));
        s[1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, off, sb.o1, 0));
        s[2] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, off, sb.o2, 0));
    }
}
// Lane l's spin of the 64-row line starting at device row line0 (line0 wave-uniform).
template <bool G3>
__device__ __forceinline__ void line3(const SpinBuf& sb, uint32_t lane, uint32_t line0, float s[3]) {
    if constexpr (G3) {
        const auto q = __builtin_amdgcn_raw_buffer_load_b96(sb.r, lane * 12u, line0 * 12u, 0);
        s[0] = __uint_as_float(q[0]);
        s[1] = __uint_as_float(q[1]);
        s[2] = __uint_as_float(q[2]);
    } else {
        const uint32_t vo = lane << 2, so = line0 << 2;
        s[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so, 0));
        s[1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so + sb.o1, 0));
        s[2] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so + sb.o2, 0));
    }
}

// One packed entry: decode, gather the neighbour, read its coupling from the LDS dictionary.
template <int NC, bool G3>
__device__ __forceinline__ void pack_entry(uint32_t w, uint32_t i, uint32_t dbits, const SpinBuf& sb,
                                           const float* sdict, float dot[3]) {
    float s[3];
    gather3<G3>(sb, pack_nb(w, i, dbits), s);
    float v[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] = sdict[pack_dict(w, dbits) * NC + k];
    accum<NC, float>(v, s, dot);
}

// Fast packed reader of the fused kernels: one uint4 (4 entries) per lane per load
// (1 KB per wavefront), then at most one uint2 tail; couplings from the block's LDS copy
// of the dictionary.
// UNR chunks (4 UNR entries, 12 UNR gathers) are loaded before any is used, so that many
// neighbour loads are in flight per wave; the launch-shape tuner picks UNR.
template <int NC, int UNR, bool G3>
__device__ __forceinline__ void exch_row_pack_fast(const Exch<float>& e, uint32_t i, const float* __restrict__ S,
                                                   const float* sdict, float dot[3]) {
    dot[0] = dot[1] = dot[2] = 0.0f;
    // rows come in whole waves (row = 64 k + lane in every caller), so the slice, its
    // descriptor and the loop bounds are wave-uniform: scalar loads and uniform branches
    const uint32_t sl = __builtin_amdgcn_readfirstlane(i >> 6), lane = i & 63u;
    const u32x2 sd = cview(reinterpret_cast<const u32x2*>(e.slice))[sl];
    const uint32_t q0 = sd.x, nch = sd.y >> 2;
    const auto* __restrict__ p = gview(reinterpret_cast<const u32x4*>(e.ent + (size_t)q0 * kSlice)) + lane;
    const uint32_t db = e.dbits;
    const SpinBuf sb = spin_buf(S, e.n);
    uint32_t c = 0;
    for (; c + UNR <= nch; c += UNR) {
        u32x4 cur[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) cur[u] = p[(size_t)(c + u) * kSlice];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            pack_entry<NC, G3>(cur[u].x, i, db, sb, sdict, dot);
            pack_entry<NC, G3>(cur[u].y, i, db, sb, sdict, dot);
            pack_entry<NC, G3>(cur[u].z, i, db, sb, sdict, dot);
            pack_entry<NC, G3>(cur[u].w, i, db, sb, sdict, dot);
        }
    }
    for (; c < nch; ++c) {
        const u32x4 cur = p[(size_t)c * kSlice];
        pack_entry<NC, G3>(cur.x, i, db, sb, sdict, dot);
        pack_entry<NC, G3>(cur.y, i, db, sb, sdict, dot);
        pack_entry<NC, G3>(cur.z, i, db, sb, sdict, dot);
        pack_entry<NC, G3>(cur.w, i, db, sb, sdict, dot);
    }
    if (sd.y & 2u) {  // tail of 2 entries per lane, [lane][2]
        const u32x2 t = ((const __attribute__((address_space(1))) u32x2*)(p - lane + (size_t)nch * kSlice))[lane];
        pack_entry<NC, G3>(t.x, i, db, sb, sdict, dot);
        pack_entry<NC, G3>(t.y, i, db, sb, sdict, dot);
    }
}

// Line-mode group record, read with scalar loads by the whole wave (constant indices only,
// so it stays in SGPRs):
//   w[0]               shift mask of the group (bit 0: unshifted, 1: next along the line,
//                      2: previous): the shifts some window of the group has couplings for
//   w[1, 1 + G)        window starts relative to the slice's first row (int32, multiples of
//                      the line length, 64 or 16)
//   w[CO, CO + 3 G NC) couplings [window][shift][component], FP32, zero where a window has
//                      none (CO = 1 + G)
//   w[W - G, W)        the next group's window starts (zero after the last group)
template <int NC>
struct LineRecW {
    static constexpr int G = line_group(NC), W = line_words(NC), NX = W - G, CO = 1 + G;
    uint32_t w[W];
    __device__ uint32_t mask() const { return w[0]; }
};
// Record of the tile reader: w[0, G) LDS byte offsets of the group's windows (the shift mask
// in the low 3 bits of w[0]), w[G, G + 3 G NC) couplings [window][shift][component].
template <int NC>
struct TileRecW {
    static constexpr int G = line_group(NC), W = tile_words(NC), CO = G;
    uint32_t w[W];
    __device__ uint32_t mask() const { return w[0] & 7u; }
    __device__ uint32_t window(int u) const { return u == 0 ? w[0] & ~7u : w[u]; }
};
// Whole 16-byte units, so the record arrives in few wide scalar loads.  CV: through the
// constant view, for kernels whose record loads follow stores, fences or atomics (the
// persistent batched kernel); the per-launch kernels measured slightly faster without it.
template <class RW, bool CV = false>
__device__ __forceinline__ RW load_rec(const uint32_t* __restrict__ p) {
    RW o;
#pragma unroll
    for (int k = 0; k < RW::W / 4; ++k) {
        u32x4 v;
        if constexpr (CV)
            v = cview(reinterpret_cast<const u32x4*>(p))[k];
        else
            v = gview(reinterpret_cast<const u32x4*>(p))[k];
        o.w[4 * k] = v.x;
        o.w[4 * k + 1] = v.y;
        o.w[4 * k + 2] = v.z;
        o.w[4 * k + 3] = v.w;
    }
    return o;
}

// A loaded window stays one 96-bit value across the pipelined loop (an aligned VGPR tuple,
// so the loads write their final registers and need no copies).
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));

template <bool G3>
__device__ __forceinline__ u32x3 line_v(const SpinBuf& sb, uint32_t lane, uint32_t line0) {
    if constexpr (G3) {
        return __builtin_amdgcn_raw_buffer_load_b96(sb.r, lane * 12u, line0 * 12u, 0);
    } else {
        const uint32_t vo = lane << 2, so = line0 << 2;
        u32x3 v;
        v.x = __builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so, 0);
        v.y = __builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so + sb.o1, 0);
        v.z = __builtin_amdgcn_raw_buffer_load_b32(sb.r, vo, so + sb.o2, 0);
        return v;
    }
}

// The G windows whose starts are ws[0, G), loaded into x.
template <int NC, bool G3>
__device__ __forceinline__ void line_loads(const SpinBuf& sb, uint32_t lane, int32_t base, const uint32_t* ws,
                                           u32x3* x) {
#pragma unroll
    for (int u = 0; u < line_group(NC); ++u)
        x[u] = line_v<G3>(sb, lane, __builtin_amdgcn_readfirstlane((uint32_t)(base + (int32_t)ws[u])));
}

// Neighbours along lines of LEN cells: lane l takes lane l + 1 (next) or l - 1 (prev) of
// its line.  Periodic lines wrap within the line: across the whole wave (wave_rol:1 /
// wave_ror:1) for 64-cell lines, within each 16-lane row (row_ror:15 / row_ror:1) for
// 16-cell lines.  Open lines shift instead (wave_shl:1 / wave_shr:1, row_shl:1 /
// row_shr:1) and the lane past a line's end reads 0 (bound_ctrl): no neighbour there.
template <int LEN, bool OPEN>
__device__ __forceinline__ float line_next(float v) {
    constexpr int ctrl = OPEN ? (LEN == 64 ? 0x130 : 0x101) : (LEN == 64 ? 0x134 : 0x12F);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, OPEN));
}
template <int LEN, bool OPEN>
__device__ __forceinline__ float line_prev(float v) {
    constexpr int ctrl = OPEN ? (LEN == 64 ? 0x138 : 0x111) : (LEN == 64 ? 0x13C : 0x121);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, OPEN));
}

// Shift K of one group: every window's couplings for that shift against its unshifted
// spins, into the partial sum of shift K.
template <int NC, int K, class RW>
__device__ __forceinline__ void line_shift_sum(const RW& r, const u32x3* x, float p[3]) {
#pragma unroll
    for (int u = 0; u < RW::G; ++u) {
        const float s[3] = {__uint_as_float(x[u].x), __uint_as_float(x[u].y), __uint_as_float(x[u].z)};
        float v[NC];
#pragma unroll
        for (int a = 0; a < NC; ++a) v[a] = __uint_as_float(r.w[RW::CO + (u * 3 + K) * NC + a]);
        accum<NC, float>(v, s, p);
    }
}

// One group's arithmetic: one uniform test per shift.
template <int NC, class RW>
__device__ __forceinline__ void line_group_sum(const RW& r, const u32x3* x, float p[3][3]) {
    const uint32_t m = r.mask();
    if (m & 1u) line_shift_sum<NC, 0>(r, x, p[0]);
    if (m & 2u) line_shift_sum<NC, 1>(r, x, p[1]);
    if (m & 4u) line_shift_sum<NC, 2>(r, x, p[2]);
}

// Line-mode reader (FMT_SELL_LINE / LINE16 / open variants): per group of G neighbour
// windows, one 64-row window load each.  The shifted neighbours' contributions are summed
// per shift at the unshifted lanes, and the two shifted partial sums are moved to the
// lanes that own them once per row (a shift along the line is linear), so a row costs
// six lane moves whatever its number of windows.  Software pipelined: a group's record
// names the next group's windows, whose loads are issued (with the scalar load of the
// next record) before the current group's arithmetic.  A slice marked kLineFallback is
// read from the packed entries.
// LFB: the encoding has slices marked kLineFallback (compiled in only then: the packed
// reader's registers would otherwise cost the line kernels occupancy).
template <int NC, bool G3, int LEN, bool OPEN, bool LFB, bool CV = false>
__device__ __forceinline__ void exch_row_line(const Exch<float>& e, uint32_t i, const float* __restrict__ S,
                                              const float* sdict, float dot[3]) {
    using RW = LineRecW<NC>;
    constexpr int G = RW::G, W = RW::W, NX = RW::NX;
    const uint32_t sl = __builtin_amdgcn_readfirstlane(i >> 6);
    const u32x2 sd = cview(reinterpret_cast<const u32x2*>(e.lslice))[sl];
    if constexpr (LFB) {
        if (sd.y == kLineFallback) {
            exch_row_pack_fast<NC, 2, G3>(e, i, S, sdict, dot);
            return;
        }
    }
    const uint32_t* __restrict__ rp = e.lent + sd.x;
    const uint32_t ngrp = sd.y, lane = i & 63u;
    const int32_t base = (int32_t)(sl << 6);
    float p[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (ngrp != 0) {
        const SpinBuf sb = spin_buf(S, e.n);
        // two window buffers used in turn (no register copies, so no early waits on the
        // loads in flight); pairs of groups run in the loop, the last one or two after it
        u32x3 xa[G], xb[G];
        RW c0 = load_rec<LineRecW<NC>, CV>(rp);
        line_loads<NC, G3>(sb, lane, base, c0.w + 1, xa);
        uint32_t r = 0;
        for (; r + 2 < ngrp; r += 2) {
            const RW c1 = load_rec<LineRecW<NC>, CV>(rp + (size_t)(r + 1) * W);
            line_loads<NC, G3>(sb, lane, base, c0.w + NX, xb);
            line_group_sum<NC>(c0, xa, p);
            const RW c2 = load_rec<LineRecW<NC>, CV>(rp + (size_t)(r + 2) * W);
            line_loads<NC, G3>(sb, lane, base, c1.w + NX, xa);
            line_group_sum<NC>(c1, xb, p);
            c0 = c2;
        }
        if (r + 1 < ngrp) {
            const RW c1 = load_rec<LineRecW<NC>, CV>(rp + (size_t)(r + 1) * W);
            line_loads<NC, G3>(sb, lane, base, c0.w + NX, xb);
            line_group_sum<NC>(c0, xa, p);
            line_group_sum<NC>(c1, xb, p);
        } else {
            line_group_sum<NC>(c0, xa, p);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) dot[c] = p[0][c] + line_next<LEN, OPEN>(p[1][c]) + line_prev<LEN, OPEN>(p[2][c]);
}

template <int FMT, int NC, class R, class V, class SpinT>
__device__ __forceinline__ void exch_any(const Exch<V>& e, uint32_t i, const SpinT* __restrict__ S, R dot[3]) {
    if constexpr (FMT == FMT_SELL_PACK || FMT == FMT_SELL_RAW)
        exch_row_sell<NC, R, V, SpinT>(e, i, S, dot);
    else if constexpr (FMT == FMT_NONE)
        dot[0] = dot[1] = dot[2] = R(0);
    else
        exch_row_ref<NC, R, V, SpinT>(e, i, S, dot);
}

struct FusedArgs {
    uint32_t n;
    Perm perm;  // device row -> reference spin index (counter noise is keyed by the latter)
    int noise;  // 0 none (T = 0), 1 counter, 2 injected
    int multi;  // per-spin species index present
    int xcd;    // XCD-aware block remap
    uint32_t il_na;  // > 1: sublattice-interleaved tile order (il_na sublattices of il_cells rows)
    uint32_t il_cells;
    int tile_sub;    // every tile lies in one sublattice block (perm.ncell % blockDim.x == 0)
    int na_sh, ncell_sh;  // log2 of perm.na and perm.ncell when powers of two (shifts), else -1
    float* th_out;   // predictor: store the thermal field it drew (or null)
    const float* th_in;  // corrector: read it instead of drawing again (or null)
    float rdt, half_rdt, sqrtT;
    float B[3];
    uint64_t seed, stream, t;
    const double* xi;
    const float* dem;
    const float* stg;
    const uint8_t* spec;
    const Species<float>* sp;
    Species<float> sp0;
    AnisG<float> an;
    Exch<float> ex;
};

// Block index -> row tile.  Consecutive workgroups are dispatched round-robin over the
// 8 XCDs of the MI355X; the remap gives each XCD a contiguous range of rows so the
// neighbour spins it gathers stay in its own L2.
__device__ __forceinline__ uint32_t tile_index(int xcd) {
    const uint32_t b = blockIdx.x;
    if (!xcd) return b;
    const uint32_t nb = gridDim.x, per = nb >> 3, rem = nb & 7u;
    const uint32_t x = b & 7u, l = b >> 3;
    return x * per + (x < rem ? x : rem) + l;
}

// Row tile of this block in the sublattice-major device order.  With the XCD remap each
// XCD takes a contiguous range of logical tiles; logical tile L is mapped to sublattice
// L % na, cell tile L / na, so an XCD's range covers one cell range of every sublattice and
// the neighbours its rows gather (all sublattices, nearby cells) are fetched into its L2
// once, not once per sublattice.  Requires il_cells % blockDim.x == 0 (checked on the host).
template <int TPB>
__device__ __forceinline__ uint32_t row_tile(const FusedArgs& a) {
    uint32_t L = tile_index(a.xcd);
    if (a.il_na > 1) {
        const uint32_t tps = a.il_cells / TPB;
        if (a.na_sh >= 0)
            L = (L & (a.il_na - 1u)) * tps + (L >> a.na_sh);
        else
            L = (L % a.il_na) * tps + L / a.il_na;
    }
    return L;
}

// Device row of this thread and the reference index that keys its noise.  When a tile
// lies inside one sublattice block, the block's sublattice and first cell are uniform
// (scalar arithmetic once per block) and the lane adds its offset; otherwise the lane
// inverts the permutation itself.
template <int TPB>
__device__ __forceinline__ void row_of_thread(const FusedArgs& a, uint32_t& i, uint32_t& iref) {
    const uint32_t d0 = row_tile<TPB>(a) * TPB;
    i = d0 + threadIdx.x;
    if (a.perm.na <= 1) {
        iref = i;
    } else if (a.tile_sub) {
        const uint32_t t = a.ncell_sh >= 0 ? d0 >> a.ncell_sh : d0 / a.perm.ncell;
        iref = (d0 - t * a.perm.ncell + threadIdx.x) * a.perm.na + t;
    } else {
        iref = a.perm.to_ref(i);
    }
}

// Row in range.  In line mode n is a multiple of 64 and tiles start at multiples of 64, so
// a wave is wholly in or out: the test is made on a wave-uniform value, which keeps the
// line reader out of divergent control flow (its record words stay scalar).
template <int FMT>
__device__ __forceinline__ bool row_live(const FusedArgs& a, uint32_t i) {
    if constexpr (is_line_fmt(FMT))
        return __builtin_amdgcn_readfirstlane(i & ~63u) < a.n;
    else
        return i < a.n;
}

// Species constants of row i: from the kernel arguments (scalar loads) for one species, not
// through a pointer that a per-spin species index would need.
__device__ __forceinline__ Species<float> species_of(const FusedArgs& a, uint32_t i) {
    Species<float> sp = a.sp0;
    if (a.multi) {
        const auto* q = gview(a.sp) + gview(a.spec)[i];
        sp.llgpf = q->llgpf;
        sp.lambda = q->lambda;
        sp.sigma = q->sigma;
        sp.k1u = q->k1u;
        sp.kc = q->kc;
        sp.k1dir[0] = q->k1dir[0];
        sp.k1dir[1] = q->k1dir[1];
        sp.k1dir[2] = q->k1dir[2];
    }
    return sp;
}

// Thermal field of row i at step t (a.noise != 0): counter noise drawn here (and stored for
// the corrector when a.th_out is set) or taken from the predictor (a.th_in), or the injected
// noise of the step.
__device__ __forceinline__ void thermal_field(const FusedArgs& a, uint32_t i, uint32_t iref, const Species<float>& sp,
                                              uint64_t t, float th[3]) {
    const uint32_t n = a.n;
    if (a.noise == 1) {
        if (a.th_in) {
            th[0] = gview(a.th_in)[i];
            th[1] = gview(a.th_in)[n + i];
            th[2] = gview(a.th_in)[2 * (size_t)n + i];
        } else {
            float z[3];
            normals3_f32(a.seed, a.stream, iref, t, z);
            const float TP = a.sqrtT * sp.sigma;
            th[0] = z[0] * TP;
            th[1] = z[1] * TP;
            th[2] = z[2] * TP;
            if (a.th_out) {
                gview(a.th_out)[i] = th[0];
                gview(a.th_out)[n + i] = th[1];
                gview(a.th_out)[2 * (size_t)n + i] = th[2];
            }
        }
    } else {
        const float TP = a.sqrtT * sp.sigma;
        th[0] = (float)gview(a.xi)[i] * TP;
        th[1] = (float)gview(a.xi)[n + i] * TP;
        th[2] = (float)gview(a.xi)[2 * (size_t)n + i] * TP;
    }
}

// Everything of f(S) at row i after the pair exchange h: four-spin term, dipolar field hd
// (already loaded when EDEM), staggered field, thermal field, applied field, anisotropy,
// and the LLG right-hand side.  The thermal field is drawn here, after the exchange: a
// global store (the predictor's hand-off of its draw) before the exchange would make the
// compiler read the read-only coupling streams with vector instead of scalar loads.
template <bool EDEM>
__device__ __forceinline__ void fused_tail(const FusedArgs& a, uint32_t i, uint32_t iref, const float* __restrict__ src,
                                           const float s[3], float h[3], float hd[3], float f[3], uint64_t t) {
    const uint32_t n = a.n;
    if (a.ex.fs_xadj) {
        float h4[3];
        fourspin_row<float, float, float>(a.ex, i, src, h4);
        h[0] += h4[0];
        h[1] += h4[1];
        h[2] += h4[2];
    }
    if (a.dem) {
        if constexpr (!EDEM) {
            hd[0] = gview(a.dem)[i];
            hd[1] = gview(a.dem)[n + i];
            hd[2] = gview(a.dem)[2 * (size_t)n + i];
        }
        h[0] += hd[0];
        h[1] += hd[1];
        h[2] += hd[2];
    }
    if (a.stg) {
        h[0] += gview(a.stg)[i];
        h[1] += gview(a.stg)[n + i];
        h[2] += gview(a.stg)[2 * (size_t)n + i];
    }
    const Species<float> sp = species_of(a, i);
    if (a.noise) {
        float th[3];
        thermal_field(a, i, iref, sp, t, th);
        h[0] += th[0];
        h[1] += th[1];
        h[2] += th[2];
    }
    h[0] += a.B[0];
    h[1] += a.B[1];
    h[2] += a.B[2];
    local_terms<float>(sp.k1u, sp.k1dir, sp.kc, a.an, s, h);
    llg_rhs<float>(sp.llgpf, sp.lambda, s, h, f);
}

// t: the step the thermal field is drawn for.  src: the spins (SoA) the field is evaluated at; gsrc: where neighbours are gathered from
// (src itself, or with G3 its AoS3 copy, which then also gives the row's own spin; src is
// then read only by the four-spin term).  EDEM: the dipolar field is loaded with the row's
// own spin, so its latency overlaps the exchange gathers instead of following them (the
// sums keep their order; the batch kernels, whose members are mostly dipolar films, use it).
template <int FMT, int NC, int UNR = 2, bool G3 = false, bool LFB = true, bool EDEM = false, bool CV = false>
__device__ __forceinline__ void fused_rhs(const FusedArgs& a, uint32_t i, uint32_t iref, const float* __restrict__ src,
                                          const float* __restrict__ gsrc, const float* sdict, float s[3], float f[3],
                                          uint64_t t) {
    const uint32_t n = a.n;
    if constexpr (G3) {
        const auto* v = gview(reinterpret_cast<const Spin3*>(gsrc)) + i;
        s[0] = v->x;
        s[1] = v->y;
        s[2] = v->z;
    } else {
        s[0] = gview(src)[i];
        s[1] = gview(src)[n + i];
        s[2] = gview(src)[2 * (size_t)n + i];
    }
    float hd[3] = {0.f, 0.f, 0.f};
    if (EDEM && a.dem) {
        hd[0] = gview(a.dem)[i];
        hd[1] = gview(a.dem)[n + i];
        hd[2] = gview(a.dem)[2 * (size_t)n + i];
    }
    float h[3];
    if constexpr (FMT == FMT_SELL_PACK)
        exch_row_pack_fast<NC, UNR, G3>(a.ex, i, gsrc, sdict, h);
    else if constexpr (is_line_fmt(FMT))
        exch_row_line<NC, G3, line_len(FMT), line_open(FMT), LFB, CV>(a.ex, i, gsrc, sdict, h);
    else
        exch_any<FMT, NC, float, float, float>(a.ex, i, src, h);
    fused_tail<EDEM>(a, i, iref, src, s, h, hd, f, t);
}

// Stage the coupling dictionary of the packed encoding in LDS (before any early exit).
// The packed reader's coupling dictionary in LDS.  Packed format: a static array.  Line
// formats: dynamic LDS, allocated by the launch (line_lds_bytes) only when some slices fall
// back to packed entries; the condition is the same for every thread of a block (one
// handle, or one batch member per block), so every thread reaches the barrier.
#define SLLG_STAGE_DICT(FMT_, NC_, a_)                                              \
    __shared__ float sdict_s_[FMT_ == FMT_SELL_PACK ? kDictLds : 1];              \
    extern __shared__ float sdict_d_[];                                           \
    float* const sdict = is_line_fmt(FMT_) ? sdict_d_ : sdict_s_;                 \
    if constexpr (FMT_ == FMT_SELL_PACK || is_line_fmt(FMT_)) {                   \
        if (FMT_ == FMT_SELL_PACK || (a_).ex.lfall) {                             \
            const uint32_t m_ = (a_).ex.ndict * NC_;                              \
            for (uint32_t k_ = threadIdx.x; k_ < m_; k_ += blockDim.x) sdict[k_] = gview((a_).ex.dict)[k_]; \
            __syncthreads();                                                      \
        }                                                                         \
    }

// Dynamic LDS bytes of a launch of format F (the dictionary, for line encodings with
// slices read from the packed entries).
template <int F>
inline size_t line_lds_bytes(const Exch<float>& e) {
    return (is_line_fmt(F) && e.lfall) ? kDictLds * sizeof(float) : 0;
}

// With G3 the kernels gather neighbours (and read the row's own spin) from the AoS3 copy
// (S3 / Y3, [3 i + c]), keep A = S + rdt/2 f(S) in the same layout, and keep S3 current: the predictor writes Y3 (and Y only when Y is
// non-null: the four-spin term or a dipolar refresh at the predicted spins reads it), the
// corrector S3 (and S only when S is non-null: the four-spin term reads it; every other
// reader of the spins first rebuilds S from S3, sync_spins).

template <int FMT, int NC, int TPB, int UNR, bool G3, bool LFB>
__global__ __launch_bounds__(TPB) void k_heun_pred(FusedArgs a, const float* __restrict__ S,
                                                    const float* __restrict__ S3, float* __restrict__ Y,
                                                    float* __restrict__ Y3, float* __restrict__ A) {
    SLLG_STAGE_DICT(FMT, NC, a)
    uint32_t i, iref;
    row_of_thread<TPB>(a, i, iref);
    if (!row_live<FMT>(a, i)) return;
    const uint32_t n = a.n;
    float s[3], f[3];
    fused_rhs<FMT, NC, UNR, G3, LFB>(a, i, iref, S, G3 ? S3 : S, sdict, s, f, a.t);
    float y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c] = s[c] + a.rdt * f[c];
    const float r = rsqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    if constexpr (G3) {  // A in the AoS3 layout too: one load in the corrector
        reinterpret_cast<Spin3*>(A)[i] =
            Spin3{s[0] + a.half_rdt * f[0], s[1] + a.half_rdt * f[1], s[2] + a.half_rdt * f[2]};
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) A[(size_t)c * n + i] = s[c] + a.half_rdt * f[c];
    }
    if (!G3 || Y) {
#pragma unroll
        for (int c = 0; c < 3; ++c) Y[(size_t)c * n + i] = y[c] * r;
    }
    if constexpr (G3) reinterpret_cast<Spin3*>(Y3)[i] = Spin3{y[0] * r, y[1] * r, y[2] * r};
}

template <int FMT, int NC, int TPB, int UNR, bool G3, bool LFB>
__global__ __launch_bounds__(TPB) void k_heun_corr(FusedArgs a, const float* __restrict__ Y,
                                                    const float* __restrict__ Y3, const float* __restrict__ A,
                                                    float* __restrict__ S, float* __restrict__ S3) {
    SLLG_STAGE_DICT(FMT, NC, a)
    uint32_t i, iref;
    row_of_thread<TPB>(a, i, iref);
    if (!row_live<FMT>(a, i)) return;
    const uint32_t n = a.n;
    float av[3];  // loaded before the field so its latency overlaps the neighbour gathers
    if constexpr (G3) {
        const Spin3 v = reinterpret_cast<const Spin3*>(A)[i];
        av[0] = v.x;
        av[1] = v.y;
        av[2] = v.z;
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) av[c] = A[(size_t)c * n + i];
    }
    float s[3], f[3];
    fused_rhs<FMT, NC, UNR, G3, LFB>(a, i, iref, Y, G3 ? Y3 : Y, sdict, s, f, a.t);
    float p[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = av[c] + a.half_rdt * f[c];
    const float r = rsqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (!G3 || S) {
#pragma unroll
        for (int c = 0; c < 3; ++c) S[(size_t)c * n + i] = p[c] * r;
    }
    if constexpr (G3) reinterpret_cast<Spin3*>(S3)[i] = Spin3{p[0] * r, p[1] * r, p[2] * r};
}

// ---------------------------------------------------------------- tile reader
// LDS-staged line reader of the Heun pair (64-cell line encodings in which every slice is a
// line, AoS3 spins).  A workgroup of NW waves takes the NW slices its stream describes
// (encode.hip, build_tiles), loads every line they read once from the AoS3 copy into LDS
// as [line][component][64] floats, and each wave then sums its slice's records with the
// windows read from LDS: a line shared by the workgroup's slices leaves the cache hierarchy
// once per workgroup instead of once per slice that reads it.  The records group the
// windows as the line records do and the sums run in the same order, so both readers give
// the same bits.

struct TileRow {
    const uint32_t* ts;  // the workgroup's stream
    uint32_t k, lane, i, iref;
    int32_t base;        // first row of the workgroup (sublattice 0, its first line)
};

template <int NW>
__device__ __forceinline__ TileRow tile_row(const FusedArgs& a) {
    TileRow r;
    const uint32_t w = tile_index(a.xcd);
    r.ts = a.ex.tent + __builtin_amdgcn_readfirstlane(a.ex.tgrp[w]);
    r.k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    r.lane = threadIdx.x & 63u;
    const uint32_t q = w * NW + r.k, na = a.perm.na;
    uint32_t sub, line, line0;
    if (a.na_sh >= 0) {
        sub = q & (na - 1u);
        line = q >> a.na_sh;
        line0 = (w * NW) >> a.na_sh;
    } else {
        sub = q % na;
        line = q / na;
        line0 = w * NW / na;
    }
    r.i = (sub * a.ex.tlps + line) * 64u + r.lane;
    r.iref = na > 1 ? (line * 64u + r.lane) * na + sub : r.i;
    r.base = (int32_t)(line0 * 64u);
    return r;
}

// Wave k stages lines [k c, (k + 1) c) of the workgroup's list, c = its length over NW
// rounded up to a multiple of 8: per 8 lines one scalar load of their first rows (the list
// starts 16-byte aligned; words past its end are the stream's own), then up to 8 line loads
// in flight, then their LDS stores.
template <int NW>
__device__ __forceinline__ void tile_stage(const TileRow& r, const float* __restrict__ G3src, uint32_t n, float* lds) {
    const SpinBuf sb = spin_buf(G3src, n);
    const uint32_t nU = r.ts[0];
    const uint32_t cpw = ((nU + NW - 1) / NW + 7) & ~7u;
    const uint32_t lo = r.k * cpw, hi = min(nU, lo + cpw);
    for (uint32_t u0 = lo; u0 < hi; u0 += 8) {
        const uint4* __restrict__ U4 = reinterpret_cast<const uint4*>(r.ts + tile_head(NW) + u0);
        const uint4 qa = U4[0], qb = U4[1];
        const uint32_t st[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
        u32x3 v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (u0 + j < hi) {
                const uint32_t row = __builtin_amdgcn_readfirstlane((uint32_t)(r.base + (int32_t)st[j]));
                v[j] = __builtin_amdgcn_raw_buffer_load_b96(sb.r, r.lane * 12u, row * 12u, 0);
            }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (u0 + j < hi) {
                float* d = lds + (u0 + j) * (kTileLineBytes / 4) + r.lane;
                d[0] = __uint_as_float(v[j].x);
                d[64] = __uint_as_float(v[j].y);
                d[128] = __uint_as_float(v[j].z);
            }
    }
}

// The windows of one record from LDS; lds4 is this lane's byte address in the first line.
template <int NC>
__device__ __forceinline__ void tile_group(const char* lds4, const TileRecW<NC>& c, float p[3][3]) {
    u32x3 x[TileRecW<NC>::G];
#pragma unroll
    for (int u = 0; u < TileRecW<NC>::G; ++u) {
        const float* q = reinterpret_cast<const float*>(lds4 + c.window(u));
        x[u].x = __float_as_uint(q[0]);
        x[u].y = __float_as_uint(q[64]);
        x[u].z = __float_as_uint(q[128]);
    }
    line_group_sum<NC>(c, x, p);
}

// Pair exchange of this wave's slice from LDS.  Two records per iteration in two register
// sets, each record's scalar load issued one record ahead (the stream has one spare record
// at its end, so the look-ahead past a slice's last record stays in bounds).  Per shift the
// contributions are summed at the unshifted lanes and moved once, as in the line reader.
template <int NC, int NW, bool OPEN>
__device__ __forceinline__ void exch_row_tile(const TileRow& r, const float* lds, float dot[3]) {
    using RW = TileRecW<NC>;
    constexpr int W = RW::W;
    const uint32_t* __restrict__ rp = r.ts + __builtin_amdgcn_readfirstlane(r.ts[2 + r.k]);
    const uint32_t ng = __builtin_amdgcn_readfirstlane(r.ts[2 + NW + r.k]);
    const char* lds4 = reinterpret_cast<const char*>(lds) + 4u * r.lane;
    float p[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    RW c0 = load_rec<RW>(rp);
    uint32_t g = 0;
    for (; g + 1 < ng; g += 2) {
        const RW c1 = load_rec<RW>(rp + (size_t)(g + 1) * W);
        tile_group<NC>(lds4, c0, p);
        const RW c2 = load_rec<RW>(rp + (size_t)(g + 2) * W);
        tile_group<NC>(lds4, c1, p);
        c0 = c2;
    }
    if (g < ng) tile_group<NC>(lds4, c0, p);
#pragma unroll
    for (int c = 0; c < 3; ++c) dot[c] = p[0][c] + line_next<64, OPEN>(p[1][c]) + line_prev<64, OPEN>(p[2][c]);
}

// Own spin (from LDS), pair exchange, then the rest of f(S).
template <int NC, int NW, bool OPEN>
__device__ __forceinline__ void tile_rhs(const FusedArgs& a, const TileRow& r, const float* lds,
                                         const float* __restrict__ src, float s[3], float f[3]) {
    const float* o = lds + (r.ts[2 + 2 * NW + r.k] >> 2) + r.lane;
    s[0] = o[0];
    s[1] = o[64];
    s[2] = o[128];
    float h[3], hd[3] = {0.f, 0.f, 0.f};
    exch_row_tile<NC, NW, OPEN>(r, lds, h);
    fused_tail<false>(a, r.i, r.iref, src, s, h, hd, f, a.t);
}

template <int NC, int NW, bool OPEN>
__global__ __launch_bounds__(NW * 64) void k_heun_pred_tile(FusedArgs a, const float* __restrict__ S,
                                                            const float* __restrict__ S3, float* __restrict__ Y,
                                                            float* __restrict__ Y3, float* __restrict__ A) {
    extern __shared__ float tl_lds[];
    const TileRow r = tile_row<NW>(a);
    tile_stage<NW>(r, S3, a.n, tl_lds);
    __syncthreads();
    const uint32_t n = a.n, i = r.i;
    float s[3], f[3];
    tile_rhs<NC, NW, OPEN>(a, r, tl_lds, S, s, f);
    float y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c] = s[c] + a.rdt * f[c];
    const float q = rsqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    reinterpret_cast<Spin3*>(A)[i] = Spin3{s[0] + a.half_rdt * f[0], s[1] + a.half_rdt * f[1], s[2] + a.half_rdt * f[2]};
    if (Y) {
#pragma unroll
        for (int c = 0; c < 3; ++c) Y[(size_t)c * n + i] = y[c] * q;
    }
    reinterpret_cast<Spin3*>(Y3)[i] = Spin3{y[0] * q, y[1] * q, y[2] * q};
}

template <int NC, int NW, bool OPEN>
__global__ __launch_bounds__(NW * 64) void k_heun_corr_tile(FusedArgs a, const float* __restrict__ Y,
                                                            const float* __restrict__ Y3, const float* __restrict__ A,
                                                            float* __restrict__ S, float* __restrict__ S3) {
    extern __shared__ float tl_lds[];
    const TileRow r = tile_row<NW>(a);
    const uint32_t n = a.n, i = r.i;
    const Spin3 av = reinterpret_cast<const Spin3*>(A)[i];  // overlaps the staging loads
    tile_stage<NW>(r, Y3, a.n, tl_lds);
    __syncthreads();
    float s[3], f[3];
    tile_rhs<NC, NW, OPEN>(a, r, tl_lds, Y, s, f);
    const float p[3] = {av.x + a.half_rdt * f[0], av.y + a.half_rdt * f[1], av.z + a.half_rdt * f[2]};
    const float q = rsqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (S) {
#pragma unroll
        for (int c = 0; c < 3; ++c) S[(size_t)c * n + i] = p[c] * q;
    }
    reinterpret_cast<Spin3*>(S3)[i] = Spin3{p[0] * q, p[1] * q, p[2] * q};
}

// SoA -> AoS3 gather copy (after the spins changed outside the Heun kernels).
__global__ __launch_bounds__(kTPB) void k_to_aos3(uint32_t n, const float* __restrict__ S, float* __restrict__ S3) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    reinterpret_cast<Spin3*>(S3)[i] = Spin3{S[i], S[n + i], S[2 * (size_t)n + i]};
}

// AoS3 -> SoA (the predictor output when only its AoS3 copy was written).
__global__ __launch_bounds__(kTPB) void k_from_aos3(uint32_t n, const float* __restrict__ S3, float* __restrict__ S) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= n) return;
    const Spin3 v = reinterpret_cast<const Spin3*>(S3)[i];
    S[i] = v.x;
    S[n + i] = v.y;
    S[2 * (size_t)n + i] = v.z;
}

// Batched Heun: one launch covers the tiles of every member of a batch (independent spin
// systems of any size, each with its own couplings, constants, temperature, field, seed
// and step counter).  Tile -> member is a table lookup made uniform across the block, so
// the member's arguments are read once per block with scalar loads.
struct BatchMember {
    FusedArgs a;
    float* S;
    float* Y;
    float* A;
    float* S3;      // AoS3 gather copies (launches with the gather copy)
    float* Y3;
    uint32_t ysoa;  // the predictor also writes the SoA Y (four-spin term, dipolar at Y)
    uint32_t pad_;
    uint32_t tile0, ntiles;
    uint64_t t0;  // step of the first step of the current launch chunk
};

// One half step of Heun for row i of member bm (a live row): the predictor (PHASE 0) or the
// corrector (PHASE 1) at step t.  CV: line records through the constant view.
template <int PHASE, int FMT, int NC, bool G3, bool CV = false>
__device__ __forceinline__ void batch_row(const BatchMember& bm, uint32_t i, uint64_t t, const float* sdict) {
    const FusedArgs& a = bm.a;  // read in place (uniform scalar loads); a local copy lands in scratch
    const uint32_t n = a.n;
    float s[3], f[3];
    if constexpr (PHASE == 0) {
        fused_rhs<FMT, NC, 2, G3, true, true, CV>(a, i, a.perm.to_ref(i), bm.S, G3 ? bm.S3 : bm.S, sdict, s, f, t);
        float y[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) y[c] = s[c] + a.rdt * f[c];
        const float r = rsqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) gview(bm.A)[(size_t)c * n + i] = s[c] + a.half_rdt * f[c];
        if (!G3 || bm.ysoa) {
#pragma unroll
            for (int c = 0; c < 3; ++c) gview(bm.Y)[(size_t)c * n + i] = y[c] * r;
        }
        if constexpr (G3) {
            auto* o = gview(reinterpret_cast<Spin3*>(bm.Y3)) + i;
            o->x = y[0] * r;
            o->y = y[1] * r;
            o->z = y[2] * r;
        }
    } else {
        float av[3];  // issued ahead of the field evaluation
#pragma unroll
        for (int c = 0; c < 3; ++c) av[c] = gview(bm.A)[(size_t)c * n + i];
        fused_rhs<FMT, NC, 2, G3, true, true, CV>(a, i, a.perm.to_ref(i), bm.Y, G3 ? bm.Y3 : bm.Y, sdict, s, f, t);
        float p[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = av[c] + a.half_rdt * f[c];
        const float r = rsqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) gview(bm.S)[(size_t)c * n + i] = p[c] * r;
        if constexpr (G3) {
            auto* o = gview(reinterpret_cast<Spin3*>(bm.S3)) + i;
            o->x = p[0] * r;
            o->y = p[1] * r;
            o->z = p[2] * r;
        }
    }
}

template <int PHASE, int FMT, int NC, int TPB, bool G3>
__global__ __launch_bounds__(TPB) void k_heun_batch(const BatchMember* __restrict__ mem,
                                                    const uint32_t* __restrict__ tile_member, uint32_t step_off,
                                                    int xcd) {
    const uint32_t tile = tile_index(xcd);
    const uint32_t m = __builtin_amdgcn_readfirstlane(tile_member[tile]);
    const BatchMember& bm = mem[m];
    SLLG_STAGE_DICT(FMT, NC, bm.a)
    const uint32_t i = (tile - bm.tile0) * TPB + threadIdx.x;
    if (!row_live<FMT>(bm.a, i)) return;
    batch_row<PHASE, FMT, NC, G3>(bm, i, bm.t0 + step_off, sdict);
}

// Wait at the e-th barrier of the launch until every block of the grid has arrived.  Every
// wave first waits until its own stores are acknowledged by the L2; then one wave releases the
// block's results at agent scope (each XCD has an L2 of its own: the release writes its dirty
// lines back) and arrives on its group's counter.  The group's last block acquires (so the
// group's releases happen before what follows), releases again, arrives on `count`, polls it
// until all groups are in, acquires and releases, and raises its group's flag; the group's
// other blocks poll that flag.  Every block then acquires at agent scope (invalidating the
// CU's L1 and the XCD's L2) before it reads what other blocks wrote.  A fence per wave would
// repeat that cache maintenance for every wave.  Every wait is bounded: after spin_limit polls
// the block raises `abort`, which every other wait also watches.  Returns false when the
// launch is being abandoned.
__device__ __forceinline__ bool grid_sync(GridBar* bar, uint32_t e, uint32_t spin_limit) {
    __shared__ uint32_t go;
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt, expcnt and lgkmcnt all zero: this wave's stores are in the L2
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t ng = gridDim.x < 8u ? gridDim.x : 8u, g = blockIdx.x & 7u;
        const uint32_t nb = (gridDim.x - g + 7u) >> 3;  // blocks of group g
        const uint32_t target = (e + 1u) * ng;
        const uint32_t old = __hip_atomic_fetch_add(&bar->group[g][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool last = old + 1u == (e + 1u) * nb;
        if (last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t* const word = last ? &bar->count : &bar->flag[g][0];
        const uint32_t want = last ? target : e + 1u;
        uint32_t it = 0, ok = 1;
        while (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
            if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || it >= spin_limit) {
                __hip_atomic_store(&bar->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            ++it;
            __builtin_amdgcn_s_sleep(2);
        }
        if (ok && __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
        if (last && ok) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(&bar->flag[g][0], e + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (last)  // the wait on `count` (one atomic per group; the other blocks' waits end with it)
            __hip_atomic_fetch_max(&bar->polls, it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        go = ok;
    }
    __syncthreads();
    return go != 0;
}

// Persistent batched Heun: one launch advances every member of the table `nsteps` steps from
// step offset step_off.  The grid holds only as many blocks as fit on the GPU at once (a
// cooperative launch); block b takes tiles in a stride over the tile range of its XCD (b % 8,
// with the XCD order) or over all tiles, the same tiles in both half steps, and a grid barrier
// separates the half steps.  Each row computes exactly what k_heun_batch computes for it.
template <int FMT, int NC, int TPB, bool G3>
__global__ __launch_bounds__(TPB) void k_heun_batch_persist(const BatchMember* __restrict__ mem,
                                                            const uint32_t* __restrict__ tile_member,
                                                            uint32_t ntiles, uint32_t step_off, uint32_t nsteps,
                                                            int xcd, GridBar* bar, uint32_t spin_limit) {
    __shared__ float sdict_s[FMT == FMT_SELL_PACK ? kDictLds : 1];
    extern __shared__ float sdict_d[];
    float* const sdict = is_line_fmt(FMT) ? sdict_d : sdict_s;
    uint32_t first, stride, end;
    if (xcd) {  // gridDim.x is a multiple of 8 here (checked on the host)
        const uint32_t x = blockIdx.x & 7u, per = ntiles >> 3, rem = ntiles & 7u;
        first = x * per + (x < rem ? x : rem);
        end = first + per + (x < rem ? 1u : 0u);
        first += blockIdx.x >> 3;
        stride = gridDim.x >> 3;
    } else {
        first = blockIdx.x;
        stride = gridDim.x;
        end = ntiles;
    }
    if (nsteps == 0) {  // residency probe: every block meets at one barrier, no data touched
        (void)grid_sync(bar, 0, spin_limit);
        return;
    }
    uint32_t staged = ~0u, e = 0;
    for (uint32_t k = 0; k < nsteps; ++k) {
        for (int phase = 0; phase < 2; ++phase) {
            for (uint32_t tile = first; tile < end; tile += stride) {
                const uint32_t m = cview(tile_member)[tile];
                const BatchMember& bm = *(const BatchMember*)cview(mem + m);
                if constexpr (FMT == FMT_SELL_PACK || is_line_fmt(FMT)) {
                    if (m != staged && (FMT == FMT_SELL_PACK || bm.a.ex.lfall)) {
                        const uint32_t nd = bm.a.ex.ndict * NC;
                        __syncthreads();  // every wave is done with the previous member's dictionary
                        for (uint32_t q = threadIdx.x; q < nd; q += TPB) sdict[q] = gview(bm.a.ex.dict)[q];
                        __syncthreads();
                        staged = m;
                    }
                }
                const uint32_t i = (tile - bm.tile0) * TPB + threadIdx.x;
                if (row_live<FMT>(bm.a, i)) {
                    if (phase == 0) batch_row<0, FMT, NC, G3, true>(bm, i, bm.t0 + step_off + k, sdict);
                    else batch_row<1, FMT, NC, G3, true>(bm, i, bm.t0 + step_off + k, sdict);
                }
            }
            if (k + 1 == nsteps && phase == 1) break;  // the launch's end orders the last half step
            if (!grid_sync(bar, e++, spin_limit)) return;
        }
    }
}

// RK4: stage m reads Yin (S for m = 0), k = rdt f(Yin);
//   m0: ACC = k/6, Yout = S + k/2;  m1: ACC += k/3, Yout = S + k/2;  m2: ACC += k/3, Yout = S + k;
//   m3: S = N(S + ACC + k/6).
template <int FMT, int NC, int M>
__global__ __launch_bounds__(kTPB) void k_rk4(FusedArgs a, const float* __restrict__ Yin, float* __restrict__ Yout,
                                              float* __restrict__ ACC, float* __restrict__ S) {
    SLLG_STAGE_DICT(FMT, NC, a)
    const uint32_t i = row_tile<kTPB>(a) * kTPB + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t n = a.n;
    float s[3], f[3];
    fused_rhs<FMT, NC>(a, i, a.perm.to_ref(i), M == 0 ? S : Yin, M == 0 ? S : Yin, sdict, s, f, a.t);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        const float k = a.rdt * f[c];
        const float s0 = (M == 0) ? s[c] : S[o];
        if constexpr (M == 0) {
            ACC[o] = k * (1.0f / 6.0f);
            Yout[o] = s0 + 0.5f * k;
        } else if constexpr (M == 1) {
            ACC[o] += k * (1.0f / 3.0f);
            Yout[o] = s0 + 0.5f * k;
        } else if constexpr (M == 2) {
            ACC[o] += k * (1.0f / 3.0f);
            Yout[o] = s0 + k;
        } else {
            s[c] = s0 + (ACC[o] + k * (1.0f / 6.0f));
        }
    }
    if constexpr (M == 3) {
        const float r = rsqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) S[(size_t)c * n + i] = s[c] * r;
    }
}

struct Rk5F {
    float b[6][5];
};
__constant__ Rk5F c_rk5f;

// Cash-Karp: stage m reads Yin (S for m = 0), k_m = rdt f(Yin), stores k_m (m <= 4),
// Yout = S + sum_{j<=m} b_{m+1,j} k_j (m <= 4); m = 5: S = N(S + sum_j c_j k_j).
template <int FMT, int NC, int M>
__global__ __launch_bounds__(kTPB) void k_rk5(FusedArgs a, const float* __restrict__ Yin, float* __restrict__ Yout,
                                              float* __restrict__ S, float* __restrict__ K0, float* __restrict__ K1,
                                              float* __restrict__ K2, float* __restrict__ K3, float* __restrict__ K4) {
    SLLG_STAGE_DICT(FMT, NC, a)
    const uint32_t i = row_tile<kTPB>(a) * kTPB + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t n = a.n;
    float s[3], f[3];
    fused_rhs<FMT, NC>(a, i, a.perm.to_ref(i), M == 0 ? S : Yin, M == 0 ? S : Yin, sdict, s, f, a.t);
    float* Ks[5] = {K0, K1, K2, K3, K4};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = (size_t)c * n + i;
        const float k = a.rdt * f[c];
        const float s0 = (M == 0) ? s[c] : S[o];
        if constexpr (M < 5) {
            Ks[M][o] = k;
            float v = s0;
#pragma unroll
            for (int j = 0; j < M; ++j) v += c_rk5f.b[M + 1][j] * Ks[j][o];
            v += c_rk5f.b[M + 1][M] * k;
            Yout[o] = v;
        } else {
            s[c] = s0 + ((37.0f / 378.0f) * K0[o] + (250.0f / 621.0f) * K2[o] + (125.0f / 594.0f) * K3[o] +
                         (512.0f / 1771.0f) * k);
        }
    }
    if constexpr (M == 5) {
        const float r = rsqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) S[(size_t)c * n + i] = s[c] * r;
    }
}

// Deterministic field in FP64 at FP32 spins Sp (thermal part excluded).
struct DetArgs {
    uint32_t n;
    int multi;
    double B[3];
    const float* dem;
    const float* stg;
    const uint8_t* spec;
    const Species<double>* sp;
    AnisG<double> an;
    Exch<double> ex;  // same entries, FP64 dictionary / values
};

template <int FMT, int NC>
__global__ __launch_bounds__(kTPB) void k_fused_detfield(DetArgs a, const float* __restrict__ Sp,
                                                         double* __restrict__ out) {
    const uint32_t i = blockIdx.x * kTPB + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t n = a.n;
    double h[3];
    exch_any<FMT, NC, double, double, float>(a.ex, i, Sp, h);
    if (a.ex.fs_xadj) {
        double h4[3];
        fourspin_row<double, double, float>(a.ex, i, Sp, h4);
        for (int c = 0; c < 3; ++c) h[c] += h4[c];
    }
    const double s[3] = {(double)Sp[i], (double)Sp[n + i], (double)Sp[2 * (size_t)n + i]};
    for (int c = 0; c < 3; ++c) {
        if (a.dem) h[c] += (double)a.dem[(size_t)c * n + i];
        if (a.stg) h[c] += (double)a.stg[(size_t)c * n + i];
        h[c] += a.B[c];
    }
    const Species<double> sp = a.sp[a.multi ? a.spec[i] : 0];
    local_terms<double>(sp.k1u, sp.k1dir, sp.kc, a.an, s, h);
    for (int c = 0; c < 3; ++c) out[(size_t)c * n + i] = h[c];
}

// ---------------------------------------------------------------- host launchers

static FusedArgs fused_args(const llg_ctx* c, const StepSched& s) {
    FusedArgs a{};
    a.n = c->n;
    a.perm = c->perm;
    if (c->T > 0.0) a.noise = (c->noise_mode == LLG_NOISE_INJECTED) ? 2 : 1;
    a.multi = c->multi_species ? 1 : 0;
    a.xcd = ((c->flags & LLG_FLAG_NO_XCD_REMAP) || !c->tune_xcd) ? 0 : 1;
    // every block size used (256..1024) divides ncell when 1024 does
    if (a.xcd && c->perm.na > 1 && c->perm.ncell % 1024u == 0) {
        a.il_na = c->perm.na;
        a.il_cells = c->perm.ncell;
    }
    a.tile_sub = (c->perm.na > 1 && c->perm.ncell % 1024u == 0) ? 1 : 0;
    auto log2_exact = [](uint32_t v) -> int { return (v && !(v & (v - 1))) ? __builtin_ctz(v) : -1; };
    a.na_sh = log2_exact(c->perm.na);
    a.ncell_sh = log2_exact(c->perm.ncell);
    a.rdt = (float)c->rdt;
    a.half_rdt = (float)(0.5 * c->rdt);
    a.sqrtT = (float)std::sqrt(c->T);
    for (int k = 0; k < 3; ++k) a.B[k] = (float)c->B[k];
    a.seed = c->seed;
    a.stream = c->stream_id;
    a.t = s.t;
    a.xi = s.xi;
    a.dem = c->dip ? c->DEM.as<float>() : nullptr;
    a.stg = c->stag_set ? c->STG.as<float>() : nullptr;
    a.spec = c->spec.as<uint8_t>();
    a.sp = c->sp32.as<Species<float>>();
    const Species<double>& s0 = c->sp64[0];
    a.sp0.llgpf = (float)s0.llgpf; a.sp0.lambda = (float)s0.lambda; a.sp0.sigma = (float)s0.sigma;
    a.sp0.k1u = (float)s0.k1u; a.sp0.kc = (float)s0.kc;
    for (int k = 0; k < 3; ++k) a.sp0.k1dir[k] = (float)s0.k1dir[k];
    a.an = c->an32;
    a.ex = c->ex32;
    return a;
}

#define SLLG_DISPATCH_FMT_NC(FMT_VAL, NC_VAL, MACRO)                                 \
    do {                                                                            \
        if ((FMT_VAL) == FMT_NONE) { MACRO(FMT_NONE, 1); }                          \
        else if ((FMT_VAL) == FMT_SELL_PACK) {                                      \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_PACK, 1); }                         \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_PACK, 3); }                    \
            else { MACRO(FMT_SELL_PACK, 9); }                                       \
        } else if ((FMT_VAL) == FMT_SELL_LINE) {                                    \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_LINE, 1); }                         \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_LINE, 3); }                    \
            else { MACRO(FMT_SELL_LINE, 9); }                                       \
        } else if ((FMT_VAL) == FMT_SELL_LINE16) {                                  \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_LINE16, 1); }                       \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_LINE16, 3); }                  \
            else { MACRO(FMT_SELL_LINE16, 9); }                                     \
        } else if ((FMT_VAL) == FMT_SELL_LINE_OPEN) {                               \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_LINE_OPEN, 1); }                    \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_LINE_OPEN, 3); }               \
            else { MACRO(FMT_SELL_LINE_OPEN, 9); }                                  \
        } else if ((FMT_VAL) == FMT_SELL_LINE16_OPEN) {                             \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_LINE16_OPEN, 1); }                  \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_LINE16_OPEN, 3); }             \
            else { MACRO(FMT_SELL_LINE16_OPEN, 9); }                                \
        } else if ((FMT_VAL) == FMT_SELL_RAW) {                                     \
            if ((NC_VAL) == 1) { MACRO(FMT_SELL_RAW, 1); }                          \
            else if ((NC_VAL) == 3) { MACRO(FMT_SELL_RAW, 3); }                     \
            else { MACRO(FMT_SELL_RAW, 9); }                                        \
        } else if ((FMT_VAL) == FMT_STENCIL) {                                      \
            if ((NC_VAL) == 1) { MACRO(FMT_STENCIL, 1); }                           \
            else if ((NC_VAL) == 3) { MACRO(FMT_STENCIL, 3); }                      \
            else { MACRO(FMT_STENCIL, 9); }                                         \
        } else {                                                                    \
            if ((NC_VAL) == 1) { MACRO(FMT_DIA, 1); }                               \
            else if ((NC_VAL) == 3) { MACRO(FMT_DIA, 3); }                          \
            else { MACRO(FMT_DIA, 9); }                                             \
        }                                                                           \
    } while (0)

// The Heun pair runs with the autotuned block size (256, 512 or 1024 threads), gather depth
// (packed encoding) and gather copy (packed and line encodings).
struct HeunBufs {
    float *S, *S3, *Y, *Y3, *A;
};

template <int F, int N, int TPB, int UNR, bool LFB>
static void launch_pred_gf(const FusedArgs& a, const HeunBufs& b, uint32_t n, bool g3, hipStream_t st) {
    const size_t lds = line_lds_bytes<F>(a.ex);
    if constexpr (F == FMT_SELL_PACK || is_line_fmt(F)) {
        if (g3) {
            k_heun_pred<F, N, TPB, UNR, true, LFB><<<blocks_for(n, TPB), TPB, lds, st>>>(a, b.S, b.S3, b.Y, b.Y3, b.A);
            return;
        }
    }
    k_heun_pred<F, N, TPB, UNR, false, LFB><<<blocks_for(n, TPB), TPB, lds, st>>>(a, b.S, b.S3, b.Y, b.Y3, b.A);
}
template <int F, int N, int TPB, int UNR, bool LFB>
static void launch_corr_gf(const FusedArgs& a, const HeunBufs& b, uint32_t n, bool g3, hipStream_t st) {
    const size_t lds = line_lds_bytes<F>(a.ex);
    if constexpr (F == FMT_SELL_PACK || is_line_fmt(F)) {
        if (g3) {
            k_heun_corr<F, N, TPB, UNR, true, LFB><<<blocks_for(n, TPB), TPB, lds, st>>>(a, b.Y, b.Y3, b.A, b.S, b.S3);
            return;
        }
    }
    k_heun_corr<F, N, TPB, UNR, false, LFB><<<blocks_for(n, TPB), TPB, lds, st>>>(a, b.Y, b.Y3, b.A, b.S, b.S3);
}
// Line encodings with slices read from the packed entries run the kernels that carry the
// packed reader; all others run without it.
template <int F, int N, int TPB, int UNR>
static void launch_pred_g(const FusedArgs& a, const HeunBufs& b, uint32_t n, bool g3, hipStream_t st) {
    if constexpr (is_line_fmt(F)) {
        if (a.ex.lfall) {
            launch_pred_gf<F, N, TPB, UNR, true>(a, b, n, g3, st);
            return;
        }
    }
    launch_pred_gf<F, N, TPB, UNR, false>(a, b, n, g3, st);
}
template <int F, int N, int TPB, int UNR>
static void launch_corr_g(const FusedArgs& a, const HeunBufs& b, uint32_t n, bool g3, hipStream_t st) {
    if constexpr (is_line_fmt(F)) {
        if (a.ex.lfall) {
            launch_corr_gf<F, N, TPB, UNR, true>(a, b, n, g3, st);
            return;
        }
    }
    launch_corr_gf<F, N, TPB, UNR, false>(a, b, n, g3, st);
}

template <int F, int N, int TPB>
static void launch_heun_pred(const FusedArgs& a, const HeunBufs& b, uint32_t n, int unr, bool g3, hipStream_t st) {
    if constexpr (F == FMT_SELL_PACK) {
        if (unr == 4) {
            launch_pred_g<F, N, TPB, 4>(a, b, n, g3, st);
            return;
        }
    }
    launch_pred_g<F, N, TPB, 2>(a, b, n, g3, st);
}

template <int F, int N, int TPB>
static void launch_heun_corr(const FusedArgs& a, const HeunBufs& b, uint32_t n, int unr, bool g3, hipStream_t st) {
    if constexpr (F == FMT_SELL_PACK) {
        if (unr == 4) {
            launch_corr_g<F, N, TPB, 4>(a, b, n, g3, st);
            return;
        }
    }
    launch_corr_g<F, N, TPB, 2>(a, b, n, g3, st);
}

// The predictor stores the thermal field it draws for the corrector of the same step
// (12 B written + 12 B read per spin instead of a second Philox + Box-Muller draw).
static bool cache_thermal(const llg_ctx* c) { return c->tune_thc && c->T > 0.0 && c->noise_mode == LLG_NOISE_COUNTER; }

// Readers of the SoA predicted spins: the four-spin term and a dipolar refresh at Y.
bool batch_needs_y_soa(const llg_ctx* c) { return c->ex32.fs_xadj || (c->dip && c->dip_every_stage); }

// The tile reader runs when chosen (line_mode 2) on an encoding that has tile streams; it
// stages from the AoS3 copy, so it implies the gather copy.
static bool use_tile(const llg_ctx* c) {
    const int f = c->ex32.fmt;
    return c->tune_line == 2 && c->ex32.tnw && (f == FMT_SELL_LINE || f == FMT_SELL_LINE_OPEN);
}

// The AoS3 gather copy is used for the packed and line encodings when the tuner chose it.
static bool use_g3(const llg_ctx* c) {
    if (use_tile(c)) return true;
    const int f = exec_fmt(c);
    return c->tune_g3 && (f == FMT_SELL_PACK || is_line_fmt(f));
}

template <int NC, int NW, bool OPEN>
static void launch_tile_nw(bool pred, const FusedArgs& a, const HeunBufs& b, hipStream_t st) {
    const uint32_t nwg = a.n / (64u * NW);
    const size_t lds = (size_t)a.ex.tumax * kTileLineBytes;
    if (pred)
        k_heun_pred_tile<NC, NW, OPEN><<<nwg, NW * 64, lds, st>>>(a, b.S, b.S3, b.Y, b.Y3, b.A);
    else
        k_heun_corr_tile<NC, NW, OPEN><<<nwg, NW * 64, lds, st>>>(a, b.Y, b.Y3, b.A, b.S, b.S3);
}
template <int NC, bool OPEN>
static void launch_tile_o(bool pred, const FusedArgs& a, const HeunBufs& b, hipStream_t st) {
    if (a.ex.tnw == 16)
        launch_tile_nw<NC, 16, OPEN>(pred, a, b, st);
    else
        launch_tile_nw<NC, 8, OPEN>(pred, a, b, st);
}
template <int NC>
static void launch_tile_c(bool pred, const FusedArgs& a, const HeunBufs& b, hipStream_t st) {
    if (a.ex.fmt == FMT_SELL_LINE_OPEN)
        launch_tile_o<NC, true>(pred, a, b, st);
    else
        launch_tile_o<NC, false>(pred, a, b, st);
}
static void launch_tile(bool pred, const FusedArgs& a, const HeunBufs& b, hipStream_t st) {
    if (a.ex.ncomp == 1)
        launch_tile_c<1>(pred, a, b, st);
    else if (a.ex.ncomp == 3)
        launch_tile_c<3>(pred, a, b, st);
    else
        launch_tile_c<9>(pred, a, b, st);
}

static int heun_bufs(llg_ctx* c, bool g3, HeunBufs& b) {
    b.S = c->S.as<float>();
    b.Y = c->Y.as<float>();
    b.A = c->A.as<float>();
    b.S3 = b.Y3 = nullptr;
    if (g3) {
        const size_t bytes = 3 * (size_t)c->n * sizeof(float);
        if (c->S3.bytes < bytes) {
            SLLG_TRY(c->S3.alloc(bytes));
            c->g3_valid = false;
        }
        if (c->Y3.bytes < bytes) SLLG_TRY(c->Y3.alloc(bytes));
        b.S3 = c->S3.as<float>();
        b.Y3 = c->Y3.as<float>();
    }
    return 0;
}

// S (SoA) from S3 when the corrector left only the AoS3 copy current.  Every reader of
// c->S outside the Heun kernels calls this first.
int sync_spins(llg_ctx* c) {
    if (c->s_soa_valid) return 0;
    k_from_aos3<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, c->S3.as<float>(), c->S.as<float>());
    SLLG_HIP(hipGetLastError());
    c->s_soa_valid = true;
    return 0;
}

int fused_y_soa(llg_ctx* c) {
    if (c->y_soa_valid) return 0;
    k_from_aos3<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, c->Y3.as<float>(), c->Y.as<float>());
    SLLG_HIP(hipGetLastError());
    c->y_soa_valid = true;
    return 0;
}

int fused_heun_pred(llg_ctx* c, const StepSched& s) {
    FusedArgs a = fused_args(c, s);
    const bool g3 = use_g3(c);
    HeunBufs b;
    SLLG_TRY(heun_bufs(c, g3, b));
    if (g3 && !c->g3_valid) {  // the spins changed outside the Heun kernels
        k_to_aos3<<<blocks_for(c->n, kTPB), kTPB, 0, c->stream>>>(c->n, b.S, b.S3);
        SLLG_HIP(hipGetLastError());
        c->g3_valid = true;
    }
    if (!g3 || c->ex32.fs_xadj) SLLG_TRY(sync_spins(c));  // this launch reads the SoA spins
    if (cache_thermal(c)) {
        if (c->HT.bytes < 3 * (size_t)c->n * sizeof(float)) SLLG_TRY(c->HT.alloc(3 * (size_t)c->n * sizeof(float)));
        a.th_out = c->HT.as<float>();
    }
    // With the AoS3 copy the corrector reads the predicted spins from Y3; the SoA Y is
    // written only for the readers that need it (four-spin term, dipolar refresh at Y).
    const bool ysoa = !g3 || c->ex32.fs_xadj || (c->dip && c->dip_every_stage);
    if (!ysoa) b.Y = nullptr;
    c->y_soa_valid = ysoa;
    if (use_tile(c)) {
        launch_tile(true, a, b, c->stream);
        SLLG_HIP(hipGetLastError());
        return 0;
    }
    const int tpb = c->tune_tpb, unr = c->tune_unr;
#define SLLG_L(F, N)                                                                          \
    do {                                                                                      \
        if (tpb == 1024) launch_heun_pred<F, N, 1024>(a, b, c->n, unr, g3, c->stream);        \
        else if (tpb == 512) launch_heun_pred<F, N, 512>(a, b, c->n, unr, g3, c->stream);     \
        else launch_heun_pred<F, N, 256>(a, b, c->n, unr, g3, c->stream);                     \
    } while (0)
    SLLG_DISPATCH_FMT_NC(exec_fmt(c), c->ex32.ncomp, SLLG_L);
#undef SLLG_L
    SLLG_HIP(hipGetLastError());
    return 0;
}

int fused_heun_corr(llg_ctx* c, const StepSched& s) {
    FusedArgs a = fused_args(c, s);
    const bool g3 = use_g3(c);
    HeunBufs b;
    SLLG_TRY(heun_bufs(c, g3, b));
    if (cache_thermal(c) && c->HT.p) a.th_in = c->HT.as<float>();
    // with the AoS3 copy the SoA spins are written only for the four-spin term, which reads
    // them; the other readers rebuild them on demand (sync_spins)
    const bool ssoa = !g3 || c->ex32.fs_xadj;
    if (!ssoa) b.S = nullptr;
    const int tpb = c->tune_tpb, unr = c->tune_unr;
#define SLLG_L(F, N)                                                                          \
    do {                                                                                      \
        if (tpb == 1024) launch_heun_corr<F, N, 1024>(a, b, c->n, unr, g3, c->stream);        \
        else if (tpb == 512) launch_heun_corr<F, N, 512>(a, b, c->n, unr, g3, c->stream);     \
        else launch_heun_corr<F, N, 256>(a, b, c->n, unr, g3, c->stream);                     \
    } while (0)
    if (use_tile(c))
        launch_tile(false, a, b, c->stream);
    else
        SLLG_DISPATCH_FMT_NC(exec_fmt(c), c->ex32.ncomp, SLLG_L);
#undef SLLG_L
    SLLG_HIP(hipGetLastError());
    c->g3_valid = g3;  // S3 is current only if the corrector wrote it
    c->s_soa_valid = ssoa;
    return 0;
}

int fused_rk_init_tables() {
    static bool done = false;
    if (done) return 0;
    Rk5F t{};
    const double b2[1] = {0.2};
    const double b3[2] = {3.0 / 40.0, 9.0 / 40.0};
    const double b4[3] = {0.3, -0.9, 1.2};
    const double b5[4] = {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0};
    const double b6[5] = {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0};
    const double* rows[6] = {nullptr, b2, b3, b4, b5, b6};
    for (int m = 1; m < 6; ++m)
        for (int j = 0; j < m; ++j) t.b[m][j] = (float)rows[m][j];
    SLLG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_rk5f), &t, sizeof(t)));
    done = true;
    return 0;
}

// Stage input ping-pong: stage m reads buffer m%2 (Y, A) and writes the other.
int fused_rk_stage(llg_ctx* c, const StepSched& s, int m) {
    FusedArgs a = fused_args(c, s);
    const dim3 g(blocks_for(c->n, kTPB)), b(kTPB);
    float* S = c->S.as<float>();
    float* P[2] = {c->Y.as<float>(), c->A.as<float>()};
    float* Yin = P[m & 1];
    float* Yout = P[(m + 1) & 1];
    float* K[6];
    for (int k = 0; k < 6; ++k) K[k] = c->K[k].as<float>();
    if (c->scheme == LLG_RK4) {
        float* ACC = K[0];
#define SLLG_L(F, N)                                                                              \
    switch (m) {                                                                                  \
        case 0: k_rk4<F, N, 0><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, ACC, S); break;              \
        case 1: k_rk4<F, N, 1><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, ACC, S); break;              \
        case 2: k_rk4<F, N, 2><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, ACC, S); break;              \
        default: k_rk4<F, N, 3><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, ACC, S); break;             \
    }
        SLLG_DISPATCH_FMT_NC(exec_fmt(c), c->ex32.ncomp, SLLG_L);
#undef SLLG_L
    } else {
#define SLLG_L(F, N)                                                                                    \
    switch (m) {                                                                                        \
        case 0: k_rk5<F, N, 0><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
        case 1: k_rk5<F, N, 1><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
        case 2: k_rk5<F, N, 2><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
        case 3: k_rk5<F, N, 3><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
        case 4: k_rk5<F, N, 4><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
        default: k_rk5<F, N, 5><<<g, b, line_lds_bytes<F>(a.ex), c->stream>>>(a, Yin, Yout, S, K[0], K[1], K[2], K[3], K[4]); break; \
    }
        SLLG_DISPATCH_FMT_NC(exec_fmt(c), c->ex32.ncomp, SLLG_L);
#undef SLLG_L
    }
    SLLG_HIP(hipGetLastError());
    return 0;
}

int fused_detfield(llg_ctx* c, const float* Sp, double* out) {
    DetArgs a{};
    a.n = c->n;
    a.multi = c->multi_species ? 1 : 0;
    for (int k = 0; k < 3; ++k) a.B[k] = c->B[k];
    a.dem = c->dip ? c->DEM.as<float>() : nullptr;
    a.stg = c->stag_set ? c->STG.as<float>() : nullptr;
    a.spec = c->spec.as<uint8_t>();
    a.sp = c->sp64d.as<Species<double>>();
    a.an = c->an64;
    a.ex = c->ex64;
    const dim3 g(blocks_for(c->n, kTPB)), b(kTPB);
    // the FP64 view of the couplings is the packed or raw encoding (never line mode)
#define SLLG_L(F, N)                                                                  \
    do {                                                                              \
        if constexpr (!is_line_fmt(F)) k_fused_detfield<F, N><<<g, b, 0, c->stream>>>(a, Sp, out); \
    } while (0)
    SLLG_DISPATCH_FMT_NC(c->ex64.fmt, c->ex64.ncomp, SLLG_L);
#undef SLLG_L
    SLLG_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- batches

size_t batch_member_size() { return sizeof(BatchMember); }

int batch_build_table(const std::vector<llg_ctx*>& members, const std::vector<uint64_t>& t0, uint32_t tpb,
                      bool g3, DevBuf& dmem, DevBuf& dmap, uint32_t& ntiles, hipStream_t s) {
    std::vector<BatchMember> hm(members.size());
    std::vector<uint32_t> map;
    uint32_t tile = 0;
    for (size_t k = 0; k < members.size(); ++k) {
        llg_ctx* c = members[k];
        SLLG_TRY(sync_spins(c));  // the batch kernels read and write the SoA spins
        const StepSched sch{t0[k], nullptr};
        hm[k].a = fused_args(c, sch);
        hm[k].S = c->S.as<float>();
        hm[k].Y = c->Y.as<float>();
        hm[k].A = c->A.as<float>();
        hm[k].S3 = hm[k].Y3 = nullptr;
        hm[k].ysoa = 1;
        hm[k].pad_ = 0;
        if (g3) {  // gather copies, brought up to date where the spins changed outside the kernels
            HeunBufs hb;
            SLLG_TRY(heun_bufs(c, true, hb));
            if (!c->g3_valid) {
                k_to_aos3<<<blocks_for(c->n, kTPB), kTPB, 0, s>>>(c->n, hb.S, hb.S3);
                SLLG_HIP(hipGetLastError());
                c->g3_valid = true;
            }
            hm[k].S3 = hb.S3;
            hm[k].Y3 = hb.Y3;
            hm[k].ysoa = batch_needs_y_soa(c) ? 1u : 0u;
        }
        hm[k].tile0 = tile;
        hm[k].ntiles = blocks_for(c->n, tpb);
        hm[k].t0 = t0[k];
        map.insert(map.end(), hm[k].ntiles, (uint32_t)k);
        tile += hm[k].ntiles;
    }
    ntiles = tile;
    const size_t mb = hm.size() * sizeof(BatchMember), tb = map.size() * sizeof(uint32_t);
    if (dmem.bytes < mb) SLLG_TRY(dmem.alloc(std::max<size_t>(mb, 64 * sizeof(BatchMember))));
    if (dmap.bytes < tb) SLLG_TRY(dmap.alloc(std::max<size_t>(tb * 2, 4096)));
    if (mb) SLLG_HIP(hipMemcpyAsync(dmem.p, hm.data(), mb, hipMemcpyHostToDevice, s));
    if (tb) SLLG_HIP(hipMemcpyAsync(dmap.p, map.data(), tb, hipMemcpyHostToDevice, s));
    SLLG_HIP(hipStreamSynchronize(s));
    return 0;
}

template <int PHASE, int F, int N, bool G3>
static void launch_batch_g3(hipStream_t s, const BatchMember* m, const uint32_t* map, uint32_t ntiles,
                            uint32_t step_off, uint32_t tpb, int xcd) {
    // members of a line format may have slices read from packed entries: reserve the dictionary
    const size_t lds = is_line_fmt(F) ? kDictLds * sizeof(float) : 0;
    if (tpb == 128) k_heun_batch<PHASE, F, N, 128, G3><<<ntiles, 128, lds, s>>>(m, map, step_off, xcd);
    else if (tpb == 512) k_heun_batch<PHASE, F, N, 512, G3><<<ntiles, 512, lds, s>>>(m, map, step_off, xcd);
    else k_heun_batch<PHASE, F, N, 256, G3><<<ntiles, 256, lds, s>>>(m, map, step_off, xcd);
}

// The gather copy applies to the packed and line readers (the others read SoA only).
template <int PHASE, int F, int N>
static void launch_batch_tpb(hipStream_t s, const BatchMember* m, const uint32_t* map, uint32_t ntiles,
                             uint32_t step_off, uint32_t tpb, int xcd, bool g3) {
    if constexpr (F == FMT_SELL_PACK || is_line_fmt(F)) {
        if (g3) {
            launch_batch_g3<PHASE, F, N, true>(s, m, map, ntiles, step_off, tpb, xcd);
            return;
        }
    }
    launch_batch_g3<PHASE, F, N, false>(s, m, map, ntiles, step_off, tpb, xcd);
}

int batch_heun_launch(hipStream_t s, int fmt, int nc, const void* dmem, const void* dmap, uint32_t ntiles,
                      uint32_t step_off, int phase, uint32_t tpb, int xcd, bool g3) {
    if (ntiles == 0) return 0;
    if (tpb != 128 && tpb != 256 && tpb != 512) SLLG_FAIL(LLG_E_INVALID, "batch tile of %u threads", tpb);
    const BatchMember* m = static_cast<const BatchMember*>(dmem);
    const uint32_t* map = static_cast<const uint32_t*>(dmap);
#define SLLG_L(F, N)                                                                   \
    do {                                                                               \
        if (phase == 0) launch_batch_tpb<0, F, N>(s, m, map, ntiles, step_off, tpb, xcd, g3); \
        else launch_batch_tpb<1, F, N>(s, m, map, ntiles, step_off, tpb, xcd, g3);            \
    } while (0)
    SLLG_DISPATCH_FMT_NC(fmt, nc, SLLG_L);
#undef SLLG_L
    SLLG_HIP(hipGetLastError());
    return 0;
}

// Clears the arrival counters and the poll maximum of a grid barrier (its abort flag stays).
__global__ void k_bar_reset(GridBar* bar) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        bar->count = 0;
        bar->polls = 0;
    }
    if (t < 8) {
        bar->group[t][0] = 0;
        bar->flag[t][0] = 0;
    }
}

// Polls of a residency probe's barrier: every block of a grid that fits is dispatched within
// microseconds, and a block that cannot become resident keeps the barrier open until then.
constexpr uint32_t kProbePolls = 1u << 16;
constexpr int kProbeDevices = 64;

// Blocks per CU of kernel f that are resident together, proven by probe launches: the
// occupancy figure first, then one fewer per failed probe.  A probe is a cooperative launch of
// blocks-per-CU x CUs blocks that only meets at one grid barrier with a bounded wait; when a
// block cannot become resident the resident ones give up after kProbePolls polls, and the rest
// leave at once when they start.  (The occupancy figure has been larger than what runs.)
static int probe_resident(hipStream_t s, const void* f, int tpb, size_t lds, int per, int ncu, GridBar* bar,
                          int* out) {
    for (int p = per; p >= 1; --p) {
        const uint32_t g = (uint32_t)p * (uint32_t)ncu;
        k_bar_reset<<<1, 64, 0, s>>>(bar);
        SLLG_HIP(hipGetLastError());
        const BatchMember* nm = nullptr;
        const uint32_t* nmap = nullptr;
        uint32_t zero = 0, lim = kProbePolls;
        int x = 0;
        void* args[] = {(void*)&nm, (void*)&nmap, (void*)&g, (void*)&zero, (void*)&zero,
                        (void*)&x, (void*)&bar, (void*)&lim};
        SLLG_HIP(hipLaunchCooperativeKernel(f, dim3(g), dim3(tpb), args, (unsigned)lds, s));
        GridBar hb{};
        SLLG_HIP(hipMemcpyAsync(&hb, bar, sizeof(hb), hipMemcpyDeviceToHost, s));
        SLLG_HIP(hipStreamSynchronize(s));
        if (!hb.abort) {
            *out = p;
            return 0;
        }
        SLLG_HIP(hipMemsetAsync(bar, 0, sizeof(GridBar), s));
        SLLG_HIP(hipStreamSynchronize(s));
    }
    SLLG_FAIL(LLG_E_DEVICE, "no grid of the persistent kernel was resident in a probe");
}

template <int F, int N, int TPB, bool G3>
static int launch_persist(hipStream_t s, const BatchMember* m, const uint32_t* map, uint32_t ntiles,
                          uint32_t step_off, uint32_t nsteps, int xcd, GridBar* bar, uint32_t spin_limit,
                          uint32_t* grid_out) {
    const void* f = reinterpret_cast<const void*>(&k_heun_batch_persist<F, N, TPB, G3>);
    const size_t lds = is_line_fmt(F) ? kDictLds * sizeof(float) : 0;
    int dev = 0, ncu = 0, coop = 0, per = 0;
    SLLG_HIP(hipGetDevice(&dev));
    SLLG_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) SLLG_FAIL(LLG_E_UNSUPPORTED, "device %d has no cooperative launches", dev);
    SLLG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // blocks per CU proven resident, per device (0: not probed yet)
    static std::mutex mu;
    static int resident[kProbeDevices] = {};
    if (dev < 0 || dev >= kProbeDevices) SLLG_FAIL(LLG_E_INVALID, "device %d beyond %d", dev, kProbeDevices);
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!resident[dev]) {
            SLLG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, TPB, lds));
            if (per < 1 || ncu < 1) SLLG_FAIL(LLG_E_DEVICE, "no resident block of the persistent kernel (%d per CU)", per);
            SLLG_TRY(probe_resident(s, f, TPB, lds, per, ncu, bar, &resident[dev]));
        }
        per = resident[dev];
    }
    if (nsteps == 0) return 0;  // residency known; nothing to launch
    uint32_t grid = std::min<uint64_t>(ntiles, (uint64_t)per * (uint64_t)ncu);
    if (xcd && grid >= 8) grid &= ~7u;  // the XCD tile ranges need every XCD to hold blocks
    else xcd = 0;
    k_bar_reset<<<1, 64, 0, s>>>(bar);
    SLLG_HIP(hipGetLastError());
    void* args[] = {(void*)&m, (void*)&map, (void*)&ntiles, (void*)&step_off, (void*)&nsteps,
                    (void*)&xcd, (void*)&bar, (void*)&spin_limit};
    SLLG_HIP(hipLaunchCooperativeKernel(f, dim3(grid), dim3(TPB), args, (unsigned)lds, s));
    if (grid_out) *grid_out = grid;
    return 0;
}

template <int F, int N>
static int persist_tpb(hipStream_t s, const BatchMember* m, const uint32_t* map, uint32_t ntiles, uint32_t step_off,
                       uint32_t nsteps, uint32_t tpb, int xcd, bool g3, GridBar* bar, uint32_t spin_limit,
                       uint32_t* grid) {
    if constexpr (F == FMT_SELL_PACK || is_line_fmt(F)) {
        if (g3) {
            if (tpb == 128) return launch_persist<F, N, 128, true>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
            if (tpb == 512) return launch_persist<F, N, 512, true>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
            return launch_persist<F, N, 256, true>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
        }
    }
    if (tpb == 128) return launch_persist<F, N, 128, false>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
    if (tpb == 512) return launch_persist<F, N, 512, false>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
    return launch_persist<F, N, 256, false>(s, m, map, ntiles, step_off, nsteps, xcd, bar, spin_limit, grid);
}

int batch_heun_persistent(hipStream_t s, int fmt, int nc, const void* dmem, const void* dmap, uint32_t ntiles,
                          uint32_t step_off, uint32_t nsteps, uint32_t tpb, int xcd, bool g3, GridBar* bar,
                          uint32_t spin_limit, uint32_t* grid) {
    if (ntiles == 0) return 0;
    if (tpb != 128 && tpb != 256 && tpb != 512) SLLG_FAIL(LLG_E_INVALID, "batch tile of %u threads", tpb);
    if (!bar) SLLG_FAIL(LLG_E_INVALID, "persistent launch without a grid barrier");
    if ((uint64_t)nsteps * 2u * ((uint64_t)ntiles + 8u) >= (1ull << 32))
        SLLG_FAIL(LLG_E_INVALID, "persistent launch of %u steps over %u tiles: barrier count overflows", nsteps, ntiles);
    const BatchMember* m = static_cast<const BatchMember*>(dmem);
    const uint32_t* map = static_cast<const uint32_t*>(dmap);
    int rc = 0;
#define SLLG_L(F, N) rc = persist_tpb<F, N>(s, m, map, ntiles, step_off, nsteps, tpb, xcd, g3, bar, spin_limit, grid)
    SLLG_DISPATCH_FMT_NC(fmt, nc, SLLG_L);
#undef SLLG_L
    return rc;
}

}  // namespace sllg
