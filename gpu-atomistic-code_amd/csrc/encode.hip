// encode.hip - host-side validation of the exchange given in the descriptor and its
// upload in the layout each path reads.
//
// FP64 plain path: the reference layouts as given (CSR u32 + FP64 [ncomp][nnz];
// DIA [ncomp][ndiag][n]; periodic stencil).
//
// FP32 path: CSR becomes sliced ELL with 64-row slices (one wavefront) so that every
// load of the coupling stream is a contiguous wave access (1 KB per uint4 load).  When the
// distinct coupling tuples fit the LDS dictionary (kDictLds floats) and every neighbour's
// row offset fits the bits the dictionary index leaves, an entry is one u32: row offset
// << dbits | dictionary index (4 B per nnz; the dictionary is staged in LDS by every
// block), and slices whose entries repeat an earlier slice's are stored once.  Otherwise
// entries are a u32 column plus ncomp FP32 values (M = 4 + 4 ncomp B per nnz).  Rows keep
// their stored order, so the per-row summation order of the reference CSR
// (src/cufields.cu:192-200) is kept.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"

namespace sllg {

namespace {

template <class T>
int upload(DevBuf& b, const T* h, size_t count, hipStream_t s) {
    SLLG_TRY(b.alloc(count * sizeof(T)));
    if (count) SLLG_HIP(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    SLLG_HIP(hipStreamSynchronize(s));
    return 0;
}

int validate_and_copy(llg_ctx* c, const llg_exchange& e) {
    ExchStore& X = *c->xs;
    HostExch& h = X.h;
    const uint32_t n = c->n;
    h.fmt = e.format;
    h.ncomp = e.ncomp ? e.ncomp : 1;
    if (h.fmt == LLG_EXCH_NONE) return 0;
    if (h.ncomp != 1 && h.ncomp != 3 && h.ncomp != 9) SLLG_FAIL(LLG_E_INVALID, "exchange ncomp must be 1, 3 or 9");
    const int nc = h.ncomp;
    if (h.fmt == LLG_EXCH_CSR) {
        if (!e.xadj) SLLG_FAIL(LLG_E_INVALID, "CSR exchange without xadj");
        if (e.nnz && (!e.adjncy || !e.csr_vals)) SLLG_FAIL(LLG_E_INVALID, "CSR exchange without adjncy/values");
        if (e.nnz > 0xFFFFFFFFull) SLLG_FAIL(LLG_E_UNSUPPORTED, "nnz above 2^32-1");
        if (e.xadj[0] != 0 || e.xadj[n] != e.nnz)
            SLLG_FAIL(LLG_E_INVALID, "CSR row pointers: xadj[0]=%u xadj[n]=%u nnz=%llu", e.xadj[0], e.xadj[n],
                      (unsigned long long)e.nnz);
        for (uint32_t i = 0; i < n; ++i)
            if (e.xadj[i + 1] < e.xadj[i]) SLLG_FAIL(LLG_E_INVALID, "CSR row pointers decrease at row %u", i);
        for (uint64_t q = 0; q < e.nnz; ++q)
            if (e.adjncy[q] >= n) SLLG_FAIL(LLG_E_INVALID, "CSR column %u out of range at nnz %llu", e.adjncy[q],
                                            (unsigned long long)q);
        h.nnz = e.nnz;
        h.xadj.assign(e.xadj, e.xadj + n + 1);
        h.adj.assign(e.adjncy, e.adjncy + e.nnz);
        h.vals.assign(e.csr_vals, e.csr_vals + (size_t)nc * e.nnz);
    } else if (h.fmt == LLG_EXCH_DIA) {
        if (e.ndiag && (!e.dia_offsets || !e.dia_vals)) SLLG_FAIL(LLG_E_INVALID, "DIA exchange without offsets/values");
        h.ndiag = e.ndiag;
        h.doff.assign(e.dia_offsets, e.dia_offsets + e.ndiag);
        h.vals.assign(e.dia_vals, e.dia_vals + (size_t)nc * e.ndiag * n);
    } else if (h.fmt == LLG_EXCH_STENCIL) {
        if ((uint64_t)e.dims[0] * e.dims[1] * e.dims[2] * e.n_atoms != n || e.n_atoms == 0)
            SLLG_FAIL(LLG_E_INVALID, "stencil dims x n_atoms != n");
        if (e.stencil_k && (!e.stencil_count || !e.stencil_entries || !e.stencil_vals))
            SLLG_FAIL(LLG_E_INVALID, "stencil without counts/entries/values");
        for (int a = 0; a < 3; ++a) { h.dims[a] = e.dims[a]; h.pbc[a] = e.pbc[a] ? 1 : 0; }
        h.na = e.n_atoms;
        h.sk = e.stencil_k;
        h.per_row = e.stencil_vals_per_row ? 1 : 0;
        h.scount.assign(e.stencil_count, e.stencil_count + h.na);
        h.sent.assign(e.stencil_entries, e.stencil_entries + (size_t)h.na * h.sk * 4);
        for (uint32_t t = 0; t < h.na; ++t) {
            if (h.scount[t] > h.sk) SLLG_FAIL(LLG_E_INVALID, "stencil count above stencil_k");
            for (uint32_t q = 0; q < h.scount[t]; ++q) {
                const int32_t* en = &h.sent[((size_t)t * h.sk + q) * 4];
                if (en[3] < 0 || (uint32_t)en[3] >= h.na) SLLG_FAIL(LLG_E_INVALID, "stencil target atom out of range");
                for (int a = 0; a < 3; ++a)
                    if (en[a] <= -(int32_t)h.dims[a] || en[a] >= (int32_t)h.dims[a])
                        SLLG_FAIL(LLG_E_INVALID, "stencil offset reaches beyond one period");
            }
        }
        const size_t nv = h.per_row ? (size_t)nc * h.sk * n : (size_t)nc * h.na * h.sk;
        h.vals.assign(e.stencil_vals, e.stencil_vals + nv);
    } else {
        SLLG_FAIL(LLG_E_INVALID, "unknown exchange format %d", h.fmt);
    }
    return 0;
}

template <class V>
void fill_common(Exch<V>& x, const HostExch& h, uint32_t n) {
    x.fmt = h.fmt;
    x.ncomp = h.ncomp;
    x.n = n;
    x.nnz = h.nnz;
    x.ndiag = h.ndiag;
    for (int a = 0; a < 3; ++a) { x.dims[a] = h.dims[a]; x.pbc[a] = h.pbc[a]; }
    x.na = h.na;
    x.sk = h.sk;
    x.per_row = h.per_row;
}

int upload_stencil_tables(llg_ctx* c) {
    ExchStore& X = *c->xs;
    const HostExch& h = X.h;
    std::vector<int4> ent((size_t)h.na * h.sk);
    for (size_t q = 0; q < ent.size(); ++q)
        ent[q] = make_int4(h.sent[q * 4], h.sent[q * 4 + 1], h.sent[q * 4 + 2], h.sent[q * 4 + 3]);
    SLLG_TRY(upload(X.a, ent.data(), ent.size(), c->stream));
    SLLG_TRY(upload(X.b, h.scount.data(), h.scount.size(), c->stream));
    return 0;
}

// Drop the host copies once the device holds the encoding (scalars stay for llg_get_info).
void release_host(HostExch& h) {
    std::vector<uint32_t>().swap(h.xadj);
    std::vector<uint32_t>().swap(h.adj);
    std::vector<double>().swap(h.vals);
    std::vector<int32_t>().swap(h.doff);
    std::vector<uint32_t>().swap(h.scount);
    std::vector<int32_t>().swap(h.sent);
}

}  // namespace

int build_exchange_fp64(llg_ctx* c, const llg_exchange& e) {
    ExchStore& X = *c->xs;
    SLLG_TRY(validate_and_copy(c, e));
    const HostExch& h = X.h;
    Exch<double>& x = X.ex64;
    fill_common(x, h, c->n);
    X.encoding = "none";
    if (h.fmt == LLG_EXCH_CSR) {
        // The CSR stored by slices of 64 rows: entry q of every row of a slice at [q][lane]
        // (column-major within the slice, the slice as long as its longest row), so one load
        // of entry q by a wave is 64 consecutive words.  Each row keeps its entries in CSR
        // order and the kernels read a row's own count from the row pointers, so the sums are
        // the CSR sums, operation for operation.
        const uint32_t n = c->n, nsl = (n + kSlice - 1) / kSlice, nc = (uint32_t)h.ncomp;
        std::vector<uint2> sl(nsl);
        uint64_t tot = 0;
        for (uint32_t t = 0; t < nsl; ++t) {
            uint32_t len = 0;
            for (uint32_t r = t * kSlice; r < std::min(n, (t + 1) * kSlice); ++r) len = std::max(len, h.xadj[r + 1] - h.xadj[r]);
            sl[t] = make_uint2((uint32_t)tot, len);
            tot += len;
        }
        if (tot > 0xFFFFFFFFull) SLLG_FAIL(LLG_E_INVALID, "exchange too large for the sliced CSR");
        const size_t nent = (size_t)tot * kSlice;
        std::vector<uint32_t> ent(nent);
        std::vector<double> val((size_t)nc * nent, 0.0);
        for (uint32_t r = 0; r < nsl * kSlice; ++r) {
            const uint2 t = sl[r / kSlice];
            const uint32_t rl = r < n ? h.xadj[r + 1] - h.xadj[r] : 0;
            for (uint32_t q = 0; q < t.y; ++q) {
                const size_t o = ((size_t)t.x + q) * kSlice + r % kSlice;
                ent[o] = q < rl ? h.adj[h.xadj[r] + q] : (r < n ? r : 0);  // padding is never read
                if (q < rl)
                    for (uint32_t k = 0; k < nc; ++k) val[(size_t)k * nent + o] = h.vals[(size_t)k * h.nnz + h.xadj[r] + q];
            }
        }
        SLLG_TRY(upload(X.a, h.xadj.data(), h.xadj.size(), c->stream));
        SLLG_TRY(upload(X.b, ent.data(), ent.size(), c->stream));
        SLLG_TRY(upload(X.c, val.data(), val.size(), c->stream));
        SLLG_TRY(upload(X.d, sl.data(), sl.size(), c->stream));
        x.xadj = X.a.as<uint32_t>();
        x.ent = X.b.as<uint32_t>();
        x.vals = X.c.as<double>();
        x.slice = X.d.as<uint2>();
        x.nent = nent;
        X.encoding = "csr-sliced64-u32-f64";
    } else if (h.fmt == LLG_EXCH_DIA) {
        SLLG_TRY(upload(X.a, h.doff.data(), h.doff.size(), c->stream));
        SLLG_TRY(upload(X.c, h.vals.data(), h.vals.size(), c->stream));
        x.doff = X.a.as<int32_t>();
        x.vals = X.c.as<double>();
        X.encoding = "dia-f64";
    } else if (h.fmt == LLG_EXCH_STENCIL) {
        SLLG_TRY(upload_stencil_tables(c));
        SLLG_TRY(upload(X.c, h.vals.data(), h.vals.size(), c->stream));
        x.sent = X.a.as<int4>();
        x.scount = X.b.as<uint32_t>();
        x.vals = X.c.as<double>();
        X.encoding = h.per_row ? "stencil-rowvals-f64" : "stencil-shared-f64";
    }
    release_host(X.h);
    return 0;
}

namespace {

struct TupleKey {
    double v[9];
    int nc;
    bool operator==(const TupleKey& o) const { return std::memcmp(v, o.v, sizeof(double) * nc) == 0; }
};
struct TupleHash {
    size_t operator()(const TupleKey& k) const {
        uint64_t h = 1469598103934665603ull;
        const unsigned char* p = reinterpret_cast<const unsigned char*>(k.v);
        for (size_t i = 0; i < sizeof(double) * (size_t)k.nc; ++i) h = (h ^ p[i]) * 1099511628211ull;
        return (size_t)h;
    }
};

// One neighbour window of a line-mode slice: its first row relative to the slice's, the
// shifts it has couplings for (bit 0/1/2: dz = 0/+1/-1) and the couplings [shift][component].
struct Line {
    int32_t lrel;
    uint32_t mask;
    double v[3][9];
};

// Windows in groups of G: full groups of one shift mask first (masks 7 down to 1), then the
// remainders together under the union of their masks; -1 marks an unused slot.
std::vector<std::vector<int>> group_windows(const std::vector<Line>& lines, int G) {
    std::vector<std::vector<int>> grp;
    std::vector<int> rest;
    for (uint32_t m = 7; m >= 1; --m) {
        std::vector<int> cls;
        for (size_t l = 0; l < lines.size(); ++l)
            if (lines[l].mask == m) cls.push_back((int)l);
        size_t k = 0;
        for (; k + (size_t)G <= cls.size(); k += (size_t)G)
            grp.emplace_back(cls.begin() + (ptrdiff_t)k, cls.begin() + (ptrdiff_t)k + G);
        rest.insert(rest.end(), cls.begin() + (ptrdiff_t)k, cls.end());
    }
    for (size_t k = 0; k < rest.size(); k += (size_t)G) {
        std::vector<int> g(rest.begin() + (ptrdiff_t)k, rest.begin() + (ptrdiff_t)std::min(k + G, rest.size()));
        g.resize((size_t)G, -1);
        grp.push_back(g);
    }
    return grp;
}

// Tile streams of the LDS-staged line reader, made from the line records of a 64-cell line
// encoding in which every slice is a line.  A workgroup takes nw slices: nw / na
// consecutive lines of each of the na sublattice blocks (slice q = w nw + k of workgroup w
// is line q / na of sublattice q % na), so the lines it reads overlap as much as the
// sublattice-major order allows.  Its stream:
//   [0]                 lines staged (nU), [1] nw
//   [2 + k]             slice k's records: first word, relative to the stream
//   [2 + nw + k]        slice k's group count
//   [2 + 2 nw + k]      LDS byte offset of slice k's own line
//   [H, H + nU)         first rows of the staged lines relative to the workgroup's first row
//                       (line q0 / na of sublattice 0, q0 = w nw), int32; H = tile_head(nw)
//   then slice k's records, tile_words(nc) words each: the LDS byte offsets of the record's
//   line_group(nc) windows (kTileLineBytes per line in staging order, so multiples of 8; the
//   first carries the record's shift mask in its low 3 bits), then the couplings
//   [window][dz 0, +1, -1][component] (FP32, zero where a window has none).  The windows are
//   grouped exactly as in the line records (group_windows, line_group(nc) to a group), so
//   both readers form the same sums.
// Streams of equal content are stored once (interior workgroups of a periodic lattice
// repeat); tile_words(nc) zero words follow the last stream, so a reader may fetch one
// record past a slice's last.  kept: the windows of each stored line-record list, by its
// first word.  Chooses nw = 8, else 16: the first that divides the slices evenly
// and stages at most kTileLdsMax bytes.  On failure the encoding keeps no tile streams.
void build_tiles(llg_ctx* c, const std::vector<uint2>& table,
                 const std::unordered_map<uint32_t, std::vector<Line>>& kept) {
    ExchStore& X = *c->xs;
    Exch<float>& x = X.ex32;
    const uint32_t n = c->n, nsl = n / kSlice, na = X.perm.na;
    const int nc = X.h.ncomp, G = line_group(nc), TW = tile_words(nc);
    if (na > 1 && X.perm.ncell % kSlice) return;
    const uint32_t lps = na > 1 ? X.perm.ncell / kSlice : nsl;
    // the groups of each stored window list, as the line records have them
    std::unordered_map<uint32_t, std::vector<std::vector<int>>> groups;
    for (const auto& kv : kept) groups.emplace(kv.first, group_windows(kv.second, G));
    const std::vector<std::vector<int>> nogroups;
    for (uint32_t nw : {8u, 16u}) {
        if (nw % na || nsl % nw || lps % (nw / na)) continue;
        const uint32_t kk = nw / na, nwg = nsl / nw, H = (uint32_t)tile_head((int)nw);
        std::vector<uint32_t> tw, st;
        std::unordered_map<uint64_t, std::vector<uint32_t>> seen;  // stream hash -> workgroups
        std::vector<uint32_t> offs(nwg);
        uint32_t umax = 0;
        bool ok = true;
        std::vector<int64_t> rows;  // staged lines' first rows, staging order
        std::unordered_map<int64_t, uint32_t> slot;
        for (uint32_t w = 0; w < nwg && ok; ++w) {
            const int64_t B = (int64_t)(w * kk) * kSlice;
            rows.clear();
            slot.clear();
            auto slot_of = [&](int64_t r) -> uint32_t {
                auto it = slot.find(r);
                if (it != slot.end()) return it->second;
                const uint32_t u = (uint32_t)rows.size();
                slot.emplace(r, u);
                rows.push_back(r);
                return u;
            };
            st.assign(H, 0u);
            std::vector<uint32_t> recs;
            for (uint32_t k = 0; k < nw; ++k) {
                const uint32_t q = w * nw + k, s = (q % na) * lps + q / na;
                const int64_t base = (int64_t)s * kSlice;
                const uint2 t = table[s];
                const auto gi = t.y ? groups.find(t.x) : groups.end();
                const std::vector<std::vector<int>>& grp = gi != groups.end() ? gi->second : nogroups;
                st[2 + nw + k] = (uint32_t)grp.size();
                st[2 + 2 * nw + k] = slot_of(base) * kTileLineBytes;
                st[2 + k] = (uint32_t)recs.size();  // made relative to the stream below
                if (!grp.empty()) {
                    const std::vector<Line>& lines = kept.at(t.x);
                    for (const std::vector<int>& g : grp) {
                        const size_t o = recs.size();
                        recs.resize(o + (size_t)TW, 0u);
                        uint32_t mask = 0;
                        for (int u = 0; u < G; ++u) {
                            const Line& ln = lines[(size_t)(g[(size_t)u] >= 0 ? g[(size_t)u] : g[0])];
                            recs[o + (size_t)u] = slot_of(base + ln.lrel) * kTileLineBytes;
                            if (g[(size_t)u] < 0) continue;
                            mask |= ln.mask;
                            for (int kz = 0; kz < 3; ++kz)
                                for (int a = 0; a < nc; ++a) {
                                    const float v = (float)ln.v[kz][a];
                                    std::memcpy(&recs[o + (size_t)(G + (u * 3 + kz) * nc + a)], &v, 4);
                                }
                        }
                        recs[o] |= mask;
                    }
                }
            }
            const uint32_t nU = (uint32_t)rows.size();
            if ((uint64_t)nU * kTileLineBytes > kTileLdsMax) {
                ok = false;
                break;
            }
            umax = std::max(umax, nU);
            st[0] = nU;
            st[1] = nw;
            for (int64_t r : rows) st.push_back((uint32_t)(int32_t)(r - B));
            while (st.size() % 4) st.push_back(0u);
            const uint32_t r0 = (uint32_t)st.size();
            for (uint32_t k = 0; k < nw; ++k) st[2 + k] += r0;
            st.insert(st.end(), recs.begin(), recs.end());
            uint64_t hsh = 1469598103934665603ull;
            for (uint32_t v : st) hsh = (hsh ^ v) * 1099511628211ull;
            bool found = false;
            auto& cand = seen[hsh];
            for (uint32_t o : cand)
                if (offs[o] + st.size() <= tw.size() && tw[offs[o]] == st[0] &&
                    std::memcmp(tw.data() + offs[o], st.data(), st.size() * 4) == 0) {
                    // equal prefix of equal length: the stored stream has the same header,
                    // so its length is the same
                    offs[w] = offs[o];
                    found = true;
                    break;
                }
            if (!found) {
                if (tw.size() + st.size() + (size_t)TW >= (1ull << 32)) {  // offsets are 32-bit
                    ok = false;
                    break;
                }
                cand.push_back(w);
                offs[w] = (uint32_t)tw.size();
                tw.insert(tw.end(), st.begin(), st.end());
            }
        }
        if (!ok) continue;
        tw.insert(tw.end(), (size_t)TW, 0u);
        if (upload(X.f, tw.data(), tw.size(), c->stream)) return;
        if (upload(X.g, offs.data(), offs.size(), c->stream)) return;
        x.tent = X.f.as<uint32_t>();
        x.tgrp = X.g.as<uint32_t>();
        x.tnw = nw;
        x.tumax = umax;
        x.tlps = lps;
        return;
    }
}

// Line mode of the FP32 path.  When the lattice's fastest cell index runs along lines of
// len = 64 or 16 cells, a 64-row slice is one line (len 64) or four consecutive lines
// (len 16), and neighbour j of lane l is at row
//   L + (l & ~(len - 1)) + ((l + dz) mod len),   dz in {0, +1, -1},
// for periodic lines, or at L + l + dz with no neighbour past a line's ends for open lines,
// of a 64-row window starting at row L (a multiple of len).  Where that holds for every
// lane with one coupling tuple, the couplings of a slice are lane-uniform: per window, one
// load of the spins, and the shifted neighbours come from a rotation (periodic) or a
// zero-filling shift (open) across the wave (len 64) or within each 16-lane row (len 16).  Slices where it does not hold (a window
// that would wrap a periodic boundary inside the slice) keep the packed entries
// (kLineFallback).  A slice becomes a list of group records read with scalar loads,
// G = line_group(ncomp) windows each, line_words(ncomp) words:
//   [0]                  shift mask of the group: bit 0/1/2 = some window of the group has
//                        couplings for dz = 0 / +1 / -1
//   [1, 1 + G)           window starts L - slice base (int32, multiples of len)
//   [1 + G, 1 + G + 3 G nc)  couplings [window][dz 0, +1, -1][component], FP32, zero where
//                        the window has none
//   [W - G, W)           the next group's window starts (zero after the last group); W is
//                        rounded up to whole 16-byte units, the words between are zero
// so the reader issues the next group's window loads before it uses the current group.
// Windows are grouped by their own shift masks (full groups of one mask first, the rest
// together under the union of their masks), so a group tests each shift once for all its
// windows; an unused slot repeats the group's first window with zero couplings.  The reader
// sums each shift's contributions unshifted and rotates the three partial sums once.
// Entries of one row that name the same (window, dz) are summed (in FP64, rounded once).
// Identical record lists are stored once; the first record is a zero record for empty
// slices.  Entries are summed window by window, not in stored CSR order.
bool build_lines(llg_ctx* c, const std::vector<uint32_t>& ref_of, const std::vector<uint32_t>& idx,
                 const std::vector<TupleKey>& tuples, uint32_t len, bool open, uint32_t& stored,
                 uint32_t& lines_row, uint32_t& nfall) {
    ExchStore& X = *c->xs;
    const HostExch& h = X.h;
    const uint32_t n = c->n;
    const int nc = h.ncomp;
    const Perm P = X.perm;
    if (n % kSlice != 0) return false;
    const uint32_t nsl = n / kSlice;
    const int G = line_group(nc), W = line_words(nc);
    std::vector<uint32_t> words((size_t)W, 0u);  // the zero record
    std::vector<uint2> table(nsl);
    std::unordered_map<uint64_t, std::vector<uint32_t>> seen;  // record-list hash -> stored slices
    std::unordered_map<uint32_t, std::vector<Line>> kept;     // windows of each stored record list
    std::vector<uint32_t> st;
    stored = 0;
    nfall = 0;
    uint64_t lines_total = 0, lined = 0;
    const int32_t lm = (int32_t)len - 1;
    // key of one coupling of row d0 + l: window start relative to d0, shift dz, tuple
    using LKey = std::tuple<int64_t, int, uint32_t>;
    // keys of row d0 + l in CSR order (csr, optional) and sorted (out, for comparing lanes)
    auto lane_keys = [&](uint32_t d0, int32_t l, std::vector<LKey>& out, std::vector<LKey>* csr) -> bool {
        out.clear();
        const uint32_t r = ref_of[d0 + (uint32_t)l];
        for (uint32_t qq = h.xadj[r]; qq < h.xadj[r + 1]; ++qq) {
            const int64_t j = P.to_dev(h.adj[qq]);
            bool found = false;
            for (int dz : {0, 1, -1}) {
                int32_t pos = (l & lm) + dz;
                if (open) {
                    if (pos < 0 || pos > lm) continue;
                } else {
                    pos &= lm;
                }
                const int64_t L = j - (int64_t)((l & ~lm) + pos);
                if (L % (int64_t)len == 0 && L >= 0 && L + kSlice <= (int64_t)n) {
                    out.emplace_back(L - (int64_t)d0, dz, idx[qq]);
                    found = true;
                    break;
                }
            }
            if (!found) return false;
        }
        if (csr) *csr = out;
        std::sort(out.begin(), out.end());
        return true;
    };
    std::vector<LKey> ref, ref_csr, lk, first, last;
    for (uint32_t s = 0; s < nsl; ++s) {
        const uint32_t d0 = s * kSlice;
        std::vector<Line> lines;
        // periodic lines: every lane has the same keys; open lines: every interior lane
        // does, and a line's first (last) lane lacks the keys that shift past its start (end)
        bool fits = lane_keys(d0, open ? 1 : 0, ref, &ref_csr);
        if (fits && open) {
            first.clear();
            last.clear();
            for (const LKey& k : ref) {
                if (std::get<1>(k) != -1) first.push_back(k);
                if (std::get<1>(k) != 1) last.push_back(k);
            }
        }
        for (int32_t l = 0; l < kSlice && fits; ++l) {
            fits = lane_keys(d0, l, lk, nullptr);
            if (!fits) break;
            const int32_t p = l & lm;
            const std::vector<LKey>& want = !open ? ref : (p == 0 ? first : (p == lm ? last : ref));
            fits = lk == want;
        }
        if (fits) {
            // windows in the order the CSR row names them; repeated neighbours add up
            for (const LKey& key : ref_csr) {
                const int32_t lrel = (int32_t)std::get<0>(key);
                const int dz = std::get<1>(key);
                Line* g = nullptr;
                for (Line& x : lines)
                    if (x.lrel == lrel) g = &x;
                if (!g) {
                    lines.push_back(Line{lrel, 0u, {}});
                    g = &lines.back();
                }
                const int k = dz == 0 ? 0 : (dz == 1 ? 1 : 2);
                g->mask |= 1u << k;
                for (int a = 0; a < nc; ++a) g->v[k][a] += tuples[std::get<2>(key)].v[a];
            }
        }
        if (!fits) {  // this slice keeps its packed entries
            table[s] = make_uint2(0u, kLineFallback);
            ++nfall;
            continue;
        }
        ++lined;
        st.clear();
        const std::vector<std::vector<int>> grp = group_windows(lines, G);
        const size_t ng = grp.size();
        auto start_of = [&](const std::vector<int>& g, int u) -> uint32_t {
            return (uint32_t)lines[(size_t)(g[(size_t)u] >= 0 ? g[(size_t)u] : g[0])].lrel;
        };
        for (size_t gi = 0; gi < ng; ++gi) {
            const std::vector<int>& g = grp[gi];
            std::vector<uint32_t> rec((size_t)W, 0u);
            for (int u = 0; u < G; ++u) {
                rec[(size_t)(1 + u)] = start_of(g, u);
                if (gi + 1 < ng) rec[(size_t)(W - G + u)] = start_of(grp[gi + 1], u);
                if (g[(size_t)u] < 0) continue;
                const Line& ln = lines[(size_t)g[(size_t)u]];
                rec[0] |= ln.mask;
                for (int k = 0; k < 3; ++k)
                    for (int a = 0; a < nc; ++a) {
                        const float v = (float)ln.v[k][a];
                        std::memcpy(&rec[(size_t)(1 + G + (u * 3 + k) * nc + a)], &v, 4);
                    }
            }
            st.insert(st.end(), rec.begin(), rec.end());
        }
        lines_total += lines.size();
        const uint32_t nrec = (uint32_t)ng;
        if (nrec == 0) {
            table[s] = make_uint2(0u, 0u);
            continue;
        }
        uint64_t hsh = 1469598103934665603ull;
        for (uint32_t w : st) hsh = (hsh ^ w) * 1099511628211ull;
        bool found = false;
        auto& cand = seen[hsh];
        for (uint32_t o : cand)
            if (table[o].y == nrec && std::memcmp(words.data() + table[o].x, st.data(), st.size() * 4) == 0) {
                table[s] = table[o];
                found = true;
                break;
            }
        if (!found) {
            cand.push_back(s);
            table[s] = make_uint2((uint32_t)words.size(), nrec);
            words.insert(words.end(), st.begin(), st.end());
            kept.emplace(table[s].x, lines);
            ++stored;
        }
    }
    // worth it only when most slices are lines
    if (lined == 0 || 4u * nfall > nsl) return false;
    if (upload(X.d, words.data(), words.size(), c->stream)) return false;
    if (upload(X.e, table.data(), table.size(), c->stream)) return false;
    Exch<float>& x = X.ex32;
    x.tnw = x.tumax = x.tlps = 0;
    if (len == 64 && nfall == 0) build_tiles(c, table, kept);
    x.fmt = len == 64 ? (open ? FMT_SELL_LINE_OPEN : FMT_SELL_LINE) : (open ? FMT_SELL_LINE16_OPEN : FMT_SELL_LINE16);
    x.lslice = X.e.as<uint2>();
    x.lent = X.d.as<uint32_t>();
    x.lfall = nfall;
    lines_row = (uint32_t)(lines_total / lined);
    return true;
}

// CSR -> sliced ELL.  Packed layout: per slice, entries in chunks of 4 per lane,
// [chunk][lane][4], so one uint4 load per lane reads 4 entries and a wave reads 1 KB
// contiguous, then a tail of 2 per lane [lane][2] (one uint2 load); slice lengths are
// rounded up to a multiple of 2 with zero-coupling entries
// (offset 0, dictionary tuple 0.0), which add exactly 0 to the field.  Raw layout:
// [entry][lane] u32 columns and [ncomp][entry][lane] FP32 values.
int build_sell(llg_ctx* c) {
    ExchStore& X = *c->xs;
    const HostExch& h = X.h;
    const uint32_t n = c->n;
    const int nc = h.ncomp;
    const Perm P = X.perm;
    const uint32_t ncell = P.ncell;
    const uint32_t nsl = (n + kSlice - 1) / kSlice;
    // device row d holds reference row P.to_ref(d)
    std::vector<uint32_t> ref_of(n);
    for (uint32_t d = 0; d < n; ++d) ref_of[d] = P.to_ref(d);
    std::vector<uint32_t> mx(nsl, 0);
    for (uint32_t s = 0; s < nsl; ++s)
        for (uint32_t d = s * kSlice; d < std::min(n, (s + 1) * kSlice); ++d) {
            const uint32_t r = ref_of[d];
            mx[s] = std::max(mx[s], h.xadj[r + 1] - h.xadj[r]);
        }
    // Packed form: few distinct coupling tuples (the LDS dictionary) and row offsets that
    // fit the bits the dictionary index leaves; 12 n bytes must fit one buffer resource.
    bool pack = (uint64_t)n * 12u < (1ull << 31);
    std::unordered_map<TupleKey, uint32_t, TupleHash> dict;
    std::vector<TupleKey> tuples;
    std::vector<uint32_t> idx;
    if (pack) {
        TupleKey z{};
        z.nc = nc;
        dict.emplace(z, 0u);  // tuple 0 is the zero coupling used for padding
        tuples.push_back(z);
        idx.resize(h.nnz);
        TupleKey last{};
        last.nc = -1;
        uint32_t last_id = 0;
        for (uint64_t q = 0; q < h.nnz && pack; ++q) {
            TupleKey k{};
            k.nc = nc;
            for (int a = 0; a < nc; ++a) k.v[a] = h.vals[(size_t)a * h.nnz + q];
            if (last.nc == nc && k == last) { idx[q] = last_id; continue; }
            auto it = dict.find(k);
            uint32_t id;
            if (it == dict.end()) {
                id = (uint32_t)tuples.size();
                if ((uint64_t)(id + 1) * nc > (uint64_t)kDictLds || id > 0xFFFu) { pack = false; break; }
                dict.emplace(k, id);
                tuples.push_back(k);
            } else {
                id = it->second;
            }
            idx[q] = id;
            last = k;
            last_id = id;
        }
    }
    uint32_t dbits = 1;
    if (pack) {
        while ((1u << dbits) < tuples.size()) ++dbits;
        const int64_t lim = (int64_t)1 << (31 - dbits);
        for (uint32_t r = 0; r < n && pack; ++r) {
            const int64_t dr = P.to_dev(r);
            for (uint32_t q = h.xadj[r]; q < h.xadj[r + 1]; ++q) {
                const int64_t off = (int64_t)P.to_dev(h.adj[q]) - dr;
                if (off < -lim || off >= lim) { pack = false; break; }
            }
        }
    }
    // slice table {first entry / 64, entries per row}
    std::vector<uint2> slice(nsl);
    uint64_t tot = 0;
    for (uint32_t s = 0; s < nsl; ++s) {
        slice[s].y = pack ? (mx[s] + 1u) / 2u * 2u : mx[s];
        slice[s].x = (uint32_t)tot;
        tot += slice[s].y;
    }
    if (tot > 0xFFFFFFFFull) SLLG_FAIL(LLG_E_INVALID, "exchange too large for the sliced-ELL encoding");
    Exch<float>& x = X.ex32;
    Exch<double>& x64 = X.ex64;
    std::vector<uint32_t> ent;
    uint64_t nent = tot * kSlice;
    if (pack) {
        // Packed entries are relative (sublattice, cell delta, coupling index), so slices of a
        // periodic lattice away from the boundary planes come out identical: each distinct
        // slice block is stored once and the slice table points repeats at it.
        const bool share = !(c->flags & LLG_FLAG_NO_SLICE_SHARING);
        std::unordered_map<uint64_t, std::vector<uint32_t>> seen;  // block hash -> stored slices
        std::vector<uint32_t> blk;
        ent.reserve(share ? 1u << 16 : nent);
        uint64_t used = 0;
        uint32_t stored = 0;
        for (uint32_t s = 0; s < nsl; ++s) {
            const uint32_t len = slice[s].y;
            blk.assign((size_t)len * kSlice, 0u);
            for (uint32_t l = 0; l < kSlice; ++l) {
                const uint32_t d = s * kSlice + l;
                const uint32_t r = d < n ? ref_of[d] : 0;
                for (uint32_t q = 0; q < len; ++q) {
                    uint32_t w = pack_word(0, 0, dbits);  // zero coupling on itself
                    if (d < n && q < h.xadj[r + 1] - h.xadj[r]) {
                        const uint32_t qq = h.xadj[r] + q;
                        w = pack_word((int32_t)((int64_t)P.to_dev(h.adj[qq]) - (int64_t)d), idx[qq], dbits);
                    }
                    blk[pack_pos(q, l, len)] = w;
                }
            }
            uint64_t hsh = 1469598103934665603ull ^ len;  // FNV-1a over the block's words
            for (uint32_t w : blk) hsh = (hsh ^ w) * 1099511628211ull;
            bool found = false;
            if (share) {
                auto& cand = seen[hsh];
                for (uint32_t o : cand)
                    if (slice[o].y == len &&
                        std::memcmp(ent.data() + (size_t)slice[o].x * kSlice, blk.data(), blk.size() * 4) == 0) {
                        slice[s].x = slice[o].x;
                        found = true;
                        break;
                    }
                if (!found) cand.push_back(s);
            }
            if (!found) {
                slice[s].x = (uint32_t)used;
                ent.insert(ent.end(), blk.begin(), blk.end());
                used += len;
                ++stored;
            }
        }
        nent = used * kSlice;
        std::vector<float> d32(tuples.size() * nc);
        std::vector<double> d64(tuples.size() * nc);
        for (size_t t = 0; t < tuples.size(); ++t)
            for (int a = 0; a < nc; ++a) {
                d64[t * nc + a] = tuples[t].v[a];
                d32[t * nc + a] = (float)tuples[t].v[a];
            }
        SLLG_TRY(upload(X.b, ent.data(), ent.size(), c->stream));
        SLLG_TRY(upload(X.c, d32.data(), d32.size(), c->stream));
        SLLG_TRY(upload(X.dict64, d64.data(), d64.size(), c->stream));
        x.fmt = FMT_SELL_PACK;
        x.ndict = (uint32_t)tuples.size();
        char buf[96];
        std::snprintf(buf, sizeof buf, "sell64x4-pack-t%u-dict%u-slices%uof%u", P.na, x.ndict, stored, nsl);
        X.encoding = buf;
    } else {
        ent.assign(nent, 0u);
        std::vector<float> v32((size_t)nc * nent, 0.0f);
        std::vector<double> v64((size_t)nc * nent, 0.0);
        for (uint32_t s = 0; s < nsl; ++s) {
            const uint32_t len = slice[s].y;
            for (uint32_t l = 0; l < kSlice; ++l) {
                const uint32_t d = s * kSlice + l;
                const uint32_t r = d < n ? ref_of[d] : 0;
                for (uint32_t q = 0; q < len; ++q) {
                    const size_t o = ((size_t)slice[s].x + q) * kSlice + l;
                    if (d < n && q < h.xadj[r + 1] - h.xadj[r]) {
                        const uint32_t qq = h.xadj[r] + q;
                        ent[o] = P.to_dev(h.adj[qq]);
                        for (int a = 0; a < nc; ++a) {
                            v64[(size_t)a * nent + o] = h.vals[(size_t)a * h.nnz + qq];
                            v32[(size_t)a * nent + o] = (float)h.vals[(size_t)a * h.nnz + qq];
                        }
                    } else {
                        ent[o] = d < n ? d : 0;
                    }
                }
            }
        }
        SLLG_TRY(upload(X.b, ent.data(), ent.size(), c->stream));
        SLLG_TRY(upload(X.c, v32.data(), v32.size(), c->stream));
        SLLG_TRY(upload(X.dict64, v64.data(), v64.size(), c->stream));
        x.fmt = FMT_SELL_RAW;
        x.ndict = 0;
        X.encoding = P.na > 1 ? "sell64-u32-f32-sublattice-major" : "sell64-u32-f32";
    }
    SLLG_TRY(upload(X.a, slice.data(), slice.size(), c->stream));
    x.slice = X.a.as<uint2>();
    x.ent = X.b.as<uint32_t>();
    x.dict = X.c.as<float>();
    x.nent = nent;
    x.ncell = ncell;
    x.dbits = pack ? dbits : 0;
    x64 = Exch<double>();
    fill_common(x64, h, n);
    x64.fmt = x.fmt;
    x64.slice = x.slice;
    x64.ent = x.ent;
    x64.dict = X.dict64.as<double>();
    x64.ndict = x.ndict;
    x64.nent = nent;
    x64.ncell = ncell;
    x64.dbits = x.dbits;
    // Line mode replaces only the FP32 kernels' view; the FP64 view (energy, deterministic
    // field) keeps the packed encoding set up above.
    if (pack && !(c->flags & (LLG_FLAG_NO_LINE_MODE | LLG_FLAG_NO_SLICE_SHARING))) {
        uint32_t lstored = 0, lines_row = 0;
        uint32_t nfall = 0;
        bool done = false;
        for (uint32_t len : {64u, 16u})
            for (bool open : {false, true})
                if (!done && build_lines(c, ref_of, idx, tuples, len, open, lstored, lines_row, nfall)) {
                    char buf[128];
                    std::snprintf(buf, sizeof buf, "line%u%s-t%u-slices%uof%u-lines%u-packed%u", len, open ? "o" : "",
                                  P.na, lstored, nsl, lines_row, nfall);
                    if (x.tnw) {  // tile streams: slices per workgroup, most lines staged
                        const size_t l = std::strlen(buf);
                        std::snprintf(buf + l, sizeof buf - l, "-tile%ux%u", x.tnw, x.tumax);
                    }
                    X.encoding = buf;
                    done = true;
                }
    }
    return 0;
}

}  // namespace

int build_exchange_fp32(llg_ctx* c, const llg_exchange& e) {
    ExchStore& X = *c->xs;
    SLLG_TRY(validate_and_copy(c, e));
    const HostExch& h = X.h;
    Exch<float>& x = X.ex32;
    fill_common(x, h, c->n);
    fill_common(X.ex64, h, c->n);
    X.encoding = "none";
    if (h.fmt == LLG_EXCH_CSR) {
        // Sublattice-major device order when the caller declares the basis size (spins
        // ordered cell * n_atoms + t, as the reference orders them).
        const uint32_t na = e.n_atoms;
        if (na > 1 && na <= 16 && c->n % na == 0 && !(c->flags & LLG_FLAG_NO_SUBLATTICE_ORDER)) {
            X.perm.na = na;
            X.perm.ncell = c->n / na;
        } else {
            X.perm.na = 1;
            X.perm.ncell = c->n;
        }
        SLLG_TRY(build_sell(c));
    } else if (h.fmt == LLG_EXCH_DIA || h.fmt == LLG_EXCH_STENCIL) {
        std::vector<float> v32(h.vals.size());
        for (size_t q = 0; q < v32.size(); ++q) v32[q] = (float)h.vals[q];
        SLLG_TRY(upload(X.c, v32.data(), v32.size(), c->stream));
        SLLG_TRY(upload(X.dict64, h.vals.data(), h.vals.size(), c->stream));
        if (h.fmt == LLG_EXCH_DIA) {
            SLLG_TRY(upload(X.a, h.doff.data(), h.doff.size(), c->stream));
            x.doff = X.ex64.doff = X.a.as<int32_t>();
            X.encoding = "dia-f32";
        } else {
            SLLG_TRY(upload_stencil_tables(c));
            x.sent = X.ex64.sent = X.a.as<int4>();
            x.scount = X.ex64.scount = X.b.as<uint32_t>();
            X.encoding = h.per_row ? "stencil-rowvals-f32" : "stencil-shared-f32";
        }
        x.vals = X.c.as<float>();
        X.ex64.vals = X.dict64.as<double>();
    }
    release_host(X.h);
    return 0;
}

// Four-spin quartets: rows and (j, k, l) in the device order of this path; both exchange
// views (the kernels' precision and the FP64 one of the reductions) point at them.
int build_fourspin(llg_ctx* c, const llg_fourspin& f) {
    ExchStore& X = *c->xs;
    const uint32_t n = c->n;
    if (!f.xadj || !f.jkl) SLLG_FAIL(LLG_E_INVALID, "four-spin term without quartets");
    if (f.xadj[0] != 0 || f.xadj[n] != f.nq) SLLG_FAIL(LLG_E_INVALID, "four-spin row pointers do not span the quartets");
    for (uint32_t i = 0; i < n; ++i)
        if (f.xadj[i + 1] < f.xadj[i]) SLLG_FAIL(LLG_E_INVALID, "four-spin row pointers decrease at row %u", i);
    for (uint64_t q = 0; q < 3 * f.nq; ++q)
        if (f.jkl[q] >= n) SLLG_FAIL(LLG_E_INVALID, "four-spin quartet %llu names spin %u", (unsigned long long)(q / 3),
                                    f.jkl[q]);
    const Perm P = X.perm;
    std::vector<uint32_t> xadj(n + 1, 0);
    std::vector<uint4> q4(std::max<uint64_t>(1, f.nq));
    for (uint32_t d = 0; d < n; ++d) {
        const uint32_t i = P.to_ref(d);
        xadj[d + 1] = xadj[d] + (f.xadj[i + 1] - f.xadj[i]);
    }
    for (uint32_t d = 0; d < n; ++d) {
        const uint32_t i = P.to_ref(d);
        uint32_t o = xadj[d];
        for (uint32_t q = f.xadj[i]; q < f.xadj[i + 1]; ++q, ++o)
            q4[o] = make_uint4(P.to_dev(f.jkl[3 * (size_t)q]), P.to_dev(f.jkl[3 * (size_t)q + 1]),
                               P.to_dev(f.jkl[3 * (size_t)q + 2]), 0u);
    }
    SLLG_TRY(upload(X.fs_x, xadj.data(), xadj.size(), c->stream));
    SLLG_TRY(upload(X.fs_q, q4.data(), q4.size(), c->stream));
    X.ex32.fs_xadj = X.ex64.fs_xadj = X.fs_x.as<uint32_t>();
    X.ex32.fs_jkl = X.ex64.fs_jkl = X.fs_q.as<uint4>();
    X.ex32.fs_jq = (float)f.jq;
    X.ex64.fs_jq = f.jq;
    X.ex32.fs_div3 = X.ex64.fs_div3 = f.cpu_form ? 0 : 1;
    return 0;
}

}  // namespace sllg
