// batch.hip - batches of independent spin systems and the continuous-batching request
// loop (include/llg_batch.h).  No reference counterpart: the reference runs one system
// per process; its request types (M(T), field ramp, spin-wave time series) are reproduced by
// llg_serve_run with the sampling semantics of src/mvt.cpp, src/ramp_field.cpp and
// src/timeseries.cpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/llg_batch.h"
#include "../../include/llg_lattice.h"
#include "kernels.hpp"

using namespace sllg;

namespace {

struct Member {
    llg_ctx* c = nullptr;
    uint64_t t = 0;
    uint64_t bytes = 0;
    bool live = false;
    int status = 0;  // LLG_E_NONFINITE once a finite check failed: kept, no longer stepped
    bool active() const { return live && status == 0; }
};

struct Group {
    int fmt = 0, nc = 1;
    std::vector<int> idx;
    DevBuf dmem, dmap;
    uint32_t ntiles = 0;
};

}  // namespace

struct llg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0, used = 0;
    uint32_t flags = 0;
    std::vector<Member> m;
    std::vector<std::unique_ptr<Group>> groups;
    bool dirty = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t gsteps = 0;
    // dipolar refreshes captured in the graph: per live dipolar member (slot, t mod refresh,
    // every-stage); a chunk whose refreshes fall elsewhere is captured anew
    std::vector<uint64_t> gsig;
    uint32_t graphs_built = 0;
    uint32_t tpb = 256;  // threads per tile (one block per tile)
    int xcd = 1;         // XCD-aware block order
    int line = 1;        // line reader for members with a line encoding (else their packed entries)
    int g3 = 0;          // gather from AoS3 copies of S and Y (packed and line readers)
    int persist = 0;     // chunks as persistent launches (grid barriers between half steps)
    int persist_graph = 1;  // with LLG_BATCH_GRAPH: capture them (0 once a capture was refused)
    bool gpersist = false;  // the captured graph holds persistent launches
    uint32_t spin_limit = LLG_BATCH_SPIN_DEFAULT;  // polls per grid-barrier wait
    DevBuf bar;          // GridBar of the persistent launches
    uint32_t persist_grid = 0;  // blocks of the last persistent launch
    DevBuf red;
    std::vector<std::shared_ptr<ExchStore>> pinned;  // coupling stores kept for future members
    std::map<const ExchStore*, uint64_t> store_est;  // admission estimate of each coupling store
};

namespace {
// Device bytes in use: member state plus every distinct coupling store, counted once, at
// the larger of its admission estimate and its size (a store that encodes smaller than
// estimated is still charged what admitting it was checked against, so a budget of one
// request's estimate holds exactly one such request at a time).
uint64_t used_bytes(const llg_batch* b) {
    uint64_t u = b->used;
    std::vector<const ExchStore*> seen;
    auto add = [&](const ExchStore* s) {
        if (s && std::find(seen.begin(), seen.end(), s) == seen.end()) {
            seen.push_back(s);
            const auto it = b->store_est.find(s);
            u += std::max<uint64_t>(s->bytes(), it == b->store_est.end() ? 0 : it->second);
        }
    };
    for (auto& mm : b->m)
        if (mm.live) add(mm.c->xs.get());
    for (auto& s : b->pinned) add(s.get());
    return u;
}
}  // namespace

namespace {

void drop_graph(llg_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->gsteps = 0;
    b->gsig.clear();
    b->gpersist = false;
}

// Where the dipolar refreshes of the next chunk fall: they depend only on each dipolar
// member's step counter modulo its refresh interval (and on every-stage refreshes).
std::vector<uint64_t> refresh_signature(const llg_batch* b) {
    std::vector<uint64_t> sig;
    for (size_t k = 0; k < b->m.size(); ++k) {
        const Member& mm = b->m[k];
        if (!mm.active() || !mm.c->dip) continue;
        sig.push_back(k);
        sig.push_back(mm.t % mm.c->dip_refresh);
        sig.push_back(mm.c->dip_every_stage ? 1u : 0u);
    }
    return sig;
}

int regroup(llg_batch* b) {
    drop_graph(b);
    b->groups.clear();
    std::map<std::pair<int, int>, Group*> by;
    for (size_t k = 0; k < b->m.size(); ++k) {
        if (!b->m[k].active()) continue;
        const llg_ctx* c = b->m[k].c;
        const int f = (is_line_fmt(c->ex32.fmt) && !b->line) ? (int)FMT_SELL_PACK : c->ex32.fmt;
        const auto key = std::make_pair(f, c->ex32.ncomp);
        if (!by.count(key)) {
            b->groups.emplace_back(new Group());
            by[key] = b->groups.back().get();
            by[key]->fmt = key.first;
            by[key]->nc = key.second;
        }
        by[key]->idx.push_back((int)k);
    }
    b->dirty = false;
    return 0;
}

// Member tables with the members' current step counters, temperatures and fields.
int upload_tables(llg_batch* b) {
    for (auto& g : b->groups) {
        std::vector<llg_ctx*> cs;
        std::vector<uint64_t> t0;
        for (int k : g->idx) {
            cs.push_back(b->m[k].c);
            t0.push_back(b->m[k].t);
        }
        SLLG_TRY(batch_build_table(cs, t0, b->tpb, b->g3 != 0, g->dmem, g->dmap, g->ntiles, b->stream));
    }
    return 0;
}

int launch_chunk(llg_batch* b, uint32_t nsteps, bool dipolar) {
    for (auto& m : b->m)
        if (m.c) {  // what the batch kernels leave current: S3 and Y3 with the gather copy, SoA Y without
            m.c->g3_valid = b->g3 != 0;
            m.c->y_soa_valid = !b->g3 || batch_needs_y_soa(m.c);
        }
    for (uint32_t k = 0; k < nsteps; ++k) {
        if (dipolar)
            for (auto& mm : b->m)
                if (mm.active() && mm.c->dip && (mm.t + k) % mm.c->dip_refresh == 0)
                    SLLG_TRY(dipolar_refresh(mm.c, mm.c->S.p, mm.c->DEM.p));
        for (auto& g : b->groups)
            SLLG_TRY(batch_heun_launch(b->stream, g->fmt, g->nc, g->dmem.p, g->dmap.p, g->ntiles, k, 0, b->tpb, b->xcd, b->g3 != 0));
        if (dipolar)
            for (auto& mm : b->m)
                if (mm.active() && mm.c->dip && mm.c->dip_every_stage)
                    SLLG_TRY(dipolar_refresh(mm.c, mm.c->Y.p, mm.c->DEM.p));
        for (auto& g : b->groups)
            SLLG_TRY(batch_heun_launch(b->stream, g->fmt, g->nc, g->dmem.p, g->dmap.p, g->ntiles, k, 1, b->tpb, b->xcd, b->g3 != 0));
    }
    return 0;
}

// Before a persistent chunk is captured: the barrier buffer, and the residency probe of every
// group's kernel shape (it synchronises the stream, so it cannot run inside a capture).
int persist_prepare(llg_batch* b) {
    if (!b->bar.p) {
        SLLG_TRY(b->bar.alloc(sizeof(GridBar)));
        SLLG_HIP(hipMemsetAsync(b->bar.p, 0, sizeof(GridBar), b->stream));
    }
    for (auto& g : b->groups)
        SLLG_TRY(batch_heun_persistent(b->stream, g->fmt, g->nc, g->dmem.p, g->dmap.p, g->ntiles, 0, 0, b->tpb,
                                       b->xcd, b->g3 != 0, static_cast<GridBar*>(b->bar.p), b->spin_limit,
                                       &b->persist_grid));
    SLLG_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

// A chunk as persistent launches: one per group and per run of steps between dipolar
// refreshes (a refresh reads every member's spins, so it runs between launches, before the
// predictor of its step, as in launch_chunk).  Batches with refreshes at the predicted spins
// (every stage) take launch_chunk.
int launch_chunk_persistent(llg_batch* b, uint32_t nsteps, bool dipolar) {
    for (auto& m : b->m)
        if (m.c) {
            m.c->g3_valid = b->g3 != 0;
            m.c->y_soa_valid = !b->g3 || batch_needs_y_soa(m.c);
        }
    if (!b->bar.p) {
        SLLG_TRY(b->bar.alloc(sizeof(GridBar)));
        SLLG_HIP(hipMemsetAsync(b->bar.p, 0, sizeof(GridBar), b->stream));
    }
    auto refresh_at = [&](uint32_t k) {
        for (auto& mm : b->m)
            if (mm.active() && mm.c->dip && (mm.t + k) % mm.c->dip_refresh == 0) return true;
        return false;
    };
    for (uint32_t k0 = 0; k0 < nsteps;) {
        if (dipolar)
            for (auto& mm : b->m)
                if (mm.active() && mm.c->dip && (mm.t + k0) % mm.c->dip_refresh == 0)
                    SLLG_TRY(dipolar_refresh(mm.c, mm.c->S.p, mm.c->DEM.p));
        uint32_t k1 = k0 + 1;
        if (dipolar)
            while (k1 < nsteps && !refresh_at(k1)) ++k1;
        else
            k1 = nsteps;
        for (auto& g : b->groups)
            SLLG_TRY(batch_heun_persistent(b->stream, g->fmt, g->nc, g->dmem.p, g->dmap.p, g->ntiles, k0, k1 - k0,
                                           b->tpb, b->xcd, b->g3 != 0, static_cast<GridBar*>(b->bar.p),
                                           b->spin_limit, &b->persist_grid));
        k0 = k1;
    }
    return 0;
}

int check_member(llg_batch* b, int k) {
    if (!b || k < 0 || (size_t)k >= b->m.size() || !b->m[k].live) SLLG_FAIL(LLG_E_INVALID, "no such member %d", k);
    return 0;
}

}  // namespace

extern "C" {

uint64_t llg_estimate_bytes(const llg_desc* d) {
    if (!d || d->n == 0) return 0;
    const uint64_t n = d->n;
    const bool f64 = d->precision == LLG_FP64_PLAIN;
    const uint64_t P = f64 ? 8 : 4;
    const uint64_t nc = d->exch.ncomp ? (uint64_t)d->exch.ncomp : 1;
    uint64_t b = 4 * 3 * n * P + 2 * n + 1024 * 64;  // S, Y, A|H, STG; species/sublattice; tables
    if (f64 || d->noise == LLG_NOISE_INJECTED) b += 3 * n * 8;  // XI
    if (f64) b += 3 * 3 * n * 8 + 8 * n * 8;                      // FN, DET, H; per-spin parameters
    if (d->scheme != LLG_HEUN) b += 6 * 3 * n * P;
    if (d->exch.format == LLG_EXCH_CSR) {
        const uint64_t ent = d->exch.nnz + 4 * n;  // padding of slices
        b += f64 ? (n + 1) * 4 + d->exch.nnz * (4 + 8 * nc) : ent * (4 + 4 * nc) + ent * 8 * nc + (n / 64 + 2) * 4;
    } else if (d->exch.format == LLG_EXCH_DIA) {
        b += (uint64_t)d->exch.ndiag * nc * n * (P + 8);
    } else if (d->exch.format == LLG_EXCH_STENCIL) {
        b += d->exch.stencil_vals_per_row ? (uint64_t)d->exch.stencil_k * nc * n * (P + 8) : 4096;
    }
    const bool fftx = d->fftx.n_int > 0;
    if (d->dip.enabled || fftx) {
        uint64_t G = 1, Gh = 1;
        for (int a = 0; a < 3; ++a) G *= 2ull * d->dip.dim[a] * d->dip.nk[a];
        Gh = G / (2ull * d->dip.dim[2] * d->dip.nk[2]) * (d->dip.dim[2] * d->dip.nk[2] + 1);
        if (fftx) {
            // one grid and spectrum pair per species, the complex 3x3 kernel per species pair,
            // and the FP64 real/complex kernel buffers that live only while it is transformed
            const uint64_t nms = d->n_species ? d->n_species : 1;
            b += 3 * n * P + 2 * 3 * nms * G * P + 2 * 3 * nms * Gh * 2 * P + nms * nms * 9 * Gh * 2 * P + 2 * n * 4;
            b += nms * nms * 9 * (G * 8 + Gh * 16);
        } else {
            b += 3 * n * P + 2 * 3 * G * P + 3 * Gh * 2 * P + 6 * Gh * P + 2 * n * 4;
        }
    }
    return b;
}

int llg_batch_create(int device, uint64_t hbm_budget_bytes, uint32_t flags, llg_batch** out) {
    if (!out) SLLG_FAIL(LLG_E_INVALID, "NULL output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) SLLG_FAIL(LLG_E_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) SLLG_FAIL(LLG_E_INVALID, "device %d not in [0, %d)", device, ndev);
    SLLG_HIP(hipSetDevice(device));
    llg_batch* b = new llg_batch();
    b->device = device;
    b->budget = hbm_budget_bytes;
    b->flags = flags;
    b->persist = (flags & LLG_BATCH_PERSISTENT) ? 1 : 0;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        SLLG_FAIL(LLG_E_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

void llg_batch_destroy(llg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto& mm : b->m)
        if (mm.c) llg_destroy(mm.c);
    b->groups.clear();
    b->red.release();
    (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"

// Admission of one member; with `store` the couplings are shared and not counted again.
static int batch_add_impl(llg_batch* b, const llg_desc* d, const std::shared_ptr<ExchStore>& store,
                          const double* sx, const double* sy, const double* sz, double T, const double B[3],
                          uint64_t t0, int* member) {
    if (!b || !d || !sx || !sy || !sz || !member) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (d->precision != LLG_FP32 || d->scheme != LLG_HEUN || d->noise != LLG_NOISE_COUNTER)
        SLLG_FAIL(LLG_E_UNSUPPORTED, "batch members are FP32 Heun handles with counter noise");
    llg_desc ds = *d;
    ds.exch.format = LLG_EXCH_NONE;
    const uint64_t est_state = llg_estimate_bytes(&ds);
    const uint64_t est = est_state + (store ? 0 : llg_estimate_bytes(d) - est_state);
    const uint64_t used = used_bytes(b);
    if (b->budget && used + est > b->budget)
        SLLG_FAIL(LLG_E_NOMEM, "HBM budget: %llu + %llu > %llu bytes", (unsigned long long)used,
                  (unsigned long long)est, (unsigned long long)b->budget);
    SLLG_HIP(hipSetDevice(b->device));
    SLLG_HIP(hipStreamSynchronize(b->stream));
    llg_ctx* c = nullptr;
    if (store)
        SLLG_TRY(create_with_store(d, store, &c));
    else
        SLLG_TRY(llg_create(d, &c));
    int rc = llg_set_spins(c, sx, sy, sz);
    if (!rc) rc = llg_set_temperature(c, T);
    if (!rc && B) rc = llg_set_applied(c, B);
    if (rc) {
        llg_destroy(c);
        return rc;
    }
    // the member runs on the batch stream from now on
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    c->stream = b->stream;
    c->own_stream = false;
    rc = dipolar_set_stream(c);
    if (rc) {
        llg_destroy(c);
        return rc;
    }
    llg_info info;
    (void)llg_get_info(c, &info);
    Member mm;
    mm.c = c;
    mm.t = t0;
    mm.bytes = std::max<uint64_t>(est_state, info.device_bytes - c->xs->bytes());  // state only
    mm.live = true;
    if (!store) b->store_est[c->xs.get()] = est - est_state;
    int slot = -1;
    for (size_t k = 0; k < b->m.size(); ++k)
        if (!b->m[k].live && !b->m[k].c) { slot = (int)k; break; }
    if (slot < 0) {
        slot = (int)b->m.size();
        b->m.push_back(mm);
    } else {
        b->m[slot] = mm;
    }
    b->used += mm.bytes;
    b->dirty = true;
    *member = slot;
    return 0;
}

extern "C" {

int llg_batch_add(llg_batch* b, const llg_desc* d, const double* sx, const double* sy, const double* sz, double T,
                  const double B[3], uint64_t t0, int* member) {
    return batch_add_impl(b, d, nullptr, sx, sy, sz, T, B, t0, member);
}

int llg_batch_add_like(llg_batch* b, const llg_desc* d, int proto_member, const double* sx, const double* sy,
                       const double* sz, double T, const double B[3], uint64_t t0, int* member) {
    SLLG_TRY(check_member(b, proto_member));
    return batch_add_impl(b, d, b->m[proto_member].c->xs, sx, sy, sz, T, B, t0, member);
}

int llg_batch_remove(llg_batch* b, int member) {
    SLLG_TRY(check_member(b, member));
    SLLG_HIP(hipStreamSynchronize(b->stream));
    Member& mm = b->m[member];
    llg_destroy(mm.c);
    b->used -= std::min(b->used, mm.bytes);
    mm = Member();
    b->dirty = true;
    return 0;
}

int llg_batch_set_temperature(llg_batch* b, int member, double T) {
    SLLG_TRY(check_member(b, member));
    return llg_set_temperature(b->m[member].c, T);
}

int llg_batch_set_applied(llg_batch* b, int member, const double B[3]) {
    SLLG_TRY(check_member(b, member));
    return llg_set_applied(b->m[member].c, B);
}

int llg_batch_step(llg_batch* b, uint32_t nsteps) {
    if (!b) SLLG_FAIL(LLG_E_INVALID, "NULL batch");
    if (nsteps == 0) return 0;
    SLLG_HIP(hipSetDevice(b->device));
    if (b->dirty) SLLG_TRY(regroup(b));
    bool dip = false, dip_graph = true, dip_stage = false;
    for (auto& mm : b->m)
        if (mm.active() && mm.c->dip) {
            dip = true;
            // refreshes made of this library's own kernels only (no hipFFT plan) are captured
            dip_graph = dip_graph && dipolar_capturable(mm.c);
            dip_stage = dip_stage || mm.c->dip_every_stage;
        }
    SLLG_TRY(upload_tables(b));
    if (b->persist && !dip_stage) {
        bool replayed = false;
        if ((b->flags & LLG_BATCH_GRAPH) && dip_graph && b->persist_graph) {
            // one hipGraph per chunk shape: barrier resets, cooperative launches and the
            // refreshes between them
            std::vector<uint64_t> sig = refresh_signature(b);
            if (!b->gexec || !b->gpersist || b->gsteps != nsteps || sig != b->gsig) {
                drop_graph(b);
                SLLG_TRY(persist_prepare(b));
                SLLG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                const int rc = launch_chunk_persistent(b, nsteps, dip);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->stream, &g);
                hipGraphExec_t ge = nullptr;
                if (!rc && e == hipSuccess && g && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess) {
                    b->graph = g;
                    b->gexec = ge;
                    b->gsteps = nsteps;
                    b->gsig.swap(sig);
                    b->gpersist = true;
                    b->graphs_built++;
                } else {
                    // the runtime would not capture the cooperative launch: launch directly
                    // from now on (the capture left nothing queued on the stream)
                    if (g) (void)hipGraphDestroy(g);
                    (void)hipGetLastError();
                    b->persist_graph = 0;
                }
            }
            if (b->gexec && b->gpersist) {
                SLLG_HIP(hipGraphLaunch(b->gexec, b->stream));
                replayed = true;
            }
        }
        if (!replayed) SLLG_TRY(launch_chunk_persistent(b, nsteps, dip));
        GridBar hb{};
        SLLG_HIP(hipMemcpyAsync(&hb, b->bar.p, sizeof(hb), hipMemcpyDeviceToHost, b->stream));
        SLLG_HIP(hipStreamSynchronize(b->stream));
        if (hb.abort) {
            SLLG_HIP(hipMemsetAsync(b->bar.p, 0, sizeof(GridBar), b->stream));
            SLLG_HIP(hipStreamSynchronize(b->stream));
            SLLG_FAIL(LLG_E_DEVICE,
                      "persistent launch abandoned: a grid barrier wait passed %u polls (the members' spins are "
                      "partly advanced; their step counters are unchanged)", b->spin_limit);
        }
    } else if ((b->flags & LLG_BATCH_GRAPH) && dip_graph) {
        std::vector<uint64_t> sig = refresh_signature(b);
        if (!b->gexec || b->gpersist || b->gsteps != nsteps || sig != b->gsig) {
            drop_graph(b);
            SLLG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int rc = launch_chunk(b, nsteps, dip);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) SLLG_FAIL(LLG_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
            b->graph = g;
            SLLG_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
            b->gsteps = nsteps;
            b->gsig.swap(sig);
            b->graphs_built++;
        }
        SLLG_HIP(hipGraphLaunch(b->gexec, b->stream));
    } else {
        SLLG_TRY(launch_chunk(b, nsteps, dip));
    }
    for (auto& mm : b->m)
        if (mm.active()) mm.t += nsteps;
    // members that asked for finite checks (LLG_FLAG_CHECK_FINITE): one with a NaN/Inf spin is
    // taken out of the launches, keeps its state and step counter, and reports the failure
    std::string bad;
    for (size_t k = 0; k < b->m.size(); ++k) {
        Member& mm = b->m[k];
        if (!mm.active() || !(mm.c->flags & LLG_FLAG_CHECK_FINITE)) continue;
        int nf = 0;
        SLLG_TRY(check_finite(mm.c, &nf));
        if (nf) {
            mm.status = LLG_E_NONFINITE;
            b->dirty = true;
            bad += (bad.empty() ? "" : ", ") + std::to_string(k);
        }
    }
    if (!bad.empty()) SLLG_FAIL(LLG_E_NONFINITE, "non-finite spins in batch member(s) %s", bad.c_str());
    return 0;
}

int llg_batch_member_status(llg_batch* b, int member, int* status) {
    SLLG_TRY(check_member(b, member));
    if (!status) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    *status = b->m[member].status;
    return 0;
}

int llg_batch_set_launch(llg_batch* b, uint32_t tpb, int xcd, int line_mode, int gather_copy) {
    if (!b) SLLG_FAIL(LLG_E_INVALID, "NULL batch");
    if (tpb != 128 && tpb != 256 && tpb != 512) SLLG_FAIL(LLG_E_INVALID, "tile of %u threads (128, 256, 512)", tpb);
    // the captured launches have the old shape
    if (tpb != b->tpb || (xcd ? 1 : 0) != b->xcd || (gather_copy ? 1 : 0) != b->g3) drop_graph(b);
    b->g3 = gather_copy ? 1 : 0;
    if ((line_mode ? 1 : 0) != b->line) b->dirty = true;        // members regroup by the reader they run
    b->tpb = tpb;
    b->xcd = xcd ? 1 : 0;
    b->line = line_mode ? 1 : 0;
    return 0;
}

int llg_batch_set_persistent(llg_batch* b, int on, uint32_t spin_limit) {
    if (!b) SLLG_FAIL(LLG_E_INVALID, "NULL batch");
    // a captured chunk holds the poll limit among its launch arguments
    if ((on ? 1 : 0) != b->persist || spin_limit != b->spin_limit) drop_graph(b);
    b->persist = on ? 1 : 0;
    b->spin_limit = spin_limit;
    return 0;
}

int llg_batch_autotune(llg_batch* b, uint32_t nsteps, llg_batch_tune* out) {
    if (!b || !out || nsteps == 0) SLLG_FAIL(LLG_E_INVALID, "NULL argument or nsteps = 0");
    std::memset(out, 0, sizeof(*out));
    SLLG_HIP(hipSetDevice(b->device));
    // save every live member's spins and step counter; the timed steps are undone afterwards
    // (and the dipolar field cached between refreshes); every candidate starts from the saved
    // step counters, so its timed window holds the same dipolar refreshes as every other's
    std::vector<DevBuf> save(b->m.size()), save_dem(b->m.size());
    std::vector<uint64_t> t_save(b->m.size());
    for (size_t k = 0; k < b->m.size(); ++k) {
        if (!b->m[k].active()) continue;
        llg_ctx* c = b->m[k].c;
        SLLG_TRY(sync_spins(c));
        SLLG_TRY(save[k].alloc(c->S.bytes));
        SLLG_HIP(hipMemcpyAsync(save[k].p, c->S.p, c->S.bytes, hipMemcpyDeviceToDevice, b->stream));
        if (c->dip) {
            SLLG_TRY(save_dem[k].alloc(c->DEM.bytes));
            SLLG_HIP(hipMemcpyAsync(save_dem[k].p, c->DEM.p, c->DEM.bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        t_save[k] = b->m[k].t;
    }
    auto rewind = [&]() {
        for (size_t k = 0; k < b->m.size(); ++k)
            if (b->m[k].active()) b->m[k].t = t_save[k];
    };
    const uint32_t tpb0 = b->tpb;
    const int xcd0 = b->xcd, line0 = b->line, g30 = b->g3;
    bool any_line = false;
    for (auto& mm : b->m) any_line = any_line || (mm.active() && is_line_fmt(mm.c->ex32.fmt));
    const int nline = any_line ? 2 : 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SLLG_HIP(hipEventCreate(&e0));
    SLLG_HIP(hipEventCreate(&e1));
    const uint32_t tpbs[3] = {128, 256, 512};
    double best = 1e300;
    uint32_t btpb = 256;
    int bxcd = 1, bline = line0, bg3 = 0, rc = 0;
    struct Cand {
        double per;
        uint32_t tpb;
        int xcd, line, g3;
    };
    std::vector<Cand> timed;  // every candidate's first timing
    for (int li = 0; li < nline && !rc; ++li)
    for (int gi = 0; gi < 2 && !rc; ++gi)
    for (int ti = 0; ti < 3 && !rc; ++ti)
        for (int xcd = 1; xcd >= 0 && !rc; --xcd) {
            const int line = nline == 2 ? 1 - li : line0;
            rc = llg_batch_set_launch(b, tpbs[ti], xcd, line, gi);
            rewind();
            if (!rc) rc = llg_batch_step(b, 2);  // warm-up (and graph capture when enabled)
            if (!rc) rc = llg_batch_step(b, nsteps);  // a graph of this length is captured here
            if (rc) break;
            if (hipEventRecord(e0, b->stream) != hipSuccess) rc = LLG_E_DEVICE;
            if (!rc) rc = llg_batch_step(b, nsteps);
            if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = LLG_E_DEVICE;
            if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = LLG_E_DEVICE;
            float ms = 0.f;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = LLG_E_DEVICE;
            if (rc) break;
            const double per = (double)ms / nsteps;
            if (tpbs[ti] == 256 && xcd == 1 && li == 0 && gi == 0) out->ms_per_step_default = per;
            out->candidates++;
            timed.push_back(Cand{per, tpbs[ti], xcd, line, gi});
            if (per < best) {
                best = per;
                btpb = tpbs[ti];
                bxcd = xcd;
                bline = line;
                bg3 = gi;
            }
        }
    // the fastest few are timed again, in turn, over twice as many steps in two passes (the
    // minimum of many short noisy timings favours a lucky candidate)
    if (!rc && timed.size() > 1) {
        std::sort(timed.begin(), timed.end(), [](const Cand& x, const Cand& y) { return x.per < y.per; });
        const size_t top = std::min<size_t>(3, timed.size());
        std::vector<double> acc(top, 0.0);
        for (int pass = 0; pass < 2 && !rc; ++pass)
            for (size_t k = 0; k < top && !rc; ++k) {
                const Cand& t = timed[k];
                rc = llg_batch_set_launch(b, t.tpb, t.xcd, t.line, t.g3);
                rewind();
                if (!rc) rc = llg_batch_step(b, 2);
                if (!rc) rc = llg_batch_step(b, 2 * nsteps);  // captures a graph of this length
                if (!rc && hipEventRecord(e0, b->stream) != hipSuccess) rc = LLG_E_DEVICE;
                if (!rc) rc = llg_batch_step(b, 2 * nsteps);
                if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = LLG_E_DEVICE;
                if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = LLG_E_DEVICE;
                float ms = 0.f;
                if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = LLG_E_DEVICE;
                acc[k] += (double)ms / (2.0 * nsteps);
            }
        if (!rc) {
            size_t kb = 0;
            for (size_t k = 1; k < top; ++k)
                if (acc[k] < acc[kb]) kb = k;
            btpb = timed[kb].tpb;
            bxcd = timed[kb].xcd;
            bline = timed[kb].line;
            bg3 = timed[kb].g3;
            best = acc[kb] / 2.0;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // restore the members' spins and step counters
    for (size_t k = 0; k < b->m.size(); ++k) {
        if (!b->m[k].active()) continue;
        llg_ctx* c = b->m[k].c;
        hipError_t e = hipMemcpyAsync(c->S.p, save[k].p, c->S.bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e != hipSuccess && !rc) rc = LLG_E_DEVICE;
        if (c->dip) {
            e = hipMemcpyAsync(c->DEM.p, save_dem[k].p, c->DEM.bytes, hipMemcpyDeviceToDevice, b->stream);
            if (e != hipSuccess && !rc) rc = LLG_E_DEVICE;
        }
        c->g3_valid = false;  // S changed under the AoS3 copy: rebuilt before the next gather-copy launch
        c->s_soa_valid = true;
        b->m[k].t = t_save[k];
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess && !rc) rc = LLG_E_DEVICE;
    if (rc) {
        (void)llg_batch_set_launch(b, tpb0, xcd0, line0, g30);
        return rc;
    }
    SLLG_TRY(llg_batch_set_launch(b, btpb, bxcd, bline, bg3));
    out->tpb = (int)btpb;
    out->xcd = bxcd;
    out->line_mode = bline;
    out->gather_copy = bg3;
    out->ms_per_step = best;
    return 0;
}

int llg_batch_reduce(llg_batch* b, const int* members, uint32_t count, uint32_t stride, double* m) {
    if (!b || (count && (!members || !m))) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    if (count == 0) return 0;
    SLLG_HIP(hipSetDevice(b->device));
    const size_t need = (size_t)count * stride * sizeof(double);
    if (b->red.bytes < need) SLLG_TRY(b->red.alloc(need * 2));
    for (uint32_t k = 0; k < count; ++k) {
        SLLG_TRY(check_member(b, members[k]));
        llg_ctx* c = b->m[members[k]].c;
        if (c->nsub * 3 > stride) SLLG_FAIL(LLG_E_INVALID, "stride %u < 3 x sublattices of member %d", stride, members[k]);
        SLLG_TRY(reduce_mag(c, b->red.as<double>() + (size_t)k * stride));
    }
    SLLG_HIP(hipMemcpyAsync(m, b->red.p, need, hipMemcpyDeviceToHost, b->stream));
    SLLG_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

int llg_batch_get_spins(llg_batch* b, int member, double* sx, double* sy, double* sz) {
    SLLG_TRY(check_member(b, member));
    return llg_get_spins(b->m[member].c, sx, sy, sz);
}

int llg_batch_get_step(llg_batch* b, int member, uint64_t* t) {
    SLLG_TRY(check_member(b, member));
    if (!t) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    *t = b->m[member].t;
    return 0;
}

int llg_batch_get_info(llg_batch* b, llg_batch_info* info) {
    if (!b || !info) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    for (auto& mm : b->m)
        if (mm.live) {
            info->members++;
            info->spins += mm.c->n;
            info->tiles += blocks_for(mm.c->n, kTPB);
        }
    info->device_bytes = used_bytes(b);
    info->budget_bytes = b->budget;
    info->graphs_built = b->graphs_built;
    info->persist_grid = b->persist_grid;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- request loop

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// Simple-cubic L^3 periodic lattice with nearest-neighbour exchange J (CSR in the
// reference order, built by the lattice builder).  Shared by all requests of one
// (L, J, mu).
struct ScLattice {
    std::vector<uint32_t> xadj, adj;
    std::vector<double> vals;
};

int build_sc(uint32_t L, double J, double mu, ScLattice& out) {
    const double xred[3] = {0, 0, 0};
    const uint32_t mp[3] = {0, 0, 0};
    llg_unitcell uc;
    std::memset(&uc, 0, sizeof uc);
    uc.n_atoms = 1;
    uc.nk[0] = uc.nk[1] = uc.nk[2] = 1;
    uc.rprim[0] = uc.rprim[4] = uc.rprim[8] = 1.0;
    uc.xred = xred;
    uc.mp = mp;
    uint32_t nint = 0;
    SLLG_TRY(llg_lattice_shells(&uc, 1, &J, nullptr, 0, &nint, nullptr));
    std::vector<llg_interaction> ints(nint);
    SLLG_TRY(llg_lattice_shells(&uc, 1, &J, ints.data(), nint, &nint, nullptr));
    const uint32_t dims[3] = {L, L, L}, pbc[3] = {1, 1, 1};
    uint64_t nnz = 0;
    SLLG_TRY(llg_lattice_build_csr(&uc, dims, pbc, nint, ints.data(), &mu, 1, nullptr, nullptr, nullptr, &nnz));
    const uint64_t n = (uint64_t)L * L * L;
    out.xadj.resize(n + 1);
    out.adj.resize(nnz);
    out.vals.resize(nnz);
    SLLG_TRY(llg_lattice_build_csr(&uc, dims, pbc, nint, ints.data(), &mu, 1, out.xadj.data(), out.adj.data(),
                                   out.vals.data(), &nnz));
    return 0;
}

struct Live {
    int req = -1, member = -1;
    uint32_t done = 0;
    uint32_t samples = 0;
    double msum[3] = {0, 0, 0};
    uint32_t mcount = 0;
    llg_ctx* solo = nullptr;  // sequential mode
};

void ramp_field(const llg_request& r, uint32_t done, double B[3]) {
    const double f = r.steps ? (double)done / (double)r.steps : 0.0;
    for (int a = 0; a < 3; ++a) B[a] = r.B0[a] + (r.B1[a] - r.B0[a]) * f;
}

void record_sample(const llg_request& r, Live& lv, const double m[3], llg_result& res, double* m_series,
                   const uint64_t* series_offset) {
    if (m_series && series_offset)
        for (int a = 0; a < 3; ++a) m_series[(series_offset[lv.req] + lv.samples) * 3 + a] = m[a];
    const uint32_t nsamp = (r.steps + r.update - 1) / r.update;
    if (lv.samples >= nsamp / 2) {
        for (int a = 0; a < 3; ++a) lv.msum[a] += m[a];
        lv.mcount++;
    }
    lv.samples++;
    for (int a = 0; a < 3; ++a) res.m_final[a] = m[a];
}

}  // namespace

extern "C" int llg_serve_run(const llg_serve_config* cfg, const llg_request* reqs, uint32_t nreq, llg_result* results,
                             double* m_series, const uint64_t* series_offset, double* sq_series,
                             const uint64_t* sq_offset, llg_serve_stats* stats) {
    if (!cfg || (nreq && (!reqs || !results)) || !stats) SLLG_FAIL(LLG_E_INVALID, "NULL argument");
    for (uint32_t i = 0; i < nreq; ++i) {
        if (reqs[i].L == 0 || reqs[i].update == 0 || reqs[i].steps == 0 || !(reqs[i].dt > 0) || !(reqs[i].mu > 0))
            SLLG_FAIL(LLG_E_INVALID, "request %u: L, steps, update, dt and mu must be positive", i);
        if (reqs[i].type > LLG_REQ_SW) SLLG_FAIL(LLG_E_INVALID, "request %u: unknown type %u", i, reqs[i].type);
    }
    if (sq_series && !sq_offset) SLLG_FAIL(LLG_E_INVALID, "sq_series needs sq_offset");
    // spin-wave requests: S(k) at (q, 0, 0), q = 0 .. L/2, on the L^3 mesh (spin (x L + y) L + z)
    std::map<uint32_t, std::vector<uint32_t>> meshes;
    auto sw_setup = [&](const llg_request& r, llg_ctx* c) -> int {
        if (r.type != LLG_REQ_SW) return 0;
        std::vector<uint32_t>& mesh = meshes[r.L];
        const uint64_t n = (uint64_t)r.L * r.L * r.L;
        if (mesh.empty()) {
            mesh.resize(3 * n);
            for (uint64_t i = 0; i < n; ++i) {
                mesh[3 * i] = (uint32_t)(i / ((uint64_t)r.L * r.L));
                mesh[3 * i + 1] = (uint32_t)((i / r.L) % r.L);
                mesh[3 * i + 2] = (uint32_t)(i % r.L);
            }
        }
        const uint32_t nk = r.L / 2 + 1, grid[3] = {r.L, r.L, r.L};
        std::vector<uint32_t> kp(3 * nk, 0u);
        for (uint32_t q = 0; q < nk; ++q) kp[3 * q] = q;
        return llg_set_sq(c, grid, mesh.data(), nk, kp.data(), 2, nullptr, nullptr, 1, nullptr);
    };
    auto sw_sample = [&](const llg_request& r, uint32_t i, uint32_t sample, llg_ctx* c) -> int {
        if (r.type != LLG_REQ_SW) return 0;
        const uint32_t nk = r.L / 2 + 1;
        std::vector<double> v(2 * nk);
        SLLG_TRY(llg_sample_sq(c, v.data()));
        if (sq_series) std::memcpy(sq_series + sq_offset[i] + (uint64_t)sample * 2 * nk, v.data(), v.size() * 8);
        return 0;
    };
    std::memset(stats, 0, sizeof(*stats));
    std::memset(results, 0, sizeof(llg_result) * nreq);
    std::vector<uint32_t> order(nreq);
    for (uint32_t i = 0; i < nreq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::map<std::tuple<uint32_t, double, double>, std::unique_ptr<ScLattice>> lattices;
    auto lattice_of = [&](const llg_request& r, ScLattice** out) -> int {
        auto key = std::make_tuple(r.L, r.J, r.mu);
        auto it = lattices.find(key);
        if (it == lattices.end()) {
            std::unique_ptr<ScLattice> L(new ScLattice());
            SLLG_TRY(build_sc(r.L, r.J, r.mu, *L));
            it = lattices.emplace(key, std::move(L)).first;
        }
        *out = it->second.get();
        return 0;
    };
    auto make_desc = [&](const llg_request& r, const ScLattice& lat, llg_species& sp, llg_desc& d, int prec) {
        llg_desc_init(&d);
        std::memset(&sp, 0, sizeof sp);
        sp.mu = r.mu;
        sp.lambda = r.lambda;
        sp.gamma_rel = 1.0;
        sp.k1u = r.K;
        sp.k1u_dir[2] = 1.0;
        d.n = r.L * r.L * r.L;
        d.n_species = 1;
        d.species = &sp;
        d.n_sublattice = 1;
        d.exch.format = LLG_EXCH_CSR;
        d.exch.ncomp = 1;
        d.exch.nnz = lat.adj.size();
        d.exch.xadj = lat.xadj.data();
        d.exch.adjncy = lat.adj.data();
        d.exch.csr_vals = lat.vals.data();
        d.dt = r.dt;
        d.scheme = LLG_HEUN;
        d.precision = prec;
        d.noise = LLG_NOISE_COUNTER;
        d.update = r.update;
        d.seed = r.seed;
        d.device = cfg->device;
    };
    const auto t_start = Clock::now();
    std::vector<double> lat_hist;
    uint64_t su = 0;
    uint32_t completed = 0, max_live = 0, reported = 0;
    // the caller's progress hook, after every progress_every-th completion and the last one
    auto progress = [&]() {
        if (!cfg->progress || !cfg->progress_every) return;
        if (completed == nreq || completed / cfg->progress_every > reported / cfg->progress_every) {
            cfg->progress(completed, nreq, since(t_start), cfg->progress_user);
            reported = completed;
        }
    };

    if (cfg->sequential) {
        // one request at a time, in arrival order, alone on the device
        for (uint32_t oi = 0; oi < nreq; ++oi) {
            const uint32_t i = order[oi];
            const llg_request& r = reqs[i];
            while (since(t_start) < r.arrival_s) std::this_thread::sleep_for(std::chrono::microseconds(200));
            llg_result& res = results[i];
            res.arrival_s = r.arrival_s;
            res.admit_s = since(t_start);
            ScLattice* lat = nullptr;
            SLLG_TRY(lattice_of(r, &lat));
            llg_species sp;
            llg_desc d;
            make_desc(r, *lat, sp, d, cfg->precision);
            llg_ctx* c = nullptr;
            SLLG_TRY(llg_create(&d, &c));
            std::vector<double> zx(d.n, 0.0), zz(d.n, 1.0);
            int rc = llg_set_spins(c, zx.data(), zx.data(), zz.data());
            if (!rc) rc = llg_set_temperature(c, r.T);
            if (!rc) rc = sw_setup(r, c);
            Live lv;
            lv.req = (int)i;
            for (uint32_t s = 0; !rc && s < r.steps; s += r.update) {
                double B[3];
                ramp_field(r, s, B);
                if (r.type != LLG_REQ_RAMP) std::memcpy(B, r.B0, sizeof B);
                double m[3];
                rc = llg_reduce(c, m, nullptr);
                if (!rc) rc = sw_sample(r, i, lv.samples, c);
                if (rc) break;
                record_sample(r, lv, m, res, m_series, series_offset);
                rc = llg_set_applied(c, B);
                if (!rc) rc = llg_step(c, s, std::min(r.update, r.steps - s));
            }
            double m[3];
            if (!rc) rc = llg_reduce(c, m, nullptr);
            llg_destroy(c);
            if (rc) return rc;
            for (int a = 0; a < 3; ++a) {
                res.m_final[a] = m[a];
                res.m_mean[a] = lv.mcount ? lv.msum[a] / lv.mcount : m[a];
            }
            res.samples = lv.samples;
            res.done_s = since(t_start);
            res.latency_s = res.done_s - r.arrival_s;
            lat_hist.push_back(res.latency_s);
            su += (uint64_t)d.n * r.steps;
            completed++;
            max_live = 1;
            progress();
        }
    } else {
        if (cfg->precision != LLG_FP32) SLLG_FAIL(LLG_E_UNSUPPORTED, "batched serving runs the FP32 path");
        llg_batch* b = nullptr;
        SLLG_TRY(llg_batch_create(cfg->device, cfg->hbm_budget_bytes, cfg->flags, &b));
        std::unique_ptr<llg_batch, void (*)(llg_batch*)> guard(b, llg_batch_destroy);
        std::vector<Live> live;
        std::map<std::tuple<uint32_t, double, double>, std::shared_ptr<ExchStore>> stores;
        uint32_t next = 0;
        uint32_t min_update = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < nreq; ++i) min_update = std::min(min_update, reqs[i].update);
        const uint32_t chunk_cap = cfg->chunk ? cfg->chunk : min_update;
        while (completed < nreq) {
            const double now = since(t_start);
            // admission: arrived requests, FIFO, while slots and memory allow, and only while every
            // live request sits at one of its sampling steps: a request admitted then starts in
            // phase with the others, so launch chunks keep the full sampling interval instead of
            // shrinking to the gaps between the requests' sampling steps
            bool in_phase = true;
            uint64_t live_spins = 0;
            for (const auto& lv : live) {
                in_phase = in_phase && lv.done % reqs[lv.req].update == 0;
                live_spins += (uint64_t)reqs[lv.req].L * reqs[lv.req].L * reqs[lv.req].L;
            }
            while (in_phase && next < nreq && reqs[order[next]].arrival_s <= now) {
                if (cfg->max_members && live.size() >= cfg->max_members) break;
                const uint32_t i = order[next];
                const llg_request& r = reqs[i];
                const uint64_t nspins = (uint64_t)r.L * r.L * r.L;
                if (cfg->max_live_spins && !live.empty() && live_spins + nspins > cfg->max_live_spins) break;
                live_spins += nspins;
                ScLattice* lat = nullptr;
                SLLG_TRY(lattice_of(r, &lat));
                llg_species sp;
                llg_desc d;
                make_desc(r, *lat, sp, d, LLG_FP32);
                std::vector<double> zx(d.n, 0.0), zz(d.n, 1.0);
                double B[3];
                std::memcpy(B, r.B0, sizeof B);
                int member = -1;
                const auto skey = std::make_tuple(r.L, r.J, r.mu);
                auto sit = stores.find(skey);
                const std::shared_ptr<ExchStore> st = sit == stores.end() ? nullptr : sit->second;
                const int rc = batch_add_impl(b, &d, st, zx.data(), zx.data(), zz.data(), r.T, B, 0, &member);
                if (rc == 0 && !st) {  // first request of this lattice: keep its couplings for later ones
                    stores[skey] = b->m[member].c->xs;
                    b->pinned.push_back(b->m[member].c->xs);
                }
                if (rc == LLG_E_NOMEM) {
                    if (live.empty()) return rc;  // does not fit even alone
                    break;                        // wait for memory to free up
                }
                if (rc) return rc;
                SLLG_TRY(sw_setup(r, b->m[member].c));
                Live lv;
                lv.req = (int)i;
                lv.member = member;
                live.push_back(lv);
                results[i].arrival_s = r.arrival_s;
                results[i].admit_s = since(t_start);
                ++next;
            }
            max_live = std::max<uint32_t>(max_live, (uint32_t)live.size());
            if (live.empty()) {
                if (next < nreq) {
                    const double wait = reqs[order[next]].arrival_s - since(t_start);
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
                }
                continue;
            }
            // samples due before this chunk (every `update` steps, before the step)
            std::vector<int> due;
            for (auto& lv : live)
                if (lv.done % reqs[lv.req].update == 0) due.push_back(lv.member);
            if (!due.empty()) {
                std::vector<double> m(due.size() * 3);
                SLLG_TRY(llg_batch_reduce(b, due.data(), (uint32_t)due.size(), 3, m.data()));
                size_t k = 0;
                for (auto& lv : live)
                    if (lv.done % reqs[lv.req].update == 0) {
                        // the batch's launches are complete (the reduction synchronised its stream)
                        SLLG_TRY(sw_sample(reqs[lv.req], (uint32_t)lv.req, lv.samples, b->m[lv.member].c));
                        record_sample(reqs[lv.req], lv, &m[3 * k], results[lv.req], m_series, series_offset);
                        ++k;
                        if (reqs[lv.req].type == LLG_REQ_RAMP) {
                            double B[3];
                            ramp_field(reqs[lv.req], lv.done, B);
                            SLLG_TRY(llg_batch_set_applied(b, lv.member, B));
                        }
                    }
            }
            uint32_t k = chunk_cap;
            for (auto& lv : live) {
                const llg_request& r = reqs[lv.req];
                k = std::min(k, r.update - lv.done % r.update);
                k = std::min(k, r.steps - lv.done);
            }
            SLLG_TRY(llg_batch_step(b, k));
            for (auto& lv : live) {
                lv.done += k;
                su += (uint64_t)reqs[lv.req].L * reqs[lv.req].L * reqs[lv.req].L * k;
            }
            // completions
            std::vector<int> fin;
            for (auto& lv : live)
                if (lv.done >= reqs[lv.req].steps) fin.push_back(lv.member);
            if (!fin.empty()) {
                std::vector<double> m(fin.size() * 3);
                SLLG_TRY(llg_batch_reduce(b, fin.data(), (uint32_t)fin.size(), 3, m.data()));
                const double tdone = since(t_start);
                size_t q = 0;
                std::vector<Live> keep;
                for (auto& lv : live) {
                    if (lv.done >= reqs[lv.req].steps) {
                        llg_result& res = results[lv.req];
                        for (int a = 0; a < 3; ++a) {
                            res.m_final[a] = m[3 * q + a];
                            res.m_mean[a] = lv.mcount ? lv.msum[a] / lv.mcount : m[3 * q + a];
                        }
                        res.samples = lv.samples;
                        res.done_s = tdone;
                        res.latency_s = tdone - res.arrival_s;
                        lat_hist.push_back(res.latency_s);
                        SLLG_TRY(llg_batch_remove(b, lv.member));
                        ++q;
                        ++completed;
                    } else {
                        keep.push_back(lv);
                    }
                }
                live.swap(keep);
                progress();
            }
        }
    }
    stats->wall_s = since(t_start);
    stats->completed = completed;
    stats->max_live = max_live;
    stats->spin_updates = su;
    stats->solves_per_s = completed / stats->wall_s;
    stats->spin_updates_per_s = (double)su / stats->wall_s;
    if (!lat_hist.empty()) {
        std::sort(lat_hist.begin(), lat_hist.end());
        auto pct = [&](double p) {
            const double x = p * (lat_hist.size() - 1);
            const size_t lo = (size_t)x;
            const size_t hi = std::min(lo + 1, lat_hist.size() - 1);
            return lat_hist[lo] + (lat_hist[hi] - lat_hist[lo]) * (x - lo);
        };
        stats->p50_latency_s = pct(0.5);
        stats->p99_latency_s = pct(0.99);
        double s = 0;
        for (double v : lat_hist) s += v;
        stats->mean_latency_s = s / lat_hist.size();
    }
    return 0;
}
