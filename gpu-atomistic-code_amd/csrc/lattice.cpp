// lattice.cpp - host-side builders declared in include/llg_lattice.h (no GPU code).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/llg_capi.h"
#include "../../include/llg_lattice.h"

namespace {

const double kTolShell = 1e-6;  // shells are distances equal within this (lattice-constant units)

double cart_dist(const llg_unitcell* uc, uint32_t ai, uint32_t aj, const int32_t dc[3]) {
    double fr[3];
    for (int a = 0; a < 3; ++a) fr[a] = uc->xred[aj * 3 + a] + dc[a] - uc->xred[ai * 3 + a];
    double r[3] = {0, 0, 0};
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) r[b] += fr[a] * uc->rprim[a * 3 + b];
    return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

bool uc_ok(const llg_unitcell* uc) { return uc && uc->n_atoms > 0 && uc->xred; }

// Interactions of atom t in CSR row order: neighbour atom (species) major, then the
// order of the interaction list (src/exch_unitcell.cpp:299-301).
std::vector<const llg_interaction*> row_order(uint32_t na, uint32_t t, uint32_t n_int, const llg_interaction* ints) {
    std::vector<const llg_interaction*> v;
    for (uint32_t s1 = 0; s1 < na; ++s1)
        for (uint32_t k = 0; k < n_int; ++k)
            if (ints[k].atom_i == t && ints[k].atom_j == s1) v.push_back(&ints[k]);
    return v;
}

}  // namespace

extern "C" {

int llg_lattice_shells(const llg_unitcell* uc, uint32_t nshell, const double* J_shell, llg_interaction* out,
                       uint32_t cap, uint32_t* nout, double* shell_dist) {
    if (!uc_ok(uc) || !nout || nshell == 0 || !J_shell) return LLG_E_INVALID;
    const uint32_t na = uc->n_atoms;
    const int R = (int)nshell + 2;
    // distinct distances over all atom pairs and offsets
    std::vector<double> d;
    for (uint32_t ai = 0; ai < na; ++ai)
        for (uint32_t aj = 0; aj < na; ++aj)
            for (int x = -R; x <= R; ++x)
                for (int y = -R; y <= R; ++y)
                    for (int z = -R; z <= R; ++z) {
                        const int32_t dc[3] = {x, y, z};
                        const double r = cart_dist(uc, ai, aj, dc);
                        if (r > kTolShell) d.push_back(r);
                    }
    std::sort(d.begin(), d.end());
    std::vector<double> shells;
    for (double r : d) {
        if (shells.empty() || r - shells.back() > kTolShell) shells.push_back(r);
        if (shells.size() == nshell) break;
    }
    if (shells.size() < nshell) return LLG_E_INVALID;
    if (shell_dist)
        for (uint32_t s = 0; s < nshell; ++s) shell_dist[s] = shells[s];
    // the search box must contain every member of the last shell
    const double rmax = shells.back() + kTolShell;
    uint32_t cnt = 0;
    for (uint32_t ai = 0; ai < na; ++ai)
        for (uint32_t aj = 0; aj < na; ++aj)
            for (uint32_t s = 0; s < nshell; ++s)
                for (int x = -R; x <= R; ++x)
                    for (int y = -R; y <= R; ++y)
                        for (int z = -R; z <= R; ++z) {
                            const int32_t dc[3] = {x, y, z};
                            const double r = cart_dist(uc, ai, aj, dc);
                            if (r > rmax || std::fabs(r - shells[s]) > kTolShell) continue;
                            if (out) {
                                if (cnt >= cap) return LLG_E_INVALID;
                                llg_interaction& it = out[cnt];
                                std::memset(&it, 0, sizeof it);
                                it.atom_i = ai;
                                it.atom_j = aj;
                                it.dc[0] = x; it.dc[1] = y; it.dc[2] = z;
                                it.shell = (int32_t)s;
                                it.J[0] = it.J[4] = it.J[8] = J_shell[s];
                            }
                            ++cnt;
                        }
    *nout = cnt;
    return 0;
}

// CSR in the order of exch::exchunitcell (src/exch_unitcell.cpp:292-420).
int llg_lattice_build_csr(const llg_unitcell* uc, const uint32_t dims[3], const uint32_t pbc[3], uint32_t n_int,
                          const llg_interaction* ints, const double* mu_atom, int ncomp, uint32_t* xadj,
                          uint32_t* adjncy, double* vals, uint64_t* nnz) {
    if (!uc_ok(uc) || !dims || !pbc || !nnz || !mu_atom || (n_int && !ints)) return LLG_E_INVALID;
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) return LLG_E_INVALID;
    if (adjncy && (!xadj || !vals)) return LLG_E_INVALID;
    const uint32_t na = uc->n_atoms;
    for (uint32_t k = 0; k < n_int; ++k)
        if (ints[k].atom_i >= na || ints[k].atom_j >= na) return LLG_E_INVALID;
    std::vector<std::vector<const llg_interaction*>> rows(na);
    for (uint32_t t = 0; t < na; ++t) rows[t] = row_order(na, t, n_int, ints);
    const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2] * na;
    const uint64_t cap = *nnz;
    uint64_t q = 0;
    uint64_t idx = 0;
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t z = 0; z < dims[2]; ++z)
                for (uint32_t t = 0; t < na; ++t, ++idx) {
                    if (adjncy) xadj[idx] = (uint32_t)q;
                    const int64_t cell[3] = {x, y, z};
                    const double den = mu_atom[t] * LLG_MU_B;
                    for (const llg_interaction* it : rows[t]) {
                        int64_t p[3];
                        bool ok = true;
                        for (int a = 0; a < 3 && ok; ++a) {
                            p[a] = cell[a] + it->dc[a];
                            const int64_t D = dims[a];
                            if (p[a] < 0 || p[a] >= D) {
                                if (!pbc[a]) ok = false;
                                else p[a] = ((p[a] % D) + D) % D;
                            }
                        }
                        if (!ok) continue;
                        if (adjncy) {
                            if (q >= cap) return LLG_E_INVALID;
                            adjncy[q] = (uint32_t)((((uint64_t)p[0] * dims[1] + p[1]) * dims[2] + p[2]) * na + it->atom_j);
                            if (ncomp == 1) {
                                vals[q] = it->J[0] / den;
                            } else if (ncomp == 3) {
                                for (int a = 0; a < 3; ++a) vals[(size_t)a * cap + q] = it->J[a * 4] / den;
                            } else {
                                for (int a = 0; a < 9; ++a) vals[(size_t)a * cap + q] = it->J[a] / den;
                            }
                        }
                        ++q;
                    }
                }
    if (adjncy) {
        xadj[n] = (uint32_t)q;
        if (q != cap) return LLG_E_INVALID;
    }
    *nnz = q;
    return 0;
}

int llg_lattice_build_stencil(const llg_unitcell* uc, uint32_t n_int, const llg_interaction* ints,
                              const double* mu_atom, int ncomp, uint32_t kmax, uint32_t* count, int32_t* entries,
                              double* vals, uint32_t* kmax_out) {
    if (!uc_ok(uc) || !count || !kmax_out || !mu_atom || (n_int && !ints)) return LLG_E_INVALID;
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) return LLG_E_INVALID;
    const uint32_t na = uc->n_atoms;
    uint32_t km = 0;
    for (uint32_t t = 0; t < na; ++t) {
        const std::vector<const llg_interaction*> r = row_order(na, t, n_int, ints);
        count[t] = (uint32_t)r.size();
        km = std::max<uint32_t>(km, count[t]);
        if (!entries) continue;
        if (count[t] > kmax || !vals) return LLG_E_INVALID;
        const double den = mu_atom[t] * LLG_MU_B;
        for (uint32_t q = 0; q < count[t]; ++q) {
            int32_t* e = &entries[((size_t)t * kmax + q) * 4];
            e[0] = r[q]->dc[0]; e[1] = r[q]->dc[1]; e[2] = r[q]->dc[2]; e[3] = (int32_t)r[q]->atom_j;
            for (int a = 0; a < ncomp; ++a) {
                const double J = ncomp == 1 ? r[q]->J[0] : (ncomp == 3 ? r[q]->J[a * 4] : r[q]->J[a]);
                vals[((size_t)a * na + t) * kmax + q] = J / den;
            }
        }
    }
    *kmax_out = km;
    return 0;
}

int llg_lattice_mesh(const llg_unitcell* uc, const uint32_t dims[3], uint32_t* mesh) {
    if (!uc_ok(uc) || !dims || !mesh) return LLG_E_INVALID;
    const uint32_t na = uc->n_atoms;
    uint64_t i = 0;
    for (uint32_t x = 0; x < dims[0]; ++x)
        for (uint32_t y = 0; y < dims[1]; ++y)
            for (uint32_t z = 0; z < dims[2]; ++z)
                for (uint32_t t = 0; t < na; ++t, ++i) {
                    const uint32_t c[3] = {x, y, z};
                    for (int a = 0; a < 3; ++a) {
                        const uint32_t nk = uc->nk[a] ? uc->nk[a] : 1;
                        const uint32_t mp = uc->mp ? uc->mp[t * 3 + a] : 0;
                        mesh[i * 3 + a] = c[a] * nk + mp;
                    }
                }
    return 0;
}

int llg_csr_to_dia(uint32_t n, const uint32_t* xadj, const uint32_t* adjncy, const double* vals, int ncomp,
                   uint64_t nnz, uint32_t* ndiag, int32_t* offsets, double* dia_vals) {
    if (!xadj || (nnz && (!adjncy || !vals)) || !ndiag) return LLG_E_INVALID;
    if (ncomp != 1 && ncomp != 3 && ncomp != 9) return LLG_E_INVALID;
    std::vector<int32_t> offs;
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t q = xadj[r]; q < xadj[r + 1]; ++q) offs.push_back((int32_t)((int64_t)adjncy[q] - r));
    std::sort(offs.begin(), offs.end());
    offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
    if (!offsets) {
        *ndiag = (uint32_t)offs.size();
        return 0;
    }
    if (*ndiag < offs.size() || !dia_vals) return LLG_E_INVALID;
    const uint32_t nd = (uint32_t)offs.size();
    std::memcpy(offsets, offs.data(), nd * sizeof(int32_t));
    std::memset(dia_vals, 0, sizeof(double) * (size_t)ncomp * nd * n);
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t q = xadj[r]; q < xadj[r + 1]; ++q) {
            const int32_t o = (int32_t)((int64_t)adjncy[q] - r);
            const uint32_t d = (uint32_t)(std::lower_bound(offs.begin(), offs.end(), o) - offs.begin());
            for (int a = 0; a < ncomp; ++a) dia_vals[((size_t)a * nd + d) * n + r] += vals[(size_t)a * nnz + q];
        }
    *ndiag = nd;
    return 0;
}

// MWC1616 (src/random.cpp:17-23), uniform x/(2^32-1) (:37-40), sphere point (src/spins.cpp:265-281).
int llg_random_spins(uint32_t a, uint32_t b, uint32_t n, double* sx, double* sy, double* sz) {
    if (!sx || !sy || !sz || a == 0 || b == 0) return LLG_E_INVALID;
    uint32_t x = a, y = b;
    auto next = [&]() {
        x = 18000u * (x & 0xFFFFu) + (x >> 16);
        y = 30903u * (y & 0xFFFFu) + (y >> 16);
        return (x << 16) + (y & 0xFFFFu);
    };
    const double rmax = 1.0 / 4294967295.0;
    for (uint32_t i = 0; i < n; ++i) {
        double v1 = 0, v2 = 0, s = 2.0;
        while (s > 1.0) {
            v1 = 2.0 * (next() * rmax) - 1.0;
            v2 = 2.0 * (next() * rmax) - 1.0;
            s = v1 * v1 + v2 * v2;
        }
        const double ss = std::sqrt(1.0 - s);
        sx[i] = 2.0 * v1 * ss;
        sy[i] = 2.0 * v2 * ss;
        sz[i] = 1.0 - 2.0 * s;
    }
    return 0;
}

}  // extern "C"
