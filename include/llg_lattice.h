/*
 * llg_lattice.h - host-side construction of the inputs the time step consumes:
 * spin order, exchange CSR / periodic stencil / DIA, mesh coordinates and the
 * reference-compatible random initial spins.  Pure CPU code in libsllg.so; no GPU.
 *
 * Reference interfaces replaced (tomostler13/gpu-atomistic-code):
 *   llg_lattice_build_csr     <- exch::exchunitcell (src/exch_unitcell.cpp:265-491): per spin,
 *                                species-major then interaction order, neighbour cell =
 *                                cell + offset with periodic wrap, value J/(mu_i mu_B) (:399)
 *   llg_lattice_shells        <- the "unitcell" exchange file (src/exch_determine_exchange_matrix.cpp:27-144),
 *                                generated here from distance shells instead of read from disk
 *   llg_csr_to_dia            <- matconv::csr_to_dia_diag/offdiag (src/matrix_conv.cpp:28-272)
 *   llg_lattice_mesh          <- geom::lu(i,0..2) (src/geom.cpp:225-229)
 *   llg_random_spins          <- spins::setSpinsRandom (src/spins.cpp:265-281) with Random::rand
 *                                (src/random.cpp:17-40) seeded (Seed, Seed+100) (src/config.cpp:64)
 */
#ifndef LLG_LATTICE_H
#define LLG_LATTICE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Unit cell: lattice vectors rprim (rows, in units of the lattice constant),
 * fractional atom coordinates xred[n_atoms][3], and the dipolar mesh (nk points per
 * cell along each axis, mp[n_atoms][3] the mesh point of each atom; mp may be NULL). */
typedef struct llg_unitcell {
    uint32_t n_atoms;
    uint32_t nk[3];
    double rprim[9];
    const double* xred;
    const uint32_t* mp;
} llg_unitcell;

/* One exchange interaction: atom_i of a cell couples to atom_j of the cell at
 * offset dc (in unit cells) with tensor J (joules, row-major, J[a*3+b] multiplies S_b). */
typedef struct llg_interaction {
    uint32_t atom_i, atom_j;
    int32_t dc[3];
    int32_t shell;
    double J[9];
} llg_interaction;

/* Enumerate the first nshell neighbour shells (by distance) of every atom of the
 * cell, with isotropic J_shell[s] (joules) on shell s.  Order: by atom_i, then atom_j,
 * then shell, then dc lexicographic.  Two-pass: with out == NULL only *nout is set.
 * shell_dist (may be NULL) receives the nshell distances in lattice-constant units. */
int llg_lattice_shells(const llg_unitcell* uc, uint32_t nshell, const double* J_shell,
                       llg_interaction* out, uint32_t cap, uint32_t* nout, double* shell_dist);

/* CSR of the dims[0] x dims[1] x dims[2] lattice (spin idx = ((x*d1+y)*d2+z)*na+t).
 * vals is component-major [ncomp][nnz] in tesla: J/(mu_atom[atom_i] mu_B); ncomp 1 takes
 * J[0] (xx), 3 takes the diagonal, 9 the full tensor.  Two-pass: adjncy == NULL returns
 * only *nnz. */
int llg_lattice_build_csr(const llg_unitcell* uc, const uint32_t dims[3], const uint32_t pbc[3],
                          uint32_t n_int, const llg_interaction* ints, const double* mu_atom,
                          int ncomp, uint32_t* xadj, uint32_t* adjncy, double* vals, uint64_t* nnz);

/* Periodic stencil of the same interactions: count[na], entries[na][kmax][4] =
 * (dx, dy, dz, atom_j), vals[ncomp][na][kmax] (tesla).  *kmax_out is the largest count;
 * with entries == NULL only counts and *kmax_out are returned. */
int llg_lattice_build_stencil(const llg_unitcell* uc, uint32_t n_int, const llg_interaction* ints,
                              const double* mu_atom, int ncomp, uint32_t kmax, uint32_t* count,
                              int32_t* entries, double* vals, uint32_t* kmax_out);

/* Mesh coordinates mesh[n][3] = (x*nk0 + mp0, y*nk1 + mp1, z*nk2 + mp2). */
int llg_lattice_mesh(const llg_unitcell* uc, const uint32_t dims[3], uint32_t* mesh);

/* CSR -> Bell-Garland DIA: offsets ascending, vals[ncomp][ndiag][n], zero padding.
 * Two-pass: offsets == NULL returns only *ndiag. */
int llg_csr_to_dia(uint32_t n, const uint32_t* xadj, const uint32_t* adjncy, const double* vals,
                   int ncomp, uint64_t nnz, uint32_t* ndiag, int32_t* offsets, double* dia_vals);

/* Reference-compatible random unit spins: MWC1616 seeded (a, b), sphere point method. */
int llg_random_spins(uint32_t a, uint32_t b, uint32_t n, double* sx, double* sy, double* sz);

#ifdef __cplusplus
}
#endif
#endif
