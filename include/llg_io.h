/*
 * llg_io.h - the reference's input files read into the descriptor's building blocks, so
 * existing inputs drive the new time step.  Pure CPU code in libsllg.so; no GPU.
 *
 * Reference interfaces replaced (tomostler13/gpu-atomistic-code):
 *   llg_cfg_*             <- libconfig::Config as the reference uses it (src/config.cpp:30):
 *                            settings, groups, arrays, lists, strings, booleans, comments
 *   llg_read_ucf          <- the unit-cell file reader in geom::initGeom
 *                            (src/geom_glob.cpp:190-286; format unit_cell_files/format.unitcell)
 *   llg_read_exch_unitcell<- exch::get_exch_unitcell, the "unitcell" exchange file with
 *                            energy-unit conversion (src/exch_determine_exchange_matrix.cpp:27-144)
 *   llg_csr_read / _write <- csr_exch_mat.dat reading (src/exch.cpp:63-132) and writing
 *                            (src/exch_unitcell.cpp:35-70)
 *   llg_four_spin_quartets<- exch::read4spin + exch::setup4SpinCSR, the four-spin quartets of
 *                            the exchange file (src/exch_read_4_spin.cpp:29-258)
 *   llg_four_csr_read /   <- 4spin_csr_exch_mat.dat reading and writing
 *   _write                   (src/exch_read_4_spin.cpp:260-381)
 *   llg_*_spins_text,     <- restart_text / grid spin files and the ParaView spin map
 *   llg_*_grid,              (src/spins.cpp:342-468, src/simple.cpp:114-141, src/util.cpp:265-370)
 *   llg_write_vtu
 * Errors: negative llg_status, message in llg_last_error() (with the file's line number for
 * syntax errors); nothing exits.
 */
#ifndef LLG_IO_H
#define LLG_IO_H
#include <stdint.h>

#include "llg_lattice.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- libconfig-format files */

typedef struct llg_cfg llg_cfg;

enum {
    LLG_CFG_MISSING = 0,
    LLG_CFG_GROUP = 1,
    LLG_CFG_ARRAY = 2,
    LLG_CFG_LIST = 3,
    LLG_CFG_INT = 4,
    LLG_CFG_FLOAT = 5,
    LLG_CFG_BOOL = 6,
    LLG_CFG_STRING = 7
};

int llg_cfg_parse_file(const char* path, llg_cfg** out);
int llg_cfg_parse_string(const char* text, llg_cfg** out);
void llg_cfg_free(llg_cfg* cfg);
/* Paths are dot-separated setting names ("system.dim"); an element of an array or list is
 * addressed by the index argument of the getters.  Names are case-sensitive, as in libconfig. */
int llg_cfg_type(const llg_cfg* cfg, const char* path);
int llg_cfg_length(const llg_cfg* cfg, const char* path, uint32_t* n);
/* Numbers (int, float or bool) of a scalar (index 0) or of an array/list element. */
int llg_cfg_get_number(const llg_cfg* cfg, const char* path, uint32_t index, double* out);
int llg_cfg_get_bool(const llg_cfg* cfg, const char* path, int* out);
/* String value; *len receives its length (without the terminator); buf may be NULL. */
int llg_cfg_get_string(const llg_cfg* cfg, const char* path, char* buf, uint32_t cap, uint32_t* len);
/* Name of the index-th setting of a group (path "" is the top level), same buffer rules. */
int llg_cfg_child_name(const llg_cfg* cfg, const char* path, uint32_t index, char* buf, uint32_t cap, uint32_t* len);

/* ---------------------------------------------------------------- unit-cell file */

typedef struct llg_ucf_atom {
    uint32_t sublattice;
    uint32_t mp[3];        /* mesh point within the cell (only with with_mesh) */
    double xred[3];        /* fractional position */
    double mu;             /* Bohr magnetons */
    double lambda;
    double gamma_rel;
    double spin0[3];       /* initial spin */
    double k1u;            /* first-order uniaxial constant (J) */
    double k1u_dir[3];
    char element[16];
} llg_ucf_atom;

/* Reads "nms nauc" then nauc atom records: sublattice, xred[3], [mp[3] if with_mesh, i.e.
 * the config sets system.Nm], mu, lambda, gamma, element, spin0[3], k1u, k1u_dir[3].
 * Checks the reference's bounds (1e-5 < mu, gamma < 1e5).  Two-pass: atoms == NULL
 * returns only the counts. */
int llg_read_ucf(const char* path, int with_mesh, uint32_t* n_species, uint32_t* n_atoms, llg_ucf_atom* atoms,
                 uint32_t cap);

/* ---------------------------------------------------------------- exchange file */

/* "unitcell" exchange file: groups exchange_<a>_<b> for atoms a, b < n_atoms with
 * Num_Interactions, units ("J"/"Joules"/"joules", "mRy", "Ry", "eV", "meV"; missing = J),
 * UnitCell<k> (cell offset, rounded half away from zero) and rows J<k>_1..3 of the
 * tensor.  Output order: atom_i, atom_j, k - the reference's CSR row order.  J is in
 * joules (row-major, J[a*3+b] multiplies S_b); shell = k.  Two-pass: out == NULL returns
 * only *nout. */
int llg_read_exch_unitcell(const llg_cfg* exch, uint32_t n_atoms, llg_interaction* out, uint32_t cap,
                           uint32_t* nout);

/* ---------------------------------------------------------------- CSR matrix file */

/* csr_exch_mat.dat: "n\nnnz\n" xadj[0..n] then per nnz "adjncy xx yy zz" (tesla).
 * Fails unless the file's n equals n.  Two-pass: adjncy == NULL returns only *nnz.
 * vals3 is [3][nnz] (xx, yy, zz). */
int llg_csr_read(const char* path, uint32_t n, uint32_t* xadj, uint32_t* adjncy, double* vals3, uint64_t* nnz);
int llg_csr_write(const char* path, uint32_t n, const uint32_t* xadj, const uint32_t* adjncy, const double* vals3,
                  uint64_t nnz);

/* ---------------------------------------------------------------- four-spin quartets */

/* Rounding of the quartet offsets to mesh points: nearest (half away from zero), or the
 * reference's static_cast<int>(x + 0.5) (src/exch_read_4_spin.cpp:114), which truncates
 * toward zero, so a negative offset comes out one mesh point too large (-1 -> 0, -2 -> -1). */
enum { LLG_FOUR_ROUND_NEAREST = 0, LLG_FOUR_ROUND_REFERENCE = 1 };

/* Four-spin quartets of every spin from the exchange file (exch::read4spin and
 * exch::setup4SpinCSR, src/exch_read_4_spin.cpp:29-258).  The file holds, per species
 * s < nms, a group FourSpinExchange_<s> with NumPermutations (at most
 * exchange.MaxQuartetPermutations), JQ (joules), InterfaceSpecies and Permute<q>_<m>
 * (q = 1..NumPermutations, m = 1, 2, 3: the offsets of spins j, k, l from spin i in unit
 * cells).  An offset becomes o * nk[c] * Scale[c] mesh points (exchange.Scale; 1 when the
 * file has none: the reference reads it only for its "permute" and "mapint" methods),
 * rounded as `rounding` says.  Spin i (mesh coordinates mesh[i][0..2] = lu(i, 0..2),
 * species sublattice[i]) gets quartet q of its species when, after one periodic wrap along
 * each axis with pbc set (grid[c] mesh points per axis), all three mesh points lie on the
 * grid and hold spins whose species is s or InterfaceSpecies; quartets are listed per spin
 * in permutation order.  xadj [n+1] and jkl [cap][3] receive the quartet CSR (jkl == NULL:
 * only *nq is returned); jq_species[nms] (may be NULL) receives JQ / (mu_s mu_B) in tesla
 * (mu_species in Bohr magnetons; the GPU term takes species 0's value, as the reference's
 * kernel does).  Two spins on one mesh point are an error. */
int llg_four_spin_quartets(const llg_cfg* exch, uint32_t nms, const double* mu_species, uint32_t n,
                           const uint32_t* mesh, const uint32_t* sublattice, const uint32_t grid[3],
                           const uint32_t nk[3], const uint32_t pbc[3], int rounding, uint32_t* xadj, uint32_t* jkl,
                           uint64_t cap, uint64_t* nq, double* jq_species);

/* 4spin_csr_exch_mat.dat: "n\nnq\n" xadj[0..n] then per quartet "j k l".  Fails unless the
 * file's n equals n.  Two-pass: jkl == NULL returns only *nq. */
int llg_four_csr_read(const char* path, uint32_t n, uint32_t* xadj, uint32_t* jkl, uint64_t* nq);
int llg_four_csr_write(const char* path, uint32_t n, const uint32_t* xadj, const uint32_t* jkl, uint64_t nq);

/* ---------------------------------------------------------------- spin configuration files */

/* restart_text: "n" then n lines "sx sy sz"; spins are normalised on reading and a spin
 * shorter than 1e-12 is an error (spins::setSpinsResText, src/spins.cpp:342-395). */
int llg_read_spins_text(const char* path, uint32_t n, double* sx, double* sy, double* sz);
int llg_write_spins_text(const char* path, uint32_t n, const double* sx, const double* sy, const double* sz);

/* grid (dw_grid.out / SpinConfigMethod "grid"): "n" then n lines "x y z sx sy sz" with the
 * mesh coordinates lu(i,0..2) of the spin (src/simple.cpp:114-141 writes it;
 * spins::setSpinsGrid, src/spins.cpp:397-468, reads it in any line order).  mesh is the
 * [n][3] coordinate table of the system; a line naming no spin is an error.  The reading
 * normalises the spin it read (the reference normalises spin i of the line counter
 * instead, src/spins.cpp:451-453). */
int llg_read_grid(const char* path, uint32_t n, const uint32_t* mesh, double* sx, double* sy, double* sz);
int llg_write_grid(const char* path, uint32_t n, const uint32_t* mesh, const double* sx, const double* sy,
                   const double* sz);

/* ParaView spin map (util::outputSpinsVTU, src/util.cpp:265-370): VTK UnstructuredGrid,
 * point data "Spin" (3 components) at the spin positions pos[n][3], one placeholder cell. */
int llg_write_vtu(const char* path, uint32_t n, const double* pos, const double* sx, const double* sy,
                  const double* sz);

#ifdef __cplusplus
}
#endif
#endif
