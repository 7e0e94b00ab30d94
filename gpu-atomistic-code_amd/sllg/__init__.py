"""Python binding of libsllg.so (ctypes over the C ABI in include/llg_capi.h).

The product is the C ABI and its HIP kernels; this module only marshals numpy arrays
into the descriptor, for the tests and the benchmark.  There is no fallback: if
libsllg.so is missing, importing the compute entry points raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsllg.so")

ABI_VERSION = 1
MAX_SPECIES = 64
GAMMA_E = 1.76e11
MU_B = 9.27e-24
K_B = 1.38e-23

HEUN, RK4, RK5 = 0, 1, 2
FP64_PLAIN, FP32 = 0, 1
EXCH_NONE, EXCH_DIA, EXCH_CSR, EXCH_STENCIL = 0, 1, 2, 3
NOISE_COUNTER, NOISE_INJECTED = 0, 1
FLAG_CHECK_FINITE = 0x1
FLAG_NO_XCD_REMAP = 0x100
FLAG_NO_SUBLATTICE_ORDER = 0x200
FLAG_NO_SLICE_SHARING = 0x400
FLAG_NO_LINE_MODE = 0x800
FLAG_DIPOLAR_FULL_FFT = 0x1000
FLAG_TUNE_FIXED_READER = 0x2000

STATUS = {0: "OK", -1: "INVALID", -2: "NOMEM", -3: "DEVICE", -4: "UNSUPPORTED", -5: "NODEVICE",
          -6: "STATE", -7: "NONFINITE"}

_SCHEMES = {"heun": HEUN, "rk4": RK4, "rk5": RK5}
_PRECS = {"fp64": FP64_PLAIN, "fp64_plain": FP64_PLAIN, "fp32": FP32}
_NOISE = {"counter": NOISE_COUNTER, "injected": NOISE_INJECTED}


class Species(C.Structure):
    _fields_ = [("mu", C.c_double), ("lambda_", C.c_double), ("gamma_rel", C.c_double), ("k1u", C.c_double),
                ("k1u_dir", C.c_double * 3), ("sublattice", C.c_uint32), ("pad_", C.c_uint32)]


class Anisotropy(C.Structure):
    _fields_ = [("k2perp", C.c_double), ("k2par", C.c_double), ("k4perp", C.c_double), ("k4par", C.c_double),
                ("k2perp_dir", C.c_double * 3), ("k2par_dir", C.c_double * 3),
                ("k4perp_dirs", C.c_double * 9), ("k4par_dirs", C.c_double * 9), ("kc", C.c_double)]


class Exchange(C.Structure):
    _fields_ = [("format", C.c_int32), ("ncomp", C.c_int32), ("nnz", C.c_uint64),
                ("xadj", C.POINTER(C.c_uint32)), ("adjncy", C.POINTER(C.c_uint32)),
                ("csr_vals", C.POINTER(C.c_double)),
                ("ndiag", C.c_uint32), ("dia_offsets", C.POINTER(C.c_int32)), ("dia_vals", C.POINTER(C.c_double)),
                ("dims", C.c_uint32 * 3), ("n_atoms", C.c_uint32), ("pbc", C.c_uint32 * 3),
                ("stencil_k", C.c_uint32), ("stencil_count", C.POINTER(C.c_uint32)),
                ("stencil_entries", C.POINTER(C.c_int32)), ("stencil_vals", C.POINTER(C.c_double)),
                ("stencil_vals_per_row", C.c_uint32), ("pad_", C.c_uint32)]


class Dipolar(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("every_stage", C.c_int32), ("refresh", C.c_uint32), ("pad_", C.c_uint32),
                ("dim", C.c_uint32 * 3), ("nk", C.c_uint32 * 3), ("abc", C.c_double * 3),
                ("mesh", C.POINTER(C.c_uint32))]


class FftInteraction(C.Structure):
    _fields_ = [("s1", C.c_uint32), ("s2", C.c_uint32), ("d", C.c_int32 * 3), ("pad_", C.c_uint32),
                ("J", C.c_double * 9)]


class FftExch(C.Structure):
    _fields_ = [("n_int", C.c_uint32), ("pad_", C.c_uint32), ("ints", C.POINTER(FftInteraction))]


class FourSpin(C.Structure):
    _fields_ = [("nq", C.c_uint64), ("xadj", C.POINTER(C.c_uint32)), ("jkl", C.POINTER(C.c_uint32)),
                ("jq", C.c_double), ("cpu_form", C.c_int32), ("pad_", C.c_uint32)]


class Desc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("n", C.c_uint32), ("n_species", C.c_uint32),
                ("n_sublattice", C.c_uint32), ("species", C.POINTER(Species)),
                ("spin_species", C.POINTER(C.c_uint32)), ("anis", Anisotropy), ("exch", Exchange),
                ("dip", Dipolar), ("dt", C.c_double), ("scheme", C.c_int32), ("precision", C.c_int32),
                ("noise", C.c_int32), ("update", C.c_uint32), ("seed", C.c_uint64), ("stream", C.c_uint64),
                ("flags", C.c_uint32), ("device", C.c_int32), ("four", FourSpin), ("fftx", FftExch)]


class Info(C.Structure):
    _fields_ = [("device_bytes", C.c_uint64), ("bytes_per_spin_update", C.c_double),
                ("dominant_kernel_bytes_per_spin", C.c_double), ("dominant_kernel", C.c_char * 48),
                ("exch_encoding", C.c_char * 64)]


class TuneResult(C.Structure):
    _fields_ = [("tpb", C.c_int), ("xcd", C.c_int), ("ms_per_step", C.c_double),
                ("ms_per_step_default", C.c_double), ("candidates", C.c_int), ("from_cache", C.c_int),
                ("unroll", C.c_int), ("thermal_cache", C.c_int), ("gather_copy", C.c_int), ("line_mode", C.c_int)]


class UnitCell(C.Structure):
    _fields_ = [("n_atoms", C.c_uint32), ("nk", C.c_uint32 * 3), ("rprim", C.c_double * 9),
                ("xred", C.POINTER(C.c_double)), ("mp", C.POINTER(C.c_uint32))]


class Interaction(C.Structure):
    _fields_ = [("atom_i", C.c_uint32), ("atom_j", C.c_uint32), ("dc", C.c_int32 * 3), ("shell", C.c_int32),
                ("J", C.c_double * 9)]


class LLGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"llg error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


_lib = None


def lib():
    """Load libsllg.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.llg_abi_version.restype = C.c_uint32
        L.llg_last_error.restype = C.c_char_p
        L.llg_device_count.argtypes = [C.POINTER(C.c_int)]
        L.llg_desc_init.argtypes = [C.POINTER(Desc)]
        L.llg_desc_init.restype = None
        L.llg_create.argtypes = [C.POINTER(Desc), C.POINTER(P)]
        L.llg_destroy.argtypes = [P]
        L.llg_destroy.restype = None
        for f in ("llg_set_spins", "llg_get_spins", "llg_set_staggered", "llg_get_det_field", "llg_get_field"):
            getattr(L, f).argtypes = [P, dp, dp, dp]
        L.llg_set_temperature.argtypes = [P, C.c_double]
        L.llg_set_applied.argtypes = [P, dp]
        L.llg_scale_staggered.argtypes = [P, C.c_double]
        L.llg_push_noise.argtypes = [P, C.c_uint32, dp]
        L.llg_step.argtypes = [P, C.c_uint64, C.c_uint32]
        L.llg_sync.argtypes = [P]
        L.llg_reduce.argtypes = [P, dp, dp]
        L.llg_reduce_mag_noalloc.argtypes = [P, dp]
        L.llg_run_sampled.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, dp]
        L.llg_noise_sample.argtypes = [P, C.c_uint64, dp]
        L.llg_stream.argtypes = [P]
        L.llg_stream.restype = P
        L.llg_profile_kernels.argtypes = [P, C.c_uint64, C.c_uint32, C.c_int, C.c_char_p, dp, C.POINTER(C.c_int)]
        L.llg_get_info.argtypes = [P, C.POINTER(Info)]
        L.llg_autotune.argtypes = [P, C.c_uint32, C.POINTER(TuneResult)]
        L.llg_set_bins.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint32), dp]
        L.llg_reduce_bins.argtypes = [P, dp]
        L.llg_set_launch.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        up = C.POINTER(C.c_uint32)
        L.llg_set_sq.argtypes = [P, up, up, C.c_uint32, up, C.c_int, dp, dp, C.c_uint32, up]
        L.llg_sample_sq.argtypes = [P, dp]
        L.llg_lattice_shells.argtypes = [C.POINTER(UnitCell), C.c_uint32, dp, C.POINTER(Interaction), C.c_uint32,
                                         up, dp]
        L.llg_lattice_build_csr.argtypes = [C.POINTER(UnitCell), up, up, C.c_uint32, C.POINTER(Interaction), dp,
                                            C.c_int, up, up, dp, C.POINTER(C.c_uint64)]
        L.llg_lattice_build_stencil.argtypes = [C.POINTER(UnitCell), C.c_uint32, C.POINTER(Interaction), dp, C.c_int,
                                                C.c_uint32, up, C.POINTER(C.c_int32), dp, up]
        L.llg_lattice_mesh.argtypes = [C.POINTER(UnitCell), up, up]
        L.llg_csr_to_dia.argtypes = [C.c_uint32, up, up, dp, C.c_int, C.c_uint64, up, C.POINTER(C.c_int32), dp]
        L.llg_random_spins.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, dp, dp, dp]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def _check(rc: int, L=None):
    if rc != 0:
        L = L or lib()
        raise LLGError(rc, L.llg_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().llg_device_count(C.byref(n)))
    return n.value


# ----------------------------------------------------------------------------- lattices

UNIT_CELLS = {
    "sc": dict(xred=[[0, 0, 0]], mp=[[0, 0, 0]], nk=(1, 1, 1)),
    "bcc": dict(xred=[[0, 0, 0], [.5, .5, .5]], mp=[[0, 0, 0], [1, 1, 1]], nk=(2, 2, 2)),
    "fcc": dict(xred=[[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]],
                mp=[[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]], nk=(2, 2, 2)),
    "bct": dict(xred=[[0, 0, 0], [.5, .5, .5]], mp=[[0, 0, 0], [1, 1, 1]], nk=(2, 2, 2)),
}


@dataclass
class UnitCellSpec:
    xred: np.ndarray
    rprim: np.ndarray
    nk: tuple
    mp: np.ndarray

    @property
    def n_atoms(self):
        return len(self.xred)

    def c(self):
        self._xred = np.ascontiguousarray(self.xred, dtype=np.float64).reshape(-1)
        self._mp = np.ascontiguousarray(self.mp, dtype=np.uint32).reshape(-1)
        uc = UnitCell()
        uc.n_atoms = self.n_atoms
        uc.nk[:] = list(self.nk)
        uc.rprim[:] = list(np.asarray(self.rprim, dtype=np.float64).reshape(-1))
        uc.xred = _p(self._xred, C.c_double)
        uc.mp = _p(self._mp, C.c_uint32)
        return uc


def unit_cell(kind: str, c_over_a: float = 1.0) -> UnitCellSpec:
    u = UNIT_CELLS[kind]
    rprim = np.eye(3)
    if kind == "bct":
        rprim[2, 2] = c_over_a
    return UnitCellSpec(np.array(u["xred"], float), rprim, tuple(u["nk"]), np.array(u["mp"], np.uint32))


def shells(uc: UnitCellSpec, J_shell: Sequence[float]):
    """Interaction list of the first len(J_shell) distance shells (isotropic J, joules)."""
    L = lib()
    ucc = uc.c()
    Js = np.ascontiguousarray(J_shell, dtype=np.float64)
    cnt = C.c_uint32(0)
    _check(L.llg_lattice_shells(C.byref(ucc), len(Js), _p(Js, C.c_double), None, 0, C.byref(cnt), None))
    arr = (Interaction * cnt.value)()
    dist = np.zeros(len(Js))
    _check(L.llg_lattice_shells(C.byref(ucc), len(Js), _p(Js, C.c_double), arr, cnt.value, C.byref(cnt),
                                _p(dist, C.c_double)))
    return arr, dist


def interactions_from(list_of_dicts):
    """Interaction list from dicts {atom_i, atom_j, dc, J (3x3 or scalar, joules)}."""
    arr = (Interaction * len(list_of_dicts))()
    for k, d in enumerate(list_of_dicts):
        arr[k].atom_i = d["atom_i"]
        arr[k].atom_j = d["atom_j"]
        arr[k].dc[:] = list(d["dc"])
        arr[k].shell = d.get("shell", 0)
        J = np.asarray(d["J"], dtype=np.float64)
        if J.ndim == 0:
            J = np.eye(3) * float(J)
        arr[k].J[:] = list(J.reshape(-1))
    return arr


def build_csr(uc: UnitCellSpec, dims, ints, mu_atom, ncomp=1, pbc=(1, 1, 1)):
    L = lib()
    ucc = uc.c()
    dims_a = np.ascontiguousarray(dims, dtype=np.uint32)
    pbc_a = np.ascontiguousarray(pbc, dtype=np.uint32)
    mu = np.ascontiguousarray(mu_atom, dtype=np.float64)
    nnz = C.c_uint64(0)
    _check(L.llg_lattice_build_csr(C.byref(ucc), _p(dims_a, C.c_uint32), _p(pbc_a, C.c_uint32), len(ints), ints,
                                   _p(mu, C.c_double), ncomp, None, None, None, C.byref(nnz)))
    n = int(np.prod(dims_a)) * uc.n_atoms
    xadj = np.zeros(n + 1, np.uint32)
    adj = np.zeros(nnz.value, np.uint32)
    vals = np.zeros(ncomp * nnz.value, np.float64)
    _check(L.llg_lattice_build_csr(C.byref(ucc), _p(dims_a, C.c_uint32), _p(pbc_a, C.c_uint32), len(ints), ints,
                                   _p(mu, C.c_double), ncomp, _p(xadj, C.c_uint32), _p(adj, C.c_uint32),
                                   _p(vals, C.c_double), C.byref(nnz)))
    return xadj, adj, vals.reshape(ncomp, -1)


def build_stencil(uc: UnitCellSpec, ints, mu_atom, ncomp=1):
    L = lib()
    ucc = uc.c()
    mu = np.ascontiguousarray(mu_atom, dtype=np.float64)
    na = uc.n_atoms
    count = np.zeros(na, np.uint32)
    km = C.c_uint32(0)
    _check(L.llg_lattice_build_stencil(C.byref(ucc), len(ints), ints, _p(mu, C.c_double), ncomp, 0,
                                       _p(count, C.c_uint32), None, None, C.byref(km)))
    k = km.value
    ent = np.zeros((na, k, 4), np.int32)
    vals = np.zeros((ncomp, na, k), np.float64)
    _check(L.llg_lattice_build_stencil(C.byref(ucc), len(ints), ints, _p(mu, C.c_double), ncomp, k,
                                       _p(count, C.c_uint32), _p(ent, C.c_int32), _p(vals, C.c_double),
                                       C.byref(km)))
    return count, ent, vals


def mesh(uc: UnitCellSpec, dims):
    L = lib()
    ucc = uc.c()
    dims_a = np.ascontiguousarray(dims, dtype=np.uint32)
    n = int(np.prod(dims_a)) * uc.n_atoms
    m = np.zeros((n, 3), np.uint32)
    _check(L.llg_lattice_mesh(C.byref(ucc), _p(dims_a, C.c_uint32), _p(m, C.c_uint32)))
    return m


def csr_to_dia(n, xadj, adj, vals):
    L = lib()
    ncomp = vals.shape[0]
    v = np.ascontiguousarray(vals, dtype=np.float64)
    nd = C.c_uint32(0)
    _check(L.llg_csr_to_dia(n, _p(xadj, C.c_uint32), _p(adj, C.c_uint32), _p(v, C.c_double), ncomp, len(adj),
                            C.byref(nd), None, None))
    offs = np.zeros(nd.value, np.int32)
    dv = np.zeros((ncomp, nd.value, n), np.float64)
    _check(L.llg_csr_to_dia(n, _p(xadj, C.c_uint32), _p(adj, C.c_uint32), _p(v, C.c_double), ncomp, len(adj),
                            C.byref(nd), _p(offs, C.c_int32), _p(dv, C.c_double)))
    return offs, dv


def random_spins(n, a=42, b=142):
    """Reference-compatible random initial spins (MWC1616 seeded (a, b))."""
    s = np.zeros((3, n))
    _check(lib().llg_random_spins(a, b, n, _p(s[0], C.c_double), _p(s[1], C.c_double), _p(s[2], C.c_double)))
    return s


# ----------------------------------------------------------------------------- system descriptor

@dataclass
class System:
    """Everything llg_create needs, as numpy arrays (kept alive while a Desc points at them)."""
    n: int
    species: list  # list of dicts: mu, lambda_, gamma_rel, k1u, k1u_dir, sublattice
    spin_species: Optional[np.ndarray] = None
    n_sublattice: int = 1
    anis: dict = field(default_factory=dict)
    exch: dict = field(default_factory=lambda: {"format": EXCH_NONE})
    dip: dict = field(default_factory=lambda: {"enabled": 0})
    dt: float = 1e-16
    scheme: str = "heun"
    precision: str = "fp32"
    noise: str = "counter"
    update: int = 1
    seed: int = 0
    stream: int = 0
    flags: int = 0
    device: int = 0
    four: dict = field(default_factory=dict)  # xadj [n+1], jkl [nq,3], jq (T), cpu_form
    # FFT exchange: list of (s1, s2, (dx, dy, dz), J 3x3 in joules); needs the mesh geometry
    # in dip (dim, nk, abc, mesh; dip["enabled"] adds the dipolar terms to the same kernel)
    fftx: list = field(default_factory=list)

    def desc(self) -> Desc:
        d = Desc()
        lib().llg_desc_init(C.byref(d))
        d.n = self.n
        self._sp = (Species * len(self.species))()
        for k, s in enumerate(self.species):
            self._sp[k].mu = s.get("mu", 2.22)
            self._sp[k].lambda_ = s.get("lambda_", 0.1)
            self._sp[k].gamma_rel = s.get("gamma_rel", 1.0)
            self._sp[k].k1u = s.get("k1u", 0.0)
            self._sp[k].k1u_dir[:] = list(s.get("k1u_dir", (0.0, 0.0, 1.0)))
            self._sp[k].sublattice = s.get("sublattice", 0)
        d.n_species = len(self.species)
        d.species = C.cast(self._sp, C.POINTER(Species))
        d.n_sublattice = self.n_sublattice
        if self.spin_species is not None:
            self._ss = np.ascontiguousarray(self.spin_species, dtype=np.uint32)
            d.spin_species = _p(self._ss, C.c_uint32)
        a = self.anis
        for key in ("k2perp", "k2par", "k4perp", "k4par", "kc"):
            setattr(d.anis, key, float(a.get(key, 0.0)))
        for key in ("k2perp_dir", "k2par_dir"):
            getattr(d.anis, key)[:] = list(a.get(key, (0.0, 0.0, 1.0)))
        for key in ("k4perp_dirs", "k4par_dirs"):
            getattr(d.anis, key)[:] = list(np.asarray(a.get(key, np.eye(3)), float).reshape(-1))
        e = self.exch
        d.exch.format = e.get("format", EXCH_NONE)
        self._keep = []
        if d.exch.format == EXCH_CSR:
            vals = np.ascontiguousarray(e["vals"], dtype=np.float64)
            d.exch.ncomp = vals.shape[0] if vals.ndim == 2 else 1
            xadj = np.ascontiguousarray(e["xadj"], dtype=np.uint32)
            adj = np.ascontiguousarray(e["adjncy"], dtype=np.uint32)
            self._keep += [vals, xadj, adj]
            d.exch.nnz = len(adj)
            d.exch.xadj, d.exch.adjncy, d.exch.csr_vals = _p(xadj, C.c_uint32), _p(adj, C.c_uint32), _p(vals, C.c_double)
            d.exch.n_atoms = int(e.get("n_atoms", 0))  # basis size: enables the sublattice-major device order
        elif d.exch.format == EXCH_DIA:
            dv = np.ascontiguousarray(e["vals"], dtype=np.float64)
            offs = np.ascontiguousarray(e["offsets"], dtype=np.int32)
            self._keep += [dv, offs]
            d.exch.ncomp = dv.shape[0]
            d.exch.ndiag = len(offs)
            d.exch.dia_offsets, d.exch.dia_vals = _p(offs, C.c_int32), _p(dv, C.c_double)
        elif d.exch.format == EXCH_STENCIL:
            cnt = np.ascontiguousarray(e["count"], dtype=np.uint32)
            ent = np.ascontiguousarray(e["entries"], dtype=np.int32)
            sv = np.ascontiguousarray(e["vals"], dtype=np.float64)
            self._keep += [cnt, ent, sv]
            d.exch.ncomp = sv.shape[0]
            d.exch.dims[:] = list(e["dims"])
            d.exch.n_atoms = ent.shape[0]
            d.exch.pbc[:] = list(e.get("pbc", (1, 1, 1)))
            d.exch.stencil_k = ent.shape[1]
            d.exch.stencil_count, d.exch.stencil_entries = _p(cnt, C.c_uint32), _p(ent, C.c_int32)
            d.exch.stencil_vals = _p(sv, C.c_double)
            d.exch.stencil_vals_per_row = int(e.get("per_row", 0))
        dp = self.dip
        if dp.get("enabled") or self.fftx:
            self._mesh = np.ascontiguousarray(dp["mesh"], dtype=np.uint32)
            d.dip.enabled = 1 if dp.get("enabled") else 0
            d.dip.every_stage = int(dp.get("every_stage", 0))
            d.dip.refresh = int(dp.get("refresh", 0))
            d.dip.dim[:] = list(dp["dim"])
            d.dip.nk[:] = list(dp["nk"])
            d.dip.abc[:] = list(dp["abc"])
            d.dip.mesh = _p(self._mesh, C.c_uint32)
        d.dt = self.dt
        d.scheme = _SCHEMES[self.scheme] if isinstance(self.scheme, str) else self.scheme
        d.precision = _PRECS[self.precision] if isinstance(self.precision, str) else self.precision
        d.noise = _NOISE[self.noise] if isinstance(self.noise, str) else self.noise
        d.update = self.update
        d.seed = self.seed
        d.stream = self.stream
        d.flags = self.flags
        d.device = self.device
        f = self.four
        if f and len(f.get("jkl", [])):
            fx = np.ascontiguousarray(f["xadj"], dtype=np.uint32)
            fj = np.ascontiguousarray(f["jkl"], dtype=np.uint32).reshape(-1, 3)
            self._keep += [fx, fj]
            d.four.nq = fj.shape[0]
            d.four.xadj = _p(fx, C.c_uint32)
            d.four.jkl = _p(fj, C.c_uint32)
            d.four.jq = float(f["jq"])
            d.four.cpu_form = int(f.get("cpu_form", 0))
        if self.fftx:
            arr = (FftInteraction * len(self.fftx))()
            for k, (s1, s2, dv, J) in enumerate(self.fftx):
                arr[k].s1, arr[k].s2 = int(s1), int(s2)
                arr[k].d[:] = [int(x) for x in dv]
                arr[k].J[:] = [float(x) for x in np.asarray(J, dtype=float).reshape(9)]
            self._keep.append(arr)
            d.fftx.n_int = len(self.fftx)
            d.fftx.ints = C.cast(arr, C.POINTER(FftInteraction))
        return d


class Context:
    """One llg_ctx: a spin system resident on one GPU."""

    def __init__(self, system: System):
        self.L = lib()
        self.sys = system
        self._desc = system.desc()
        h = C.c_void_p()
        _check(self.L.llg_create(C.byref(self._desc), C.byref(h)), self.L)
        self.h = h
        self.n = system.n
        self.nsub = system.n_sublattice

    def close(self):
        if getattr(self, "h", None):
            self.L.llg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _v3(self, fn, s=None):
        if s is None:
            s = np.zeros((3, self.n))
            _check(fn(self.h, _p(s[0], C.c_double), _p(s[1], C.c_double), _p(s[2], C.c_double)), self.L)
            return s
        s = np.ascontiguousarray(s, dtype=np.float64)
        _check(fn(self.h, _p(s[0], C.c_double), _p(s[1], C.c_double), _p(s[2], C.c_double)), self.L)

    def set_spins(self, s):
        self._v3(self.L.llg_set_spins, s)

    def get_spins(self):
        return self._v3(self.L.llg_get_spins)

    def set_temperature(self, T):
        _check(self.L.llg_set_temperature(self.h, float(T)), self.L)

    def set_applied(self, B):
        b = np.ascontiguousarray(B, dtype=np.float64)
        _check(self.L.llg_set_applied(self.h, _p(b, C.c_double)), self.L)

    def set_staggered(self, h=None):
        if h is None:
            _check(self.L.llg_set_staggered(self.h, None, None, None), self.L)
        else:
            self._v3(self.L.llg_set_staggered, h)

    def scale_staggered(self, r):
        _check(self.L.llg_scale_staggered(self.h, float(r)), self.L)

    def push_noise(self, xi):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        nsteps = xi.shape[0] if xi.ndim == 3 else 1
        _check(self.L.llg_push_noise(self.h, nsteps, _p(xi, C.c_double)), self.L)

    def step(self, t0, nsteps):
        _check(self.L.llg_step(self.h, int(t0), int(nsteps)), self.L)

    def sync(self):
        _check(self.L.llg_sync(self.h), self.L)

    def det_field(self):
        return self._v3(self.L.llg_get_det_field)

    def field(self):
        return self._v3(self.L.llg_get_field)

    def reduce(self, energy=True):
        m = np.zeros(self.nsub * 3)
        e = C.c_double(0.0)
        _check(self.L.llg_reduce(self.h, _p(m, C.c_double), C.byref(e) if energy else None), self.L)
        return m.reshape(self.nsub, 3), (e.value if energy else None)

    def reduce_mag_noalloc(self):
        """<M> per sublattice through llg_reduce_mag_noalloc (no device allocation)."""
        m = np.zeros(self.nsub * 3)
        _check(self.L.llg_reduce_mag_noalloc(self.h, _p(m, C.c_double)), self.L)
        return m.reshape(self.nsub, 3)

    def run_sampled(self, t0, nsteps, every):
        nrec = (nsteps + every - 1) // every
        out = np.zeros((max(nrec, 1), self.nsub, 3))
        _check(self.L.llg_run_sampled(self.h, int(t0), int(nsteps), int(every), _p(out, C.c_double)), self.L)
        return out[:nrec]

    def noise_sample(self, t):
        xi = np.zeros((3, self.n))
        _check(self.L.llg_noise_sample(self.h, int(t), _p(xi, C.c_double)), self.L)
        return xi

    def stream_ptr(self) -> int:
        return self.L.llg_stream(self.h) or 0

    def profile_kernels(self, t0, nsteps, max_kinds=16):
        names = C.create_string_buffer(48 * max_kinds)
        ms = np.zeros(max_kinds)
        nk = C.c_int(0)
        _check(self.L.llg_profile_kernels(self.h, int(t0), int(nsteps), max_kinds, names, _p(ms, C.c_double),
                                          C.byref(nk)), self.L)
        raw = names.raw
        out = {}
        for k in range(nk.value):
            nm = raw[48 * k:48 * (k + 1)].split(b"\0", 1)[0].decode()
            out[nm] = float(ms[k])
        return out

    def set_bins(self, nbins, bin_of_spin, weight=None):
        """Upload a bin map (reference spin order) for llg_reduce_bins."""
        b = np.ascontiguousarray(bin_of_spin, dtype=np.uint32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        _check(self.L.llg_set_bins(self.h, int(nbins), _p(b, C.c_uint32), _p(w, C.c_double)), self.L)
        self._nbins = int(nbins)

    def reduce_bins(self):
        out = np.zeros((self._nbins, 3))
        _check(self.L.llg_reduce_bins(self.h, _p(out, C.c_double)), self.L)
        return out

    def set_sq(self, grid, mesh, kpts, axis=2, ua=None, ub=None, ngroups=1, group=None):
        """Wave-vectors (mesh indices) of the transverse structure factor sampled by sample_sq."""
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        m = np.ascontiguousarray(mesh, dtype=np.uint32).reshape(-1, 3)
        k = np.ascontiguousarray(kpts, dtype=np.uint32).reshape(-1, 3)
        a = None if ua is None else np.ascontiguousarray(ua, dtype=np.float64)
        b = None if ub is None else np.ascontiguousarray(ub, dtype=np.float64)
        gr = None if group is None else np.ascontiguousarray(group, dtype=np.uint32)
        _check(self.L.llg_set_sq(self.h, _p(g, C.c_uint32), _p(m, C.c_uint32), len(k), _p(k, C.c_uint32), int(axis),
                                 _p(a, C.c_double), _p(b, C.c_double), int(ngroups), _p(gr, C.c_uint32)), self.L)
        self._sq = (int(ngroups), len(k))

    def sample_sq(self):
        """Complex array [ngroups, nk]."""
        out = np.zeros(self._sq + (2,))
        _check(self.L.llg_sample_sq(self.h, _p(out, C.c_double)), self.L)
        return out[..., 0] + 1j * out[..., 1]

    def autotune(self, nsteps=20):
        """Pick the FP32 Heun launch shape by timing the candidates on this system."""
        r = TuneResult()
        _check(self.L.llg_autotune(self.h, int(nsteps), C.byref(r)), self.L)
        return {"tpb": r.tpb, "xcd": r.xcd, "unroll": r.unroll, "thermal_cache": r.thermal_cache,
                "gather_copy": r.gather_copy, "line_mode": r.line_mode,
                "ms_per_step": r.ms_per_step,
                "ms_per_step_default": r.ms_per_step_default, "candidates": r.candidates,
                "from_cache": bool(r.from_cache)}

    def set_launch(self, tpb, xcd=1, unroll=2, thermal_cache=1, gather_copy=1, line_mode=1):
        _check(self.L.llg_set_launch(self.h, int(tpb), int(xcd), int(unroll), int(thermal_cache), int(gather_copy),
                                     int(line_mode)), self.L)

    def info(self):
        i = Info()
        _check(self.L.llg_get_info(self.h, C.byref(i)), self.L)
        return {"device_bytes": i.device_bytes, "bytes_per_spin_update": i.bytes_per_spin_update,
                "dominant_kernel_bytes_per_spin": i.dominant_kernel_bytes_per_spin,
                "dominant_kernel": i.dominant_kernel.decode(), "exch_encoding": i.exch_encoding.decode()}
