"""The reference's input files -> a System for the new time step.

Thin wrappers over include/llg_io.h (libconfig-format parser, unit-cell file, "unitcell"
exchange file, csr_exch_mat.dat) and `load_reference_input`, which maps a reference
configuration file onto the descriptor:

  system.dim, system.Unitcellfile, system.Nm, system.abc, system.rprim1..3,
  system.Periodic_Boundaries, system.Include_dipole   (src/geom_glob.cpp, src/config.cpp)
  exchange.Read_exch_method = "unitcell", exchange.Exchfile, or exchange.ReadExchangeMatrix
  with exchange.ExchangeMatrixFilename                 (src/exch_routines.cpp:38-114)
  the exchange file's FourSpin quartets (FourSpinExchange_<s>), or 4spin_csr_exch_mat.dat
  with exchange.ReadExchangeMatrix4Spin                (src/exch_read_4_spin.cpp:29-381)
  llg.dt, llg.intscheme, llg.applied, llg.update, llg.staggered_field.sapplied<k>,
  llg.SpinConfigMethod ("random", "restart_text", "grid", or the unit cell's initial
  spins)                                               (src/llg.cpp:237-500)
  anis.k2perp/k2par/k4perp/k4par and their directions (src/anis.cpp)
  Seed                                                  (src/config.cpp:64)

The simulation driver settings (sim.sim_type and its group) are returned as plain values
for the caller's driver loop.
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import (EXCH_CSR, Interaction, System, UnitCellSpec, _check, _p, build_csr, lib, random_spins)

CFG_MISSING, CFG_GROUP, CFG_ARRAY, CFG_LIST, CFG_INT, CFG_FLOAT, CFG_BOOL, CFG_STRING = range(8)


class UcfAtom(C.Structure):
    _fields_ = [("sublattice", C.c_uint32), ("mp", C.c_uint32 * 3), ("xred", C.c_double * 3), ("mu", C.c_double),
                ("lambda_", C.c_double), ("gamma_rel", C.c_double), ("spin0", C.c_double * 3), ("k1u", C.c_double),
                ("k1u_dir", C.c_double * 3), ("element", C.c_char * 16)]


_bound = False


def _io():
    global _bound
    L = lib()
    if not _bound:
        P = C.c_void_p
        up = C.POINTER(C.c_uint32)
        dp = C.POINTER(C.c_double)
        L.llg_cfg_parse_file.argtypes = [C.c_char_p, C.POINTER(P)]
        L.llg_cfg_parse_string.argtypes = [C.c_char_p, C.POINTER(P)]
        L.llg_cfg_free.argtypes = [P]
        L.llg_cfg_free.restype = None
        L.llg_cfg_type.argtypes = [P, C.c_char_p]
        L.llg_cfg_length.argtypes = [P, C.c_char_p, up]
        L.llg_cfg_get_number.argtypes = [P, C.c_char_p, C.c_uint32, dp]
        L.llg_cfg_get_bool.argtypes = [P, C.c_char_p, C.POINTER(C.c_int)]
        L.llg_cfg_get_string.argtypes = [P, C.c_char_p, C.c_char_p, C.c_uint32, up]
        L.llg_cfg_child_name.argtypes = [P, C.c_char_p, C.c_uint32, C.c_char_p, C.c_uint32, up]
        L.llg_read_ucf.argtypes = [C.c_char_p, C.c_int, up, up, C.POINTER(UcfAtom), C.c_uint32]
        L.llg_read_exch_unitcell.argtypes = [P, C.c_uint32, C.POINTER(Interaction), C.c_uint32, up]
        L.llg_csr_read.argtypes = [C.c_char_p, C.c_uint32, up, up, dp, C.POINTER(C.c_uint64)]
        L.llg_csr_write.argtypes = [C.c_char_p, C.c_uint32, up, up, dp, C.c_uint64]
        for nm in ("llg_read_spins_text", "llg_write_spins_text"):
            getattr(L, nm).argtypes = [C.c_char_p, C.c_uint32, dp, dp, dp]
        for nm in ("llg_read_grid", "llg_write_grid"):
            getattr(L, nm).argtypes = [C.c_char_p, C.c_uint32, up, dp, dp, dp]
        L.llg_write_vtu.argtypes = [C.c_char_p, C.c_uint32, dp, dp, dp, dp]
        L.llg_four_spin_quartets.argtypes = [P, C.c_uint32, dp, C.c_uint32, up, up, up, up, up, C.c_int, up, up,
                                             C.c_uint64, C.POINTER(C.c_uint64), dp]
        L.llg_four_csr_read.argtypes = [C.c_char_p, C.c_uint32, up, up, C.POINTER(C.c_uint64)]
        L.llg_four_csr_write.argtypes = [C.c_char_p, C.c_uint32, up, up, C.c_uint64]
        _bound = True
    return L


class Config:
    """A parsed libconfig document (llg_cfg)."""

    def __init__(self, path: Optional[str] = None, text: Optional[str] = None):
        self.L = _io()
        self.h = C.c_void_p()
        if path is not None:
            _check(self.L.llg_cfg_parse_file(path.encode(), C.byref(self.h)), self.L)
            self.path = path
        else:
            _check(self.L.llg_cfg_parse_string(text.encode(), C.byref(self.h)), self.L)
            self.path = None

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.llg_cfg_free(self.h)
            self.h = C.c_void_p()

    def type(self, path: str) -> int:
        return self.L.llg_cfg_type(self.h, path.encode())

    def exists(self, path: str) -> bool:
        return self.type(path) != CFG_MISSING

    def length(self, path: str) -> int:
        n = C.c_uint32()
        _check(self.L.llg_cfg_length(self.h, path.encode(), C.byref(n)), self.L)
        return n.value

    def children(self, path: str):
        """Setting names of a group, in file order."""
        out = []
        for k in range(self.length(path)):
            n = C.c_uint32()
            _check(self.L.llg_cfg_child_name(self.h, path.encode(), k, None, 0, C.byref(n)), self.L)
            buf = C.create_string_buffer(n.value + 1)
            _check(self.L.llg_cfg_child_name(self.h, path.encode(), k, buf, n.value + 1, None), self.L)
            out.append(buf.value.decode())
        return out

    def number(self, path: str, index: int = 0) -> float:
        v = C.c_double()
        _check(self.L.llg_cfg_get_number(self.h, path.encode(), index, C.byref(v)), self.L)
        return v.value

    def get(self, path: str, default=None):
        """Python value of a scalar or array setting (default when missing)."""
        t = self.type(path)
        if t == CFG_MISSING:
            return default
        if t == CFG_STRING:
            n = C.c_uint32()
            _check(self.L.llg_cfg_get_string(self.h, path.encode(), None, 0, C.byref(n)), self.L)
            buf = C.create_string_buffer(n.value + 1)
            _check(self.L.llg_cfg_get_string(self.h, path.encode(), buf, n.value + 1, None), self.L)
            return buf.value.decode()
        if t == CFG_BOOL:
            b = C.c_int()
            _check(self.L.llg_cfg_get_bool(self.h, path.encode(), C.byref(b)), self.L)
            return bool(b.value)
        if t == CFG_INT:
            return int(self.number(path))
        if t == CFG_FLOAT:
            return self.number(path)
        if t == CFG_ARRAY:
            return [self.number(path, k) for k in range(self.length(path))]
        raise ValueError(f"{path}: groups and lists are read through their members")


def read_ucf(path: str, with_mesh: bool = False):
    L = _io()
    ns, na = C.c_uint32(), C.c_uint32()
    _check(L.llg_read_ucf(path.encode(), int(with_mesh), C.byref(ns), C.byref(na), None, 0), L)
    atoms = (UcfAtom * na.value)()
    _check(L.llg_read_ucf(path.encode(), int(with_mesh), C.byref(ns), C.byref(na), atoms, na.value), L)
    return ns.value, list(atoms)


def read_exch_unitcell(cfg: Config, n_atoms: int):
    L = _io()
    cnt = C.c_uint32()
    _check(L.llg_read_exch_unitcell(cfg.h, n_atoms, None, 0, C.byref(cnt)), L)
    out = (Interaction * cnt.value)()
    if cnt.value:
        _check(L.llg_read_exch_unitcell(cfg.h, n_atoms, out, cnt.value, C.byref(cnt)), L)
    return out


def csr_read(path: str, n: int):
    L = _io()
    nnz = C.c_uint64()
    _check(L.llg_csr_read(path.encode(), n, None, None, None, C.byref(nnz)), L)
    xadj = np.zeros(n + 1, np.uint32)
    adj = np.zeros(nnz.value, np.uint32)
    vals = np.zeros((3, nnz.value))
    _check(L.llg_csr_read(path.encode(), n, _p(xadj, C.c_uint32), _p(adj, C.c_uint32), _p(vals, C.c_double),
                          C.byref(nnz)), L)
    return xadj, adj, vals


def csr_write(path: str, xadj, adjncy, vals3):
    L = _io()
    xadj = np.ascontiguousarray(xadj, np.uint32)
    adj = np.ascontiguousarray(adjncy, np.uint32)
    v = np.ascontiguousarray(vals3, np.float64).reshape(3, -1)
    _check(L.llg_csr_write(path.encode(), len(xadj) - 1, _p(xadj, C.c_uint32), _p(adj, C.c_uint32),
                           _p(v, C.c_double), len(adj)), L)


FOUR_ROUND = {"nearest": 0, "reference": 1}


def four_spin_quartets(ex: Config, nms: int, mu_species, mesh_xyz, sublattice, grid, nk, pbc, rounding="nearest"):
    """Four-spin quartets of every spin from the exchange file's FourSpinExchange_<s> groups
    (llg_four_spin_quartets; src/exch_read_4_spin.cpp:29-258).  rounding "nearest" rounds the
    quartet offsets to mesh points half away from zero; "reference" repeats the reference's
    static_cast<int>(x + 0.5), which moves negative offsets one mesh point up.  Returns
    (xadj [n+1], jkl [nq, 3], jq per species in tesla)."""
    L = _io()
    u32 = lambda a: np.ascontiguousarray(a, np.uint32)
    mesh_xyz, sub = u32(mesh_xyz).reshape(-1, 3), u32(sublattice)
    n = len(sub)
    g, k, p = u32(grid), u32(nk), u32(pbc)
    mu = np.ascontiguousarray(mu_species, np.float64)
    nq = C.c_uint64()
    args = (ex.h, nms, _p(mu, C.c_double), n, _p(mesh_xyz, C.c_uint32), _p(sub, C.c_uint32), _p(g, C.c_uint32),
            _p(k, C.c_uint32), _p(p, C.c_uint32), FOUR_ROUND[rounding])
    _check(L.llg_four_spin_quartets(*args, None, None, 0, C.byref(nq), None), L)
    xadj = np.zeros(n + 1, np.uint32)
    jkl = np.zeros((nq.value, 3), np.uint32)
    jq = np.zeros(nms)
    _check(L.llg_four_spin_quartets(*args, _p(xadj, C.c_uint32), _p(jkl, C.c_uint32), nq.value, C.byref(nq),
                                    _p(jq, C.c_double)), L)
    return xadj, jkl, jq


def four_csr_read(path: str, n: int):
    """4spin_csr_exch_mat.dat -> (xadj [n+1], jkl [nq, 3])."""
    L = _io()
    nq = C.c_uint64()
    _check(L.llg_four_csr_read(path.encode(), n, None, None, C.byref(nq)), L)
    xadj = np.zeros(n + 1, np.uint32)
    jkl = np.zeros((nq.value, 3), np.uint32)
    _check(L.llg_four_csr_read(path.encode(), n, _p(xadj, C.c_uint32), _p(jkl, C.c_uint32), C.byref(nq)), L)
    return xadj, jkl


def four_csr_write(path: str, xadj, jkl):
    L = _io()
    xadj = np.ascontiguousarray(xadj, np.uint32)
    jkl = np.ascontiguousarray(jkl, np.uint32).reshape(-1, 3)
    _check(L.llg_four_csr_write(path.encode(), len(xadj) - 1, _p(xadj, C.c_uint32), _p(jkl, C.c_uint32), len(jkl)), L)


def read_spins_text(path: str, n: int):
    """restart_text spin file -> (3, n), normalised."""
    L = _io()
    s = np.zeros((3, n))
    _check(L.llg_read_spins_text(path.encode(), n, _p(s[0], C.c_double), _p(s[1], C.c_double), _p(s[2], C.c_double)), L)
    return s


def write_spins_text(path: str, s):
    L = _io()
    s = np.ascontiguousarray(s, np.float64)
    _check(L.llg_write_spins_text(path.encode(), s.shape[1], _p(s[0], C.c_double), _p(s[1], C.c_double),
                                  _p(s[2], C.c_double)), L)


def read_grid(path: str, mesh):
    """grid spin file (x y z sx sy sz per line, any order) -> (3, n) in spin order."""
    L = _io()
    m = np.ascontiguousarray(mesh, np.uint32).reshape(-1, 3)
    n = m.shape[0]
    s = np.zeros((3, n))
    _check(L.llg_read_grid(path.encode(), n, _p(m, C.c_uint32), _p(s[0], C.c_double), _p(s[1], C.c_double),
                           _p(s[2], C.c_double)), L)
    return s


def write_grid(path: str, mesh, s):
    L = _io()
    m = np.ascontiguousarray(mesh, np.uint32).reshape(-1, 3)
    s = np.ascontiguousarray(s, np.float64)
    _check(L.llg_write_grid(path.encode(), m.shape[0], _p(m, C.c_uint32), _p(s[0], C.c_double), _p(s[1], C.c_double),
                            _p(s[2], C.c_double)), L)


def write_vtu(path: str, pos, s):
    L = _io()
    p = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    s = np.ascontiguousarray(s, np.float64)
    _check(L.llg_write_vtu(path.encode(), p.shape[0], _p(p, C.c_double), _p(s[0], C.c_double), _p(s[1], C.c_double),
                           _p(s[2], C.c_double)), L)


@dataclass
class ReferenceInput:
    system: System
    spins: np.ndarray               # (3, n) initial spins
    dims: tuple
    n_atoms: int
    applied: tuple                  # llg.applied (T)
    staggered: Optional[np.ndarray]  # (3, n) per-spin staggered field or None
    sim: dict = field(default_factory=dict)  # sim.sim_type and the driver group's scalars
    atoms: list = field(default_factory=list)


def _ncomp_of(ints, n):
    J = np.array([list(ints[k].J) for k in range(n)]).reshape(n, 3, 3) if n else np.zeros((0, 3, 3))
    off = J.copy()
    off[:, [0, 1, 2], [0, 1, 2]] = 0.0
    if np.any(off != 0.0):
        return 9
    d = J[:, [0, 1, 2], [0, 1, 2]]
    if np.any(d[:, 0] != d[:, 1]) or np.any(d[:, 1] != d[:, 2]):
        return 3
    return 1


def load_reference_input(cff_path: str, precision: str = "fp32", scheme: Optional[str] = None,
                         noise: str = "counter", four_rounding: str = "nearest") -> ReferenceInput:
    """Build the System a reference configuration file describes (paths inside it are
    relative to the configuration file's directory, as when the reference runs there).
    With exchange.FourSpin set in the exchange file, the four-spin quartets are built from
    its FourSpinExchange_<s> groups (rounding of their offsets: four_spin_quartets), or read
    from 4spin_csr_exch_mat.dat when the configuration sets exchange.ReadExchangeMatrix4Spin;
    the term takes species 0's JQ, as the reference's GPU kernel does."""
    base = os.path.dirname(os.path.abspath(cff_path))
    cfg = Config(cff_path)

    def rel(p):
        return p if os.path.isabs(p) else os.path.join(base, p)

    dims = tuple(int(x) for x in cfg.get("system.dim"))
    with_mesh = cfg.exists("system.Nm")
    nms, atoms = read_ucf(rel(cfg.get("system.Unitcellfile")), with_mesh)
    na = len(atoms)
    rprim = np.eye(3)
    for k in range(3):
        v = cfg.get(f"system.rprim{k + 1}")
        if v is not None:
            rprim[k] = v
    pbc = tuple(int(bool(b)) for b in cfg.get("system.Periodic_Boundaries", [True, True, True]))
    nk = tuple(int(x) for x in cfg.get("system.Nm", [1, 1, 1]))
    uc = UnitCellSpec(np.array([list(a.xred) for a in atoms], float), rprim, nk,
                      np.array([list(a.mp) for a in atoms], np.uint32))
    n = int(np.prod(dims)) * na
    mu_atom = [a.mu for a in atoms]

    exfile = cfg.get("exchange.Exchfile")
    ex = Config(rel(exfile)) if exfile is not None else None
    if cfg.get("exchange.ReadExchangeMatrix", False):
        xadj, adj, vals = csr_read(rel(cfg.get("exchange.ExchangeMatrixFilename", "csr_exch_mat.dat")), n)
        ncomp = 3 if np.any(vals[0] != vals[1]) or np.any(vals[1] != vals[2]) else 1
        vals = vals if ncomp == 3 else vals[:1]
    else:
        method = cfg.get("exchange.Read_exch_method", "unitcell")
        if method != "unitcell":
            raise NotImplementedError(f"exchange.Read_exch_method = {method!r}: only 'unitcell' and "
                                      f"exchange.ReadExchangeMatrix are read")
        if cfg.get("exchange.TruncateExchange", False):
            raise NotImplementedError("exchange.TruncateExchange = TRUE is not read")
        if ex is None:
            raise ValueError("exchange.Exchfile is not set")
        ints = read_exch_unitcell(ex, na)
        ncomp = _ncomp_of(ints, len(ints))
        xadj, adj, vals = build_csr(uc, dims, ints, mu_atom, ncomp=ncomp, pbc=pbc)

    species = [dict(mu=a.mu, lambda_=a.lambda_, gamma_rel=a.gamma_rel, k1u=a.k1u, k1u_dir=tuple(a.k1u_dir),
                    sublattice=a.sublattice) for a in atoms]
    spin_species = np.tile(np.arange(na, dtype=np.uint32), n // na)
    anis = {}
    for key in ("k2perp", "k2par", "k4perp", "k4par"):
        if cfg.exists(f"anis.{key}"):
            anis[key] = cfg.number(f"anis.{key}")
    for key in ("k2perp", "k2par"):
        v = cfg.get(f"anis.{key}dir")
        if v is not None:
            anis[f"{key}_dir"] = tuple(v)
    for key in ("k4perp", "k4par"):
        rows = [cfg.get(f"anis.{key}dir{k}") for k in (1, 2, 3)]
        if any(r is not None for r in rows):
            anis[f"{key}_dirs"] = np.array([r if r is not None else [0.0, 0.0, 0.0] for r in rows])
    sch = scheme or cfg.get("llg.intscheme", "heun").lower()
    seed = int(cfg.get("Seed", 0))
    sysd = System(n=n, species=species, spin_species=spin_species if na > 1 else None, n_sublattice=nms,
                  anis=anis, exch=dict(format=EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals, n_atoms=na),
                  dt=float(cfg.get("llg.dt", 1e-16)), scheme=sch, precision=precision, noise=noise,
                  update=int(cfg.get("llg.update", 1)), seed=seed)
    if cfg.get("system.Include_dipole", False):
        from . import mesh as _mesh
        sysd.dip = dict(enabled=1, mesh=_mesh(uc, dims), dim=dims, nk=nk, abc=tuple(cfg.get("system.abc")))
    if ex is not None and ex.get("exchange.FourSpin", False):
        from . import mesh as _mesh
        sub = np.array([atoms[i % na].sublattice for i in range(n)], np.uint32)
        mu_s = np.zeros(nms)
        for a in atoms:
            mu_s[a.sublattice] = a.mu
        grid = tuple(int(d) * int(k) for d, k in zip(dims, nk))
        fx, fj, jq = four_spin_quartets(ex, nms, mu_s, _mesh(uc, dims), sub, grid, nk, pbc, four_rounding)
        if cfg.get("exchange.ReadExchangeMatrix4Spin", False):
            fx, fj = four_csr_read(rel("4spin_csr_exch_mat.dat"), n)
        sysd.four = dict(xadj=fx, jkl=fj, jq=float(jq[0]), cpu_form=0)

    scm = cfg.get("llg.SpinConfigMethod")
    if scm == "random":
        s0 = random_spins(n, seed, seed + 100)
    elif scm == "restart_text":
        s0 = read_spins_text(rel(cfg.get("llg.SpinFile")), n)
    elif scm == "grid":
        from . import mesh as _mesh
        s0 = read_grid(rel(cfg.get("llg.SpinFile")), _mesh(uc, dims))
    elif scm is None or scm == "align":
        s0 = np.array([list(atoms[i % na].spin0) for i in range(n)], float).T.copy()
        s0 /= np.linalg.norm(s0, axis=0)
    else:
        raise NotImplementedError(f"llg.SpinConfigMethod = {scm!r} is not read")

    stag = None
    if cfg.exists("llg.staggered_field"):
        sof = np.array([cfg.get(f"llg.staggered_field.sapplied{k + 1}") for k in range(nms)], float)
        sub = np.array([atoms[i % na].sublattice for i in range(n)])
        stag = sof[sub].T.copy()
    applied = tuple(cfg.get("llg.applied", [0.0, 0.0, 0.0]))
    sim = {}
    st = cfg.get("sim.sim_type")
    if st is not None:
        sim["sim_type"] = st
        if cfg.type(st) == CFG_GROUP:
            for name in cfg.children(st):
                t = cfg.type(f"{st}.{name}")
                if t in (CFG_INT, CFG_FLOAT, CFG_BOOL, CFG_STRING, CFG_ARRAY):
                    sim[name] = cfg.get(f"{st}.{name}")
    return ReferenceInput(system=sysd, spins=s0, dims=dims, n_atoms=na, applied=applied, staggered=stag,
                          sim=sim, atoms=atoms)
