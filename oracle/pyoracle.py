"""ctypes wrapper of liboracle.so - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker or the timed CPU baseline.  OracleContext mirrors sllg.Context
call for call so a test can drive both with the same sequence.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "gpu-atomistic-code_amd"))
import sllg  # noqa: E402  (descriptor types only)

LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_RANDOM = os.path.join(HERE, "_ref", "ref_random")
COMPAT_LLGCPU = 0x1

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.or_create.argtypes = [C.POINTER(sllg.Desc), C.c_uint32, C.POINTER(P)]
        L.or_destroy.argtypes = [P]
        L.or_destroy.restype = None
        for f in ("or_set_spins", "or_get_spins", "or_set_staggered", "or_get_det_field", "or_get_field",
                  "or_get_dipolar"):
            getattr(L, f).argtypes = [P, dp, dp, dp]
        L.or_set_temperature.argtypes = [P, C.c_double]
        L.or_set_applied.argtypes = [P, dp]
        L.or_scale_staggered.argtypes = [P, C.c_double]
        L.or_push_noise.argtypes = [P, C.c_uint32, dp]
        L.or_step.argtypes = [P, C.c_uint64, C.c_uint32]
        L.or_reduce.argtypes = [P, dp, dp]
        L.or_noise_sample.argtypes = [P, C.c_uint64, dp]
        L.or_run_sampled.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, dp]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_philox4x32_10.restype = None
        L.or_counter_normals_f64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint64, dp]
        L.or_counter_normals_f64.restype = None
        L.or_mwc_seed.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.or_mwc_seed.restype = None
        L.or_mwc_uniform.argtypes = [C.c_void_p]
        L.or_mwc_uniform.restype = C.c_double
        L.or_mwc_normal.argtypes = [C.c_void_p]
        L.or_mwc_normal.restype = C.c_double
        L.or_random_spins.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, dp, dp, dp]
        L.or_random_spins.restype = None
        up = C.POINTER(C.c_uint32)
        L.or_build_csr.argtypes = [C.POINTER(sllg.UnitCell), up, up, C.c_uint32, C.POINTER(sllg.Interaction), dp,
                                   C.c_int, up, up, dp, C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


class OracleContext:
    def __init__(self, system: "sllg.System", options: int = 0):
        self.L = lib()
        self._desc = system.desc()
        h = C.c_void_p()
        _chk(self.L.or_create(C.byref(self._desc), options, C.byref(h)))
        self.h = h
        self.n = system.n
        self.nsub = system.n_sublattice

    def close(self):
        if getattr(self, "h", None):
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _v3(self, fn, s=None):
        if s is None:
            s = np.zeros((3, self.n))
            _chk(fn(self.h, _p(s[0]), _p(s[1]), _p(s[2])))
            return s
        s = np.ascontiguousarray(s, dtype=np.float64)
        _chk(fn(self.h, _p(s[0]), _p(s[1]), _p(s[2])))

    def set_spins(self, s):
        self._v3(self.L.or_set_spins, s)

    def get_spins(self):
        return self._v3(self.L.or_get_spins)

    def set_temperature(self, T):
        _chk(self.L.or_set_temperature(self.h, float(T)))

    def set_applied(self, B):
        b = np.ascontiguousarray(B, dtype=np.float64)
        _chk(self.L.or_set_applied(self.h, _p(b)))

    def set_staggered(self, h=None):
        if h is None:
            _chk(self.L.or_set_staggered(self.h, None, None, None))
        else:
            self._v3(self.L.or_set_staggered, h)

    def scale_staggered(self, r):
        _chk(self.L.or_scale_staggered(self.h, float(r)))

    def push_noise(self, xi):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        nsteps = xi.shape[0] if xi.ndim == 3 else 1
        _chk(self.L.or_push_noise(self.h, nsteps, _p(xi)))

    def step(self, t0, nsteps):
        _chk(self.L.or_step(self.h, int(t0), int(nsteps)))

    def det_field(self):
        return self._v3(self.L.or_get_det_field)

    def field(self):
        return self._v3(self.L.or_get_field)

    def dipolar(self):
        return self._v3(self.L.or_get_dipolar)

    def reduce(self, energy=True):
        m = np.zeros(self.nsub * 3)
        e = C.c_double(0.0)
        _chk(self.L.or_reduce(self.h, _p(m), C.byref(e) if energy else None))
        return m.reshape(self.nsub, 3), (e.value if energy else None)

    def run_sampled(self, t0, nsteps, every):
        nrec = (nsteps + every - 1) // every
        out = np.zeros((max(nrec, 1), self.nsub, 3))
        _chk(self.L.or_run_sampled(self.h, int(t0), int(nsteps), int(every), _p(out)))
        return out[:nrec]

    def noise_sample(self, t):
        xi = np.zeros((3, self.n))
        _chk(self.L.or_noise_sample(self.h, int(t), _p(xi)))
        return xi


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def counter_normals(seed, stream, spin, t):
    z = np.zeros(3)
    lib().or_counter_normals_f64(seed, stream, spin, t, _p(z))
    return z


class MWC:
    def __init__(self, a, b):
        self.buf = C.create_string_buffer(32)
        lib().or_mwc_seed(self.buf, a, b)

    def uniform(self):
        return lib().or_mwc_uniform(self.buf)

    def normal(self):
        return lib().or_mwc_normal(self.buf)


def random_spins(a, b, n):
    s = np.zeros((3, n))
    lib().or_random_spins(a, b, n, _p(s[0]), _p(s[1]), _p(s[2]))
    return s


def build_csr(uc: "sllg.UnitCellSpec", dims, ints, mu_atom, ncomp=1, pbc=(1, 1, 1)):
    L = lib()
    ucc = uc.c()
    up = C.POINTER(C.c_uint32)
    dims_a = np.ascontiguousarray(dims, dtype=np.uint32)
    pbc_a = np.ascontiguousarray(pbc, dtype=np.uint32)
    mu = np.ascontiguousarray(mu_atom, dtype=np.float64)
    nnz = C.c_uint64(0)
    _chk(L.or_build_csr(C.byref(ucc), dims_a.ctypes.data_as(up), pbc_a.ctypes.data_as(up), len(ints), ints, _p(mu),
                        ncomp, None, None, None, C.byref(nnz)))
    n = int(np.prod(dims_a)) * uc.n_atoms
    xadj = np.zeros(n + 1, np.uint32)
    adj = np.zeros(nnz.value, np.uint32)
    vals = np.zeros(ncomp * nnz.value)
    _chk(L.or_build_csr(C.byref(ucc), dims_a.ctypes.data_as(up), pbc_a.ctypes.data_as(up), len(ints), ints, _p(mu),
                        ncomp, xadj.ctypes.data_as(up), adj.ctypes.data_as(up), _p(vals), C.byref(nnz)))
    return xadj, adj, vals.reshape(ncomp, -1)
