"""ctypes binding of include/llg_batch.h: batches of independent spin systems and the
continuous-batching request loop."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Desc, LLGError, _check, _p, lib

BATCH_GRAPH = 0x1
BATCH_PERSISTENT = 0x2
SPIN_DEFAULT = 0x1000000  # polls per grid-barrier wait of a persistent launch
REQ_MT, REQ_RAMP, REQ_SW = 0, 1, 2


class BatchInfo(C.Structure):
    _fields_ = [("members", C.c_uint32), ("tiles", C.c_uint32), ("spins", C.c_uint64), ("device_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64), ("graphs_built", C.c_uint32), ("persist_grid", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("type", C.c_uint32), ("L", C.c_uint32), ("steps", C.c_uint32), ("update", C.c_uint32),
                ("T", C.c_double), ("J", C.c_double), ("K", C.c_double), ("mu", C.c_double), ("lambda_", C.c_double),
                ("dt", C.c_double), ("B0", C.c_double * 3), ("B1", C.c_double * 3), ("arrival_s", C.c_double),
                ("seed", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("latency_s", C.c_double), ("m_final", C.c_double * 3), ("m_mean", C.c_double * 3),
                ("samples", C.c_uint32), ("pad_", C.c_uint32)]


ProgressFn = C.CFUNCTYPE(None, C.c_uint32, C.c_uint32, C.c_double, C.c_void_p)


class ServeConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("flags", C.c_uint32), ("hbm_budget_bytes", C.c_uint64),
                ("max_members", C.c_uint32), ("chunk", C.c_uint32), ("sequential", C.c_int), ("precision", C.c_int),
                ("progress", ProgressFn), ("progress_user", C.c_void_p), ("progress_every", C.c_uint32),
                ("max_live_spins", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("solves_per_s", C.c_double), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double), ("spin_updates_per_s", C.c_double),
                ("completed", C.c_uint32), ("max_live", C.c_uint32), ("spin_updates", C.c_uint64)]


class BatchTune(C.Structure):
    _fields_ = [("tpb", C.c_int), ("xcd", C.c_int), ("ms_per_step", C.c_double),
                ("ms_per_step_default", C.c_double), ("candidates", C.c_int), ("line_mode", C.c_int),
                ("gather_copy", C.c_int)]


_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.llg_estimate_bytes.argtypes = [C.POINTER(Desc)]
    L.llg_estimate_bytes.restype = C.c_uint64
    L.llg_batch_create.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(P)]
    L.llg_batch_destroy.argtypes = [P]
    L.llg_batch_destroy.restype = None
    L.llg_batch_add.argtypes = [P, C.POINTER(Desc), dp, dp, dp, C.c_double, dp, C.c_uint64, C.POINTER(C.c_int)]
    L.llg_batch_remove.argtypes = [P, C.c_int]
    L.llg_batch_set_temperature.argtypes = [P, C.c_int, C.c_double]
    L.llg_batch_set_applied.argtypes = [P, C.c_int, dp]
    L.llg_batch_step.argtypes = [P, C.c_uint32]
    L.llg_batch_member_status.argtypes = [P, C.c_int, C.POINTER(C.c_int)]
    L.llg_batch_reduce.argtypes = [P, C.POINTER(C.c_int), C.c_uint32, C.c_uint32, dp]
    L.llg_batch_get_spins.argtypes = [P, C.c_int, dp, dp, dp]
    L.llg_batch_get_step.argtypes = [P, C.c_int, C.POINTER(C.c_uint64)]
    L.llg_batch_get_info.argtypes = [P, C.POINTER(BatchInfo)]
    L.llg_batch_set_launch.argtypes = [P, C.c_uint32, C.c_int, C.c_int, C.c_int]
    L.llg_batch_autotune.argtypes = [P, C.c_uint32, C.POINTER(BatchTune)]
    L.llg_batch_set_persistent.argtypes = [P, C.c_int, C.c_uint32]
    L.llg_serve_run.argtypes = [C.POINTER(ServeConfig), C.POINTER(Request), C.c_uint32, C.POINTER(Result), dp,
                                C.POINTER(C.c_uint64), dp, C.POINTER(C.c_uint64), C.POINTER(ServeStats)]
    _bound = True
    return L


def estimate_bytes(system) -> int:
    L = _bind()
    d = system.desc()
    return int(L.llg_estimate_bytes(C.byref(d)))


class Batch:
    def __init__(self, device=0, budget_bytes=0, flags=0):
        self.L = _bind()
        h = C.c_void_p()
        _check(self.L.llg_batch_create(device, int(budget_bytes), flags, C.byref(h)), self.L)
        self.h = h
        self._sys = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.llg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, system, s0, T=0.0, B=(0.0, 0.0, 0.0), t0=0):
        d = system.desc()
        s0 = np.ascontiguousarray(s0, dtype=np.float64)
        b = np.ascontiguousarray(B, dtype=np.float64)
        m = C.c_int(-1)
        _check(self.L.llg_batch_add(self.h, C.byref(d), _p(s0[0], C.c_double), _p(s0[1], C.c_double),
                                    _p(s0[2], C.c_double), float(T), _p(b, C.c_double), int(t0), C.byref(m)), self.L)
        self._sys[m.value] = system
        return m.value

    def remove(self, member):
        _check(self.L.llg_batch_remove(self.h, member), self.L)
        self._sys.pop(member, None)

    def set_temperature(self, member, T):
        _check(self.L.llg_batch_set_temperature(self.h, member, float(T)), self.L)

    def set_applied(self, member, B):
        b = np.ascontiguousarray(B, dtype=np.float64)
        _check(self.L.llg_batch_set_applied(self.h, member, _p(b, C.c_double)), self.L)

    def step(self, nsteps):
        _check(self.L.llg_batch_step(self.h, int(nsteps)), self.L)

    def member_status(self, member):
        """0, or the status (-7, NONFINITE) of a member taken out by a failed finite check."""
        st = C.c_int(0)
        _check(self.L.llg_batch_member_status(self.h, member, C.byref(st)), self.L)
        return st.value

    def set_launch(self, tpb, xcd=1, line_mode=1, gather_copy=0):
        _check(self.L.llg_batch_set_launch(self.h, int(tpb), int(xcd), int(line_mode), int(gather_copy)), self.L)

    def set_persistent(self, on=True, spin_limit=SPIN_DEFAULT):
        """Advance each chunk in persistent launches (one cooperative launch per group of
        members and per run of steps between dipolar refreshes, grid barriers between half
        steps, every wait bounded by spin_limit polls)."""
        _check(self.L.llg_batch_set_persistent(self.h, 1 if on else 0, int(spin_limit)), self.L)

    def autotune(self, nsteps=10):
        """Time every tile size, block order and gather copy (and both readers when a member
        has a line encoding) on the live members (their spins and step counters are
        restored) and keep the fastest; returns the choice."""
        r = BatchTune()
        _check(self.L.llg_batch_autotune(self.h, int(nsteps), C.byref(r)), self.L)
        return {f: getattr(r, f) for f, _ in BatchTune._fields_}

    def reduce(self, members, stride=3):
        ms = np.ascontiguousarray(members, dtype=np.int32)
        out = np.zeros(len(ms) * stride)
        _check(self.L.llg_batch_reduce(self.h, ms.ctypes.data_as(C.POINTER(C.c_int)), len(ms), stride,
                                       _p(out, C.c_double)), self.L)
        return out.reshape(len(ms), stride)

    def get_spins(self, member):
        n = self._sys[member].n
        s = np.zeros((3, n))
        _check(self.L.llg_batch_get_spins(self.h, member, _p(s[0], C.c_double), _p(s[1], C.c_double),
                                          _p(s[2], C.c_double)), self.L)
        return s

    def get_step(self, member):
        t = C.c_uint64(0)
        _check(self.L.llg_batch_get_step(self.h, member, C.byref(t)), self.L)
        return t.value

    def info(self):
        i = BatchInfo()
        _check(self.L.llg_batch_get_info(self.h, C.byref(i)), self.L)
        return {f: getattr(i, f) for f, _ in BatchInfo._fields_}


def make_requests(specs):
    """specs: list of dicts with Request fields (B0/B1 as 3-sequences)."""
    arr = (Request * len(specs))()
    for k, s in enumerate(specs):
        r = arr[k]
        r.type = s.get("type", REQ_MT)
        r.L = s["L"]
        r.steps = s["steps"]
        r.update = s.get("update", 100)
        r.T = s.get("T", 0.0)
        r.J = s.get("J", 3e-21)
        r.K = s.get("K", 1e-23)
        r.mu = s.get("mu", 2.22)
        r.lambda_ = s.get("lambda_", 0.1)
        r.dt = s.get("dt", 1e-16)
        r.B0[:] = list(s.get("B0", (0.0, 0.0, 0.0)))
        r.B1[:] = list(s.get("B1", (0.0, 0.0, 0.0)))
        r.arrival_s = s.get("arrival_s", 0.0)
        r.seed = s.get("seed", k + 1)
    return arr


def serve(requests, device=0, budget_bytes=0, max_members=0, chunk=0, sequential=False, precision=1, graph=False,
          want_series=False, progress=None, progress_every=0, max_live_spins=0, persistent=False):
    """Run llg_serve_run; returns (results list of dicts, stats dict, series list or None).

    With want_series, series[i] is request i's <M> samples (ceil(steps/update) x 3) and, for
    a spin-wave request, results[i]["sq"] its S(k) samples (ceil(steps/update) x (L/2+1))
    as complex numbers at k = (q, 0, 0).  progress(completed, nreq, wall_s) is called after
    every progress_every-th completion and after the last one."""
    L = _bind()
    n = len(requests)
    hook = ProgressFn(lambda done, total, wall, user: progress(done, total, wall)) if progress else ProgressFn()
    flags = (BATCH_GRAPH if graph else 0) | (BATCH_PERSISTENT if persistent else 0)
    cfg = ServeConfig(device, flags, int(budget_bytes), max_members, chunk,
                      1 if sequential else 0, precision, hook, None, progress_every if progress else 0,
                      int(max_live_spins))
    res = (Result * n)()
    st = ServeStats()
    series = offs = sq = sq_offs = None
    counts = [(r.steps + r.update - 1) // r.update for r in requests]
    if want_series:
        offs = np.zeros(n, np.uint64)
        offs[1:] = np.cumsum(counts)[:-1]
        series = np.zeros(int(sum(counts)) * 3)
        sqn = [counts[k] * (r.L // 2 + 1) * 2 if r.type == REQ_SW else 0 for k, r in enumerate(requests)]
        sq_offs = np.zeros(n, np.uint64)
        sq_offs[1:] = np.cumsum(sqn)[:-1]
        sq = np.zeros(max(int(sum(sqn)), 1))
    u64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64)) if a is not None else None
    _check(L.llg_serve_run(C.byref(cfg), requests, n, res, _p(series, C.c_double) if series is not None else None,
                           u64(offs), _p(sq, C.c_double) if sq is not None else None, u64(sq_offs), C.byref(st)), L)
    out = [dict(arrival_s=r.arrival_s, admit_s=r.admit_s, done_s=r.done_s, latency_s=r.latency_s,
                m_final=list(r.m_final), m_mean=list(r.m_mean), samples=r.samples) for r in res]
    stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
    ser = None
    if series is not None:
        ser = []
        for k, r in enumerate(requests):
            c = counts[k]
            o = int(offs[k])
            ser.append(series[3 * o:3 * (o + c)].reshape(c, 3))
            if r.type == REQ_SW:
                nk = r.L // 2 + 1
                q = int(sq_offs[k])
                v = sq[q:q + 2 * nk * c].reshape(c, nk, 2)
                out[k]["sq"] = v[..., 0] + 1j * v[..., 1]
    return out, stats, ser
