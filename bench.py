"""Benchmark of the stochastic-LLG time step on one MI355X.

Default workload (the 1,048,576-spin case): fcc 64x64x64 conventional cells x 4 atoms,
periodic, CSR exchange to the 3rd shell (42 nnz/row), cubic anisotropy, Heun with the
FP32 fused field assembly + integrator, T = 300 K counter-based thermal noise, spins
random from seed 42.  A "step" is one full Heun step (predictor + corrector) of all
spins; value = spin-updates/s = n * K / (timed seconds) with the state resident in HBM.

The launch shape (reader: LDS-staged line reader, line reader or packed entries; block
size; XCD order; thermal hand-off; AoS3 copy) is picked by the library's autotuner before
timing and reported in launch_tune.

Also measured in the same run and carried in the same JSON line:
  roofline       the longer Heun kernel's algorithmic bytes / its average launch duration (HIP
                 events around each launch on the handle's stream; per_kernel gives both kernels
                 and their shares of the timed step), and its HBM
                 traffic per launch from two rocprofv3 PMC child runs of this file (FETCH_SIZE,
                 WRITE_SIZE), each scaled by the same counter on a 1 GiB device copy
  by_temperature the same system at T = 0, 300 and 600 K
  beyond_cache   fcc 192 x 192 x 64 x 4 = 9,437,184 spins, whose stepping state (S, Y, A: 340 MB)
                 exceeds the 256 MB Infinity Cache
  unshared_encoding  the same system with every 64-row slice's entries stored (no translation
                 sharing), packed reader: the HBM-bound form of the step
  fp64_plain     the FP64 plain path (reference kernel structure, CSR with 3 FP64 values per
                 nnz as the reference stores diagonal exchange, rows stored by 64-row slices)
                 on the 1M-spin lattice
  cpu_baseline   the FP64 oracle on one host core over a bounded sample
  serving        the serving stream: 4,096 Poisson requests at 200/s of simple-cubic M(T)
                 and field-ramp solves, L log-uniform in [10, 100], 2,000-20,000 Heun steps,
                 through the batched FP32 request loop (solves/s, p50/p99 latency), a
                 sub-capacity stream (256 requests at 10/s), and the sequential FP64 plain
                 solves of a seeded sample of the stream (every 256th request, run back to back)

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--mode headline|copy-probe|traffic-probe|plain-probe|accuracy|rk4|batched|serve|fft-exchange]
  python bench.py --summarise trace|pmc <rocprofv3 results.db> [kernel ...]
"""
import argparse
import glob
import json
import os
import platform
import sqlite3
import subprocess
import sys
import tempfile
import time
from collections import defaultdict

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-atomistic-code_amd"))

PEAK_HBM_GBS = 8000.0
METRIC = "spin-updates/s @1M spins FP32 vs FP64 plain; batched solves/s + p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# ---------------------------------------------------------------- rocprofv3 results


def rp_trace(db):
    """Kernel trace of a rocprofv3 results database: (name, calls, mean, min, max us)."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration) from kernels "
                       "group by name order by sum(duration) desc").fetchall()
    return [(name, cnt, avg / 1e3, mn / 1e3, mx / 1e3) for name, cnt, avg, mn, mx in rows]


def rp_pmc(db, kernels=None):
    """Mean counter value per dispatch: {(kernel name, counter): (mean, dispatches)}."""
    con = sqlite3.connect(db)
    acc = defaultdict(list)
    for name, counter, value in con.execute("select kernel_name, counter_name, value from counters_collection"):
        if kernels and not any(k in name for k in kernels):
            continue
        acc[(name, counter)].append(value)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def summarise(argv):
    mode, db = argv[0], argv[1]
    if mode == "trace":
        print(f"{'kernel':60s} {'calls':>6s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s}")
        for name, cnt, avg, mn, mx in rp_trace(db):
            print(f"{name[:60]:60s} {cnt:6d} {avg:10.3f} {mn:10.3f} {mx:10.3f}")
    else:
        # full names: template arguments tell a kernel's instances apart
        for (name, counter), (mean, cnt) in sorted(rp_pmc(db, argv[2:]).items()):
            print(f"{name} {counter} dispatches={cnt} mean={mean:.1f}")


# ---------------------------------------------------------------- timing helpers


def time_steps(ctx, torch, t0, steps):
    """K steps bracketed by device synchronisation; HIP events on the handle's stream."""
    stream = torch.cuda.ExternalStream(ctx.stream_ptr())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    ctx.step(t0, steps)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    return e0.elapsed_time(e1) / 1e3, wall


def cpu_baseline(L, budget_s=15.0):
    """FP64 oracle (single-thread restatement of the reference CPU step) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from sllg import workloads as W

    sysd, s0 = W.fcc_3shell(L=L, precision="fp64")
    o = pyoracle.OracleContext(sysd)
    o.set_spins(s0)
    o.set_temperature(300.0)
    o.step(0, 1)  # warm-up
    steps = 0
    t = time.perf_counter()
    while time.perf_counter() - t < budget_s:
        o.step(1 + steps, 1)
        steps += 1
    dt = time.perf_counter() - t
    o.close()
    return {"value": sysd.n * steps / dt, "unit": "spin-updates/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle Heun, same fcc lattice ({sysd.n} spins, 42 nnz/row), T=300 K, "
                      f"{steps} steps after 1 warm-up step ({dt:.1f} s) on one core of {cpu_model()}"}


LAUNCH_KEYS = ("tpb", "xcd", "unroll", "thermal_cache", "gather_copy", "line_mode")


def launch_of(tune):
    return [int(tune[k]) for k in LAUNCH_KEYS]


# ---------------------------------------------------------------- HBM traffic (PMC child runs)


def traffic_probe(args):
    """Child of the PMC runs: the headline system at the given launch shape for a few steps,
    then a 1 GiB device copy (the counter scale)."""
    import torch
    import sllg
    from sllg import workloads as W

    sysd, s0 = W.fcc_3shell(L=args.L, precision="fp32")
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(args.T)
    g.set_launch(*args.launch)
    g.step(0, 50)  # the clocks settle before the profiled steps
    g.step(50, 100)
    g.sync()
    g.close()
    nbytes = 1 << 30
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)


def plain_probe(args):
    """The FP64 plain path on the headline lattice (CSR with 3 FP64 values per entry) for 20
    steps after 3, for a kernel trace of its SpMV and integrator kernels."""
    import torch
    import sllg
    from sllg import workloads as W

    sysp, sp0 = W.fcc_3shell(L=args.L, precision="fp64", ncomp=3)
    p = sllg.Context(sysp)
    p.set_spins(sp0)
    p.set_temperature(args.T)
    p.step(0, 3)
    p.step(3, 20)
    p.sync()
    p.close()
    torch.cuda.synchronize()
    print("plain probe done", flush=True)


def measure_traffic(args, tune):
    """HBM bytes per launch of the Heun kernels from two rocprofv3 PMC runs of traffic_probe
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads
    for the 1 GiB copy.  Returns (dict, None) or (None, reason)."""
    out = {}
    with tempfile.TemporaryDirectory(prefix="sllg_pmc_") as d:
        env = dict(os.environ, TMPDIR=d)
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            od = os.path.join(d, ctr)
            cmd = ["rocprofv3", "--pmc", ctr, "-d", od, "-o", ctr.lower(), "--", sys.executable,
                   os.path.join(ROOT, "bench.py"), "--mode", "traffic-probe", "--L", str(args.L), "--T", str(args.T),
                   "--launch", *[str(v) for v in launch_of(tune)]]
            log(f"bench: PMC run {ctr}")
            try:
                r = subprocess.run(cmd, cwd=d, env=env, capture_output=True, text=True, timeout=180)
            except (OSError, subprocess.TimeoutExpired) as e:
                return None, f"rocprofv3 {ctr}: {e}"
            dbs = glob.glob(os.path.join(od, "**", "*.db"), recursive=True)
            if "traffic probe done" not in r.stdout or not dbs:
                return None, f"rocprofv3 {ctr}: rc {r.returncode}, no results ({r.stderr[-300:]})"
            vals = defaultdict(list)  # per dispatch
            for db in dbs:
                con = sqlite3.connect(db)
                for name, counter, value in con.execute(
                        "select kernel_name, counter_name, value from counters_collection"):
                    if counter != ctr:
                        continue
                    key = ("copy" if "copyBuffer" in name else "corr" if "k_heun_corr" in name
                           else "pred" if "k_heun_pred" in name else None)
                    if key:
                        vals[key].append(value)
            if not all(vals.get(k) for k in ("copy", "corr", "pred")):
                return None, f"{ctr}: kernels missing from the PMC results ({sorted(vals)})"
            # the 1 GiB copies are the largest copy dispatches (uploads are copies too)
            scale = (1 << 30) / (max(vals["copy"]) * 1024.0)
            out[ctr] = {"scale": scale, "corr_KiB": float(np.mean(vals["corr"])), "pred_KiB": float(np.mean(vals["pred"]))}
    f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
    res = {"fetch_scale": f["scale"], "write_scale": w["scale"]}
    for k in ("corr", "pred"):
        res[k] = {"fetch_bytes": f[k + "_KiB"] * 1024 * f["scale"], "write_bytes": w[k + "_KiB"] * 1024 * w["scale"]}
        res[k]["hbm_bytes"] = res[k]["fetch_bytes"] + res[k]["write_bytes"]
    return res, None


# ---------------------------------------------------------------- headline


def headline(args):
    import torch
    import sllg
    from sllg import workloads as W

    assert sllg.device_count() > 0, "no GPU visible"
    L = args.L
    T = args.T
    sysd, s0 = W.fcc_3shell(L=L, precision="fp32")
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(T)
    tune = None if args.no_tune else g.autotune(20)
    log(f"bench: launch {tune}")
    g.step(0, args.warmup)
    sec, wall = time_steps(g, torch, args.warmup, args.steps)
    value = sysd.n * args.steps / sec
    ms_step = sec * 1e3 / args.steps
    # Per-kernel launch durations from HIP events around each launch on the handle's stream
    # (they agree with a rocprofv3 kernel trace).  Both Heun kernels move the same algorithmic
    # bytes per launch; the roofline is stated for the longer one.  A measured launch carries
    # a little dispatch time of its own, so the two durations add up to more than the timed
    # step; each kernel's share of the timed step is given beside them.
    prof = g.profile_kernels(args.warmup + args.steps, 50)
    info = g.info()
    pk, ck = prof.get("heun_predictor"), prof.get("heun_corrector")
    bytes_launch = info["dominant_kernel_bytes_per_spin"] * sysd.n
    per_kernel = {}
    for key, name, kms in (("pred", "k_heun_pred", pk), ("corr", "k_heun_corr", ck)):
        if kms:
            per_kernel[key] = {"kernel": name, "launch_ms": kms,
                               "step_share_ms": ms_step * kms / (pk + ck) if (pk and ck) else None,
                               "achieved_GBs": bytes_launch / (kms / 1e3) / 1e9}
    dom_key = max(per_kernel, key=lambda k: per_kernel[k]["launch_ms"])
    dom_ms = per_kernel[dom_key]["launch_ms"]
    achieved = per_kernel[dom_key]["achieved_GBs"]
    # the same system at the three temperatures of the case
    by_T = {}
    t_next = args.warmup + args.steps + 50
    for TT in (0.0, 300.0, 600.0):
        g.set_temperature(TT)
        g.step(t_next, 5)
        tsec, _ = time_steps(g, torch, t_next + 5, 100)
        t_next += 105
        by_T[f"{TT:g}K"] = {"spin_updates_per_s": sysd.n * 100 / tsec, "ms_per_step": tsec * 10}
    g.close()
    log(f"bench: headline {value:.3e} su/s, {ms_step * 1e3:.2f} us/step")

    traffic = traffic_err = None
    if not args.no_pmc and tune is not None:
        traffic, traffic_err = measure_traffic(args, tune)
        log(f"bench: traffic {traffic if traffic else traffic_err}")

    # Same system with every slice's entries stored (no sharing of repeated slices): the
    # coupling stream then comes from HBM on every launch.
    unshared = None
    if not args.no_unshared:
        sysd.flags = sllg.FLAG_NO_SLICE_SHARING | sllg.FLAG_NO_LINE_MODE
        u = sllg.Context(sysd)
        u.set_spins(s0)
        u.set_temperature(T)
        if tune is not None:
            u.autotune(20)
        u.step(0, args.warmup)
        usteps = max(50, args.steps // 4)
        usec, _ = time_steps(u, torch, args.warmup, usteps)
        uprof = u.profile_kernels(args.warmup + usteps, 20)
        uinfo = u.info()
        uck = max(uprof.get("heun_predictor", 0.0), uprof.get("heun_corrector", 0.0))  # the longer launch
        u_ms = uck
        u_bytes = uinfo["dominant_kernel_bytes_per_spin"] * sysd.n
        u_ach = u_bytes / (u_ms / 1e3) / 1e9
        unshared = {"value": sysd.n * usteps / usec, "ms_per_step": usec * 1e3 / usteps,
                    "exch_encoding": uinfo["exch_encoding"], "kernel_ms": u_ms,
                    "algorithmic_bytes_per_spin": uinfo["dominant_kernel_bytes_per_spin"],
                    "achieved_GBs": u_ach, "frac": u_ach / PEAK_HBM_GBS}
        u.close()
        sysd.flags = 0

    # A lattice whose stepping state does not fit the 256 MB Infinity Cache.
    beyond = None
    if not args.no_beyond:
        bd = (192, 192, 64)
        bsys, bs0 = W.fcc_3shell(dims=bd, precision="fp32")
        bg = sllg.Context(bsys)
        bg.set_spins(bs0)
        bg.set_temperature(T)
        btune = bg.autotune(10) if tune is not None else None
        bg.step(0, 5)
        bsteps = max(20, args.steps // 4)
        bsec, _ = time_steps(bg, torch, 5, bsteps)
        bprof = bg.profile_kernels(5 + bsteps, 20)
        binfo = bg.info()
        bck = max(bprof.get("heun_predictor", 0.0), bprof.get("heun_corrector", 0.0))  # the longer launch
        b_ms = bck
        b_ach = binfo["dominant_kernel_bytes_per_spin"] * bsys.n / (b_ms / 1e3) / 1e9
        beyond = {"workload": f"fcc {bd[0]}x{bd[1]}x{bd[2]} conventional x4 = {bsys.n} spins, same couplings",
                  "stepping_state_MB": 36 * bsys.n / 1e6, "infinity_cache_MB": 256,
                  "value": bsys.n * bsteps / bsec, "ms_per_step": bsec * 1e3 / bsteps, "kernel_ms": b_ms,
                  "achieved_GBs": b_ach, "frac": b_ach / PEAK_HBM_GBS, "exch_encoding": binfo["exch_encoding"],
                  "launch_tune": btune}
        bg.close()
        log(f"bench: beyond cache {beyond['value']:.3e} su/s")

    # FP64 plain path (reference structure) on the same lattice: CSR xx, yy, zz per nnz.
    sysp, sp0 = W.fcc_3shell(L=L, precision="fp64", ncomp=3)
    p = sllg.Context(sysp)
    p.set_spins(sp0)
    p.set_temperature(T)
    p.step(0, 3)
    psteps = max(10, args.steps // 20)
    psec, _ = time_steps(p, torch, 3, psteps)
    plain = sysp.n * psteps / psec
    pprof = p.profile_kernels(3 + psteps, 5)
    p.close()

    cpu = None if args.no_cpu else cpu_baseline(L, args.cpu_budget)
    serving = None if args.no_serve else serving_block(args)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"fcc {L}x{L}x{L} conventional x4 = {sysd.n} spins, periodic, CSR 42 nnz/row "
                               f"(J1=5e-21, J2=1e-21, J3=-4e-22 J), cubic Kc=5e-24 J, lambda=0.05, dt=1e-16 s, "
                               f"T={T} K counter noise, Heun, FP32 field+integrator, FP64 reductions",
                   "n_spins": sysd.n, "exch_encoding": info["exch_encoding"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic[dom_key]["hbm_bytes"] if traffic else None,
                     "kernel": per_kernel[dom_key]["kernel"], "kernel_ms": dom_ms,
                     "algorithmic_bytes_per_spin": info["dominant_kernel_bytes_per_spin"],
                     "per_kernel": per_kernel,
                     "traffic_detail": traffic if traffic else traffic_err},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": plain, "unit": "spin-updates/s", "ms_per_step": psec * 1e3 / psteps,
                       "fp32_over_fp64_plain": value / plain, "kernels_ms": pprof},
        "kernels_ms_events": prof,
        "by_temperature": by_T,
        "launch_tune": tune,
        "unshared_encoding": unshared,
        "beyond_cache": beyond,
        "serving": serving,
        "wall_s_timed": wall,
    }
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------- serving


def config5_specs(seed, nreq, rate, lmin=10, lmax=100, smin=2000, smax=20000, types=(0, 1)):
    """Poisson arrivals at `rate`/s; L = round(10^U(log10 lmin, log10 lmax)); T ~ U(0, 1.5 Tc) with
    Tc = 1.44 J / k_B; steps ~ U{smin..smax}; the request type drawn from `types` (0 M(T) at
    fixed T, 1 field ramp from -0.5 T to +0.5 T along z, 2 spin wave with S(k) samples); a
    seed per request; <M> sampled every 100 steps."""
    rng = np.random.default_rng(seed)
    tc = 1.44 * 3e-21 / 1.38e-23
    t = np.cumsum(rng.exponential(1.0 / rate, nreq))
    specs = []
    for k in range(nreq):
        typ = int(types[int(rng.integers(len(types)))])
        L = int(round(10 ** rng.uniform(np.log10(lmin), np.log10(lmax))))
        specs.append(dict(type=typ, L=L, steps=int(rng.integers(smin, smax + 1)), update=100,
                          T=float(rng.uniform(0, 1.5 * tc)), arrival_s=float(t[k]), seed=1000 + k,
                          B0=(0.0, 0.0, -0.5) if typ == 1 else (0.0, 0.0, 0.0),
                          B1=(0.0, 0.0, 0.5) if typ == 1 else (0.0, 0.0, 0.0)))
    return specs


def _serve_warmup():
    from sllg import batch as SB

    SB.serve(SB.make_requests([dict(L=8, steps=200, update=100, T=100.0, arrival_s=0.0, seed=1),
                               dict(L=12, steps=200, update=100, T=100.0, arrival_s=0.0, seed=2)]), graph=True)


def _progress(tag):
    return lambda done, total, wall: log(f"bench: {tag} {done}/{total} solves done, {wall:.1f} s")


def _stats(st, keys=("solves_per_s", "p50_latency_s", "p99_latency_s", "mean_latency_s", "spin_updates_per_s",
                     "completed", "max_live", "wall_s")):
    return {k: st[k] for k in keys}


def serving_block(args):
    """The serving half of the metric on the serving stream (batched FP32 request loop), a
    sub-capacity stream, and the sequential FP64 plain comparator on a seeded sample."""
    import sllg
    from sllg import batch as SB

    specs = config5_specs(args.seed, args.requests, args.rate, args.lmin, args.lmax, args.smin, args.smax)
    _serve_warmup()
    _, st, _ = SB.serve(SB.make_requests(specs), budget_bytes=int(args.budget_gb * 2**30), graph=True,
                        progress=_progress("serving stream"), progress_every=256, max_live_spins=args.max_live_spins)
    mean_su = float(np.mean([s["L"] ** 3 * s["steps"] for s in specs]))
    out = {"workload": f"{args.requests} Poisson requests at {args.rate:g}/s, simple-cubic M(T) and field-ramp "
                       f"solves, L log-uniform in [{args.lmin}, {args.lmax}], {args.smin}-{args.smax} Heun steps, "
                       f"T uniform in [0, 1.5 Tc], NN J=3e-21 J, K=1e-23 J, FP32, seed {args.seed}",
           "batched_fp32": _stats(st), "mean_spin_updates_per_request": mean_su,
           "admission": f"FIFO by arrival within an HBM budget of {args.budget_gb:g} GiB and "
                        f"{args.max_live_spins} live spins (0: no spin cap)"}
    # sub-capacity: the same kind of stream at an arrival rate the device keeps up with
    sub = config5_specs(args.seed + 1, args.sub_requests, args.sub_rate, args.lmin, args.lmax, args.smin, args.smax)
    _, ss, _ = SB.serve(SB.make_requests(sub), budget_bytes=int(args.budget_gb * 2**30), graph=True,
                        progress=_progress("sub-capacity stream"), progress_every=64,
                        max_live_spins=args.max_live_spins)
    out["sub_capacity"] = {"requests": args.sub_requests, "rate_per_s": args.sub_rate, "batched_fp32": _stats(ss)}
    # sequential FP64 plain solves of a seeded sample of the stream, back to back
    if args.seq_every > 0:
        sample = [dict(specs[k], arrival_s=0.0) for k in range(0, len(specs), args.seq_every)]
        _, sq, _ = SB.serve(SB.make_requests(sample), sequential=True, precision=sllg.FP64_PLAIN,
                            progress=_progress("sequential FP64 sample"), progress_every=1)
        proj = sq["spin_updates_per_s"] / mean_su
        out["sequential_fp64_plain"] = {
            "sample": f"{len(sample)} requests (every {args.seq_every}th of the stream), run back to back",
            "sample_solves_per_s": sq["solves_per_s"], "spin_updates_per_s": sq["spin_updates_per_s"],
            "projected_stream_solves_per_s": proj,
            "note": "projected = the sample's spin-updates/s over the stream's mean spin-updates per request"}
        out["batched_over_sequential_solves_per_s"] = st["solves_per_s"] / proj
    return out


def accuracy(args):
    """FP32 fused Heun vs the FP64 plain path (bit-identical to the FP64 oracle) on the
    1M-spin fcc case at T = 0 from the same state, beside two FP64 controls that show what
    the dynamics does to a perturbation alone: the start with 1e-15 added to every x
    component, and the start rounded to FP32 (a perturbation of the size FP32 storage makes,
    ~3e-8 per component), each against the unperturbed FP64 run."""
    import sllg
    from sllg import workloads as W

    sysd, s0 = W.fcc_3shell(L=args.L, precision="fp32")
    checkpoints = [c for c in (100, 1000, 3000, 10000, 30000) if c <= args.steps]

    def run(prec, s, ncomp=1):
        sd, _ = W.fcc_3shell(L=args.L, precision=prec, ncomp=ncomp)
        g = sllg.Context(sd)
        g.set_spins(s)
        g.set_temperature(0.0)
        out, t = {}, 0
        for c in checkpoints:
            g.step(t, c - t)
            t = c
            out[c] = (g.get_spins(), g.reduce(energy=False)[0][0])
            print(f"accuracy: {prec} step {c}", file=sys.stderr, flush=True)
        g.close()
        return out

    def compare(start, s_init):
        ref = run("fp64", s_init)
        f32 = run("fp32", s_init)
        sp = s_init.copy()
        sp[0] += 1e-15
        sp /= np.linalg.norm(sp, axis=0)
        lya = run("fp64", sp)
        sr = s_init.astype(np.float32).astype(np.float64)
        sr /= np.linalg.norm(sr, axis=0)
        rnd = run("fp64", sr)
        rows = []
        for c in checkpoints:
            rows.append({"step": c,
                         "fp32_max_abs_dS": float(np.abs(f32[c][0] - ref[c][0]).max()),
                         "fp32_abs_dM": float(np.abs(f32[c][1] - ref[c][1]).max()),
                         "fp64_perturbed_max_abs_dS": float(np.abs(lya[c][0] - ref[c][0]).max()),
                         "fp64_perturbed_abs_dM": float(np.abs(lya[c][1] - ref[c][1]).max()),
                         "fp64_fp32_rounded_start_max_abs_dS": float(np.abs(rnd[c][0] - ref[c][0]).max()),
                         "fp64_fp32_rounded_start_abs_dM": float(np.abs(rnd[c][1] - ref[c][1]).max()),
                         "M_fp64": [float(x) for x in ref[c][1]]})
        print(json.dumps({"mode": "accuracy", "start": start, "n_spins": sysd.n, "T": 0.0, "rows": rows}),
              flush=True)

    if args.start in ("random", "both"):
        # uniform random start: a quench, chaotic (the FP64 control shows the growth)
        compare("random", s0)
    if args.start in ("ordered", "both"):
        # ferromagnet along z with a Gaussian tilt of 0.05 rad per component: damped spin waves
        rng = np.random.default_rng(args.seed)
        so = np.zeros_like(s0)
        so[0] = 0.05 * rng.standard_normal(sysd.n)
        so[1] = 0.05 * rng.standard_normal(sysd.n)
        so[2] = 1.0
        so /= np.linalg.norm(so, axis=0)
        compare("ordered", so)


def rk4_dmi(args):
    """bct 96x96x48 x 2 antiferromagnet, 3x3 tensors with per-bond DMI as a periodic stencil
    with per-row tensors, easy-plane anisotropy, T = 0, RK4, Neel + 1e-3 rad tilt: FP32
    spin-updates/s and the FP32-vs-FP64 trajectory error."""
    import torch
    import sllg
    from sllg import workloads as W

    out = {"mode": "rk4_dmi"}
    res = {}
    for prec in ("fp32", "fp64"):
        sysd, s0 = W.bct_afm_dmi(precision=prec)
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.step(0, 5)
        steps = args.steps if prec == "fp32" else max(20, args.steps // 10)
        sec, _ = time_steps(g, torch, 5, steps)
        out[prec] = {"spin_updates_per_s": sysd.n * steps / sec, "ms_per_step": sec * 1e3 / steps,
                     "kernels_ms": g.profile_kernels(5 + steps, 5), "info": g.info()}
        g.close()
    n_err = args.err_steps
    for prec in ("fp32", "fp64"):
        sysd, s0 = W.bct_afm_dmi(precision=prec)
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.step(0, n_err)
        res[prec] = (g.get_spins(), g.reduce(energy=False)[0])
        g.close()
    out["n_spins"] = sysd.n
    out["error_after_steps"] = n_err
    out["fp32_max_abs_dS"] = float(np.abs(res["fp32"][0] - res["fp64"][0]).max())
    out["fp32_abs_dM"] = float(np.abs(res["fp32"][1] - res["fp64"][1]).max())
    out["fp32_over_fp64"] = out["fp32"]["spin_updates_per_s"] / out["fp64"]["spin_updates_per_s"]
    print(json.dumps(out), flush=True)


def batched(args):
    """64 replicas of the 128x128x16 film (NN exchange + FFT dipolar on the 256x256x32 padded
    grid) at T_r = 900 r/63 K, Heun FP32, one batched launch per half step, chunks of 100 steps
    replayed from hipGraphs that hold the dipolar refreshes (every `update` = 100 steps); then
    the same steps as persistent chunks (one cooperative launch per 100 steps, also replayed
    from hipGraphs, with the refreshes between the launches)."""
    import sllg
    from sllg import batch as SB
    from sllg import workloads as W

    R = args.replicas
    B = SB.Batch(flags=SB.BATCH_GRAPH)
    t = time.perf_counter()
    ids = []
    for r in range(R):
        sysd, s0 = W.sc_film_dip(Lx=args.film[0], Ly=args.film[1], Lz=args.film[2], precision="fp32", seed=100 + r,
                                 stream=r)
        ids.append(B.add(sysd, s0, T=900.0 * r / max(R - 1, 1)))
    setup = time.perf_counter() - t
    B.step(10)
    B.reduce(ids)
    tune = B.autotune(10) if not args.no_tune else None
    steps = args.steps
    t = time.perf_counter()
    done = 0
    while done < steps:
        k = min(100, steps - done)
        B.step(k)
        done += k
    m = B.reduce(ids)  # synchronises
    sec = time.perf_counter() - t
    info = B.info()
    # the same number of steps again as persistent chunks (one cooperative launch per chunk of
    # 100 steps, grid barriers between the half steps, refreshes between the chunks; the batch
    # replays graphs, so each chunk is a captured hipGraph)
    B.set_persistent(True)
    B.step(100)
    B.reduce(ids)
    t = time.perf_counter()
    done = 0
    while done < steps:
        k = min(100, steps - done)
        B.step(k)
        done += k
    B.reduce(ids)
    sec_p = time.perf_counter() - t
    pgrid = B.info()["persist_grid"]
    print(json.dumps({"mode": "batched", "replicas": R, "spins_total": info["spins"], "steps": steps,
                      "spin_updates_per_s": info["spins"] * steps / sec, "replica_steps_per_s": R * steps / sec,
                      "ms_per_step": sec * 1e3 / steps, "setup_s": setup, "device_bytes": info["device_bytes"],
                      "launch_tune": tune,
                      "persistent": {"ms_per_step": sec_p * 1e3 / steps,
                                     "spin_updates_per_s": info["spins"] * steps / sec_p, "grid": pgrid,
                                     "tiles": info["tiles"]},
                      "m_z_lowest_T": float(m[0][2]), "m_z_highest_T": float(m[-1][2])}), flush=True)


def serve_mode(args):
    """A request stream (config5_specs with the three request types, spin-wave requests
    sampling S(k)) through the batched FP32 loop; optionally the sequential FP64 plain solves
    of the same requests in arrival order."""
    import sllg
    from sllg import batch as SB

    specs = config5_specs(args.seed, args.requests, args.rate, args.lmin, args.lmax, args.smin, args.smax,
                          types=(0, 1, 2))
    reqs = SB.make_requests(specs)
    _serve_warmup()
    _, st, _ = SB.serve(reqs, budget_bytes=int(args.budget_gb * 2**30), graph=True,
                        progress=_progress("stream"), progress_every=64, max_live_spins=args.max_live_spins)
    out = {"mode": "serve", "requests": args.requests, "rate_per_s": args.rate, "batched_fp32": _stats(st)}
    if args.sequential:
        _, sq, _ = SB.serve(reqs, sequential=True, precision=sllg.FP64_PLAIN, progress=_progress("sequential"),
                            progress_every=16)
        out["sequential_fp64_plain"] = _stats(sq)
        out["solves_per_s_ratio"] = st["solves_per_s"] / sq["solves_per_s"]
    print(json.dumps(out), flush=True)


def fft_exchange(args):
    """FFT exchange (the reference's Exchange_method "fft"): open 128x128x16 sc film, NN
    exchange and dipolar in one species-resolved convolution evaluated every stage, FP32
    Heun, against the same physics as CSR exchange plus the dipolar convolution refreshed
    every stage; spin-updates/s of both and their trajectory difference at T = 0."""
    import sllg
    from sllg import workloads as W

    out = {"mode": "fft_exchange"}
    res = {}
    for name, csr in (("fft_exchange", False), ("csr_plus_dipolar", True)):
        sysd, s0 = W.sc_film_fftx(Lx=args.film[0], Ly=args.film[1], Lz=args.film[2], precision="fp32", as_csr=csr)
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.set_temperature(0.0)
        g.step(0, 20)
        g.sync()
        t = time.perf_counter()
        g.step(20, args.steps)
        g.sync()
        sec = time.perf_counter() - t
        res[name] = g.get_spins()
        out[name] = {"spin_updates_per_s": sysd.n * args.steps / sec, "ms_per_step": sec * 1e3 / args.steps,
                     "kernels_ms": g.profile_kernels(20 + args.steps, 10)}
        g.close()
    out["n_spins"] = int(np.prod(args.film))
    out["max_abs_dS_after_steps"] = float(np.abs(res["fft_exchange"] - res["csr_plus_dipolar"]).max())
    print(json.dumps(out), flush=True)


def copy_probe(args):
    """Device copy of a known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch

    nbytes = 1 << 30
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--mode", default="headline",
                    choices=["headline", "copy-probe", "traffic-probe", "plain-probe", "accuracy", "rk4", "batched",
                             "serve", "fft-exchange"])
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--T", type=float, default=300.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="headline, batched: skip the launch-shape autotuner")
    ap.add_argument("--no-unshared", action="store_true",
                    help="headline: skip the comparison run with repeated slices stored separately")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--err-steps", type=int, default=2000, help="rk4: steps for the FP32-vs-FP64 error")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--film", type=int, nargs=3, default=[128, 128, 16])
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--sub-requests", type=int, default=256, help="headline: requests of the sub-capacity stream")
    ap.add_argument("--sub-rate", type=float, default=10.0, help="headline: arrival rate of the sub-capacity stream")
    ap.add_argument("--seq-every", type=int, default=256,
                    help="headline: the sequential FP64 plain sample takes every N-th request (0: none)")
    ap.add_argument("--no-pmc", action="store_true", help="headline: skip the rocprofv3 PMC traffic runs")
    ap.add_argument("--no-beyond", action="store_true", help="headline: skip the lattice beyond the Infinity Cache")
    ap.add_argument("--launch", type=int, nargs=6, default=None, help="traffic-probe: launch shape")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--lmin", type=int, default=10)
    ap.add_argument("--lmax", type=int, default=100)
    ap.add_argument("--smin", type=int, default=2000)
    ap.add_argument("--smax", type=int, default=20000)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-live-spins", type=int, default=1 << 24,
                    help="serving: FIFO admission holds the live requests' spins below this (0: no cap)")
    ap.add_argument("--sequential", action="store_true", help="serve: also run the sequential FP64 plain comparator")
    ap.add_argument("--no-serve", action="store_true", help="headline: skip the serving measurement")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--start", default="both", choices=["random", "ordered", "both"], help="accuracy: initial state")
    if len(sys.argv) > 1 and sys.argv[1] == "--summarise":
        summarise(sys.argv[2:])
        return
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(replicas and requests would be placed whole per GPU, one large system split into slabs "
              "with a halo exchange per field evaluation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    modes = {"copy-probe": copy_probe, "traffic-probe": traffic_probe, "plain-probe": plain_probe,
             "accuracy": accuracy, "rk4": rk4_dmi,
             "batched": batched, "serve": serve_mode, "headline": headline, "fft-exchange": fft_exchange}
    modes[args.mode](args)


if __name__ == "__main__":
    main()
