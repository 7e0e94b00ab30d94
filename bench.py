"""Benchmark of the stochastic-LLG time step on one MI355X.

Default workload (the 1,048,576-spin case): fcc 64x64x64 conventional cells x 4 atoms,
periodic, CSR exchange to the 3rd shell (42 nnz/row), cubic anisotropy, Heun with the
FP32 fused field assembly + integrator, T = 300 K counter-based thermal noise, spins
random from seed 42.  A "step" is one full Heun step (predictor + corrector) of all
spins; value = spin-updates/s = n * K / (timed seconds) with the state resident in HBM.

Also measured in the same run: the FP64 plain path (reference kernel structure, CSR with
3 FP64 values per nnz as the reference stores diagonal exchange) on the same lattice,
the per-kernel mean launch time from HIP events on the handle's stream (roofline), the
FP64 oracle on one host core over a bounded sample (cpu_baseline), and the serving half of
the metric: 96 M(T) / field-ramp / spin-wave requests (L 10-40, 1000-5000 steps, Poisson
arrivals at 200/s) through the batched FP32 request loop, solves/s and p50/p99 latency,
against the sequential FP64 plain solves of the same requests (`serving`).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode headline|copy-probe|accuracy|rk4|batched|serve|fft-exchange]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-atomistic-code_amd"))

PEAK_HBM_GBS = 8000.0
METRIC = "spin-updates/s @1M spins FP32 vs FP64 plain; batched solves/s + p99 latency"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_heun_fp32_fcc1M.json")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


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
    g.step(0, args.warmup)
    sec, wall = time_steps(g, torch, args.warmup, args.steps)
    value = sysd.n * args.steps / sec
    prof = g.profile_kernels(args.warmup + args.steps, 20)
    info = g.info()
    # dominant kernel: the corrector (same algorithmic bytes as the predictor)
    dom_ms = prof.get("heun_corrector", max(prof.values()))
    bytes_launch = info["dominant_kernel_bytes_per_spin"] * sysd.n
    achieved = bytes_launch / (dom_ms / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    g.close()

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
        u_ms = uprof.get("heun_corrector", max(uprof.values()))
        u_bytes = uinfo["dominant_kernel_bytes_per_spin"] * sysd.n
        u_ach = u_bytes / (u_ms / 1e3) / 1e9
        unshared = {"value": sysd.n * usteps / usec, "ms_per_step": usec * 1e3 / usteps,
                    "exch_encoding": uinfo["exch_encoding"], "kernel_ms": u_ms,
                    "algorithmic_bytes_per_spin": uinfo["dominant_kernel_bytes_per_spin"],
                    "achieved_GBs": u_ach, "frac": u_ach / PEAK_HBM_GBS}
        u.close()
        sysd.flags = 0

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
    serving = None
    if not args.no_serve:
        # the metric's serving half: 96 requests (L 10-40, 1000-5000 steps, 200/s), batched
        # FP32 loop vs the sequential FP64 plain solves of the same requests
        reqs = serve_requests(args.seed, 96, 200.0, 10, 40, 1000, 5000)
        srv = serve_run(reqs, args.budget_gb, True)
        b, q = srv["batched_fp32"], srv["sequential_fp64_plain"]
        serving = {"requests": 96, "rate_per_s": 200.0, "L": [10, 40], "steps": [1000, 5000],
                   "types": "M(T), field ramp, spin wave", "batched_solves_per_s": b["solves_per_s"],
                   "p50_latency_s": b["p50_latency_s"], "p99_latency_s": b["p99_latency_s"],
                   "sequential_fp64_plain_solves_per_s": q["solves_per_s"],
                   "sequential_p99_latency_s": q["p99_latency_s"], "solves_per_s_ratio": srv["solves_per_s_ratio"]}
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
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
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": "k_heun_corr", "kernel_ms": dom_ms,
                     "algorithmic_bytes_per_spin": info["dominant_kernel_bytes_per_spin"]},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": plain, "unit": "spin-updates/s", "ms_per_step": psec * 1e3 / psteps,
                       "fp32_over_fp64_plain": value / plain, "kernels_ms": pprof},
        "kernels_ms": prof,
        "launch_tune": tune,
        "unshared_encoding": unshared,
        "serving": serving,
        "wall_s_timed": wall,
    }
    print(json.dumps(line), flush=True)


def accuracy(args):
    """FP32 fused Heun vs the FP64 plain path (bit-identical to the FP64 oracle) on the
    1M-spin fcc case at T = 0 from the same random state, and FP64 vs FP64 with a 1e-15
    perturbation of one component (the Lyapunov growth the dynamics imposes)."""
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
        rows = []
        for c in checkpoints:
            rows.append({"step": c,
                         "fp32_max_abs_dS": float(np.abs(f32[c][0] - ref[c][0]).max()),
                         "fp32_abs_dM": float(np.abs(f32[c][1] - ref[c][1]).max()),
                         "fp64_perturbed_max_abs_dS": float(np.abs(lya[c][0] - ref[c][0]).max()),
                         "fp64_perturbed_abs_dM": float(np.abs(lya[c][1] - ref[c][1]).max()),
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
    grid) at T_r = 900 r/63 K, Heun FP32, one batched launch per half step (hipGraph when
    no dipolar refresh falls in the chunk), dipolar refreshed every `update` = 100 steps."""
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
    print(json.dumps({"mode": "batched", "replicas": R, "spins_total": info["spins"], "steps": steps,
                      "spin_updates_per_s": info["spins"] * steps / sec, "replica_steps_per_s": R * steps / sec,
                      "ms_per_step": sec * 1e3 / steps, "setup_s": setup, "device_bytes": info["device_bytes"],
                      "launch_tune": tune,
                      "m_z_lowest_T": float(m[0][2]), "m_z_highest_T": float(m[-1][2])}), flush=True)


def serve_requests(seed, nreq, rate, lmin, lmax, smin, smax):
    """Poisson stream of simple-cubic M(T), field-ramp and spin-wave requests (one third
    each; spin-wave requests sample S(k) at (q, 0, 0) every `update` steps), L log-uniform in
    [lmin, lmax], T uniform in [0, 1.5 Tc], steps uniform in [smin, smax] (whole hundreds)."""
    import sllg
    from sllg import batch as SB

    rng = np.random.default_rng(seed)
    Tc = 1.44 * 3e-21 / sllg.K_B
    specs = []
    t = 0.0
    for k in range(nreq):
        t += rng.exponential(1.0 / rate)
        L = int(round(10 ** rng.uniform(np.log10(lmin), np.log10(lmax))))
        typ = (SB.REQ_MT, SB.REQ_RAMP, SB.REQ_SW)[int(rng.integers(3))]
        specs.append(dict(type=typ, L=L, steps=int(rng.integers(smin, smax + 1)) // 100 * 100,
                          update=100, T=float(rng.uniform(0, 1.5 * Tc)), arrival_s=t, seed=1000 + k,
                          B0=(0, 0, -0.5) if typ == SB.REQ_RAMP else (0, 0, 0), B1=(0, 0, 0.5) if typ == SB.REQ_RAMP else (0, 0, 0)))
    return SB.make_requests(specs)


def serve_run(reqs, budget_gb, sequential):
    """Batched FP32 request loop (graph-captured chunks) after an untimed warm-up (the first
    serve call in a process pays module loading and first launches); optionally the
    sequential FP64 plain solves of the same requests."""
    import sllg
    from sllg import batch as SB

    SB.serve(SB.make_requests([dict(L=8, steps=200, update=100, T=100.0, arrival_s=0.0, seed=1),
                               dict(L=12, steps=200, update=100, T=100.0, arrival_s=0.0, seed=2)]), graph=True)
    _, st, _ = SB.serve(reqs, budget_bytes=int(budget_gb * 2**30), graph=True)
    out = {"batched_fp32": st}
    if sequential:
        _, sq, _ = SB.serve(reqs, sequential=True, precision=sllg.FP64_PLAIN)
        out["sequential_fp64_plain"] = sq
        out["solves_per_s_ratio"] = st["solves_per_s"] / sq["solves_per_s"]
    return out


def serve_mode(args):
    """Request stream of serve_requests: batched FP32 loop vs sequential FP64 plain solves,
    p50/p99 latency."""
    reqs = serve_requests(args.seed, args.requests, args.rate, args.lmin, args.lmax, args.smin, args.smax)
    out = {"mode": "serve", "requests": args.requests, "rate_per_s": args.rate}
    out.update(serve_run(reqs, args.budget_gb, args.sequential))
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
                    choices=["headline", "copy-probe", "accuracy", "rk4", "batched", "serve", "fft-exchange"])
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
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--lmin", type=int, default=10)
    ap.add_argument("--lmax", type=int, default=100)
    ap.add_argument("--smin", type=int, default=2000)
    ap.add_argument("--smax", type=int, default=20000)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--sequential", action="store_true", help="serve: also run the sequential FP64 plain comparator")
    ap.add_argument("--no-serve", action="store_true", help="headline: skip the serving measurement")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--start", default="both", choices=["random", "ordered", "both"], help="accuracy: initial state")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(replicas and requests would be placed whole per GPU, one large system split into slabs "
              "with a halo exchange per field evaluation); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    modes = {"copy-probe": copy_probe, "accuracy": accuracy, "rk4": rk4_dmi, "batched": batched,
             "serve": serve_mode, "headline": headline, "fft-exchange": fft_exchange}
    modes[args.mode](args)


if __name__ == "__main__":
    main()
