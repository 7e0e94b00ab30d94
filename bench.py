"""Benchmark of the stochastic-LLG time step on one MI355X.

Default workload (the 1,048,576-spin case): fcc 64x64x64 conventional cells x 4 atoms,
periodic, CSR exchange to the 3rd shell (42 nnz/row), cubic anisotropy, Heun with the
FP32 fused field assembly + integrator, T = 300 K counter-based thermal noise, spins
random from seed 42.  A "step" is one full Heun step (predictor + corrector) of all
spins; value = spin-updates/s = n * K / (timed seconds) with the state resident in HBM.

Also measured in the same run: the FP64 plain path (reference kernel structure, CSR with
3 FP64 values per nnz as the reference stores diagonal exchange) on the same lattice,
the per-kernel mean launch time from HIP events on the handle's stream (roofline), and
the FP64 oracle on one host core over a bounded sample (cpu_baseline).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode headline|copy-probe]
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
        "wall_s_timed": wall,
    }
    print(json.dumps(line), flush=True)


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
    ap.add_argument("--mode", default="headline", choices=["headline", "copy-probe"])
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--T", type=float, default=300.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy-probe":
        copy_probe(args)
    else:
        headline(args)


if __name__ == "__main__":
    main()
