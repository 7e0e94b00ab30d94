"""GPU parity at the five workload configurations, plus the staggered-field ramp, the
launch tuners' state handling and the failure paths.

  config 1  simple-cubic 32^3, NN CSR, FP64 plain Heun, 10^4 steps: T = 0 and T = 300 K with
            the same injected noise (chunks of 100 steps), <M> every 100 steps, vs the oracle.
  config 3  bct two-sublattice antiferromagnet, 3x3 tensors + per-bond DMI on the per-row
            periodic stencil: small lattice vs the oracle (Heun and RK4, FP64 and FP32), and
            the full 96 x 96 x 48 x 2 lattice FP32 vs FP64 (RK4).
  config 4  128 x 128 x 16 dipolar films batched at 0-900 K: every member equals its
            single-handle run; graph replay (refreshes inside the graph) equals eager launches.
  config 5  a seeded Poisson mix of 32 requests, L log-uniform in [10, 100], 2,000-20,000 steps:
            a sample of requests gives the same result alone and in the mix.

Bars (in each assertion): FP64 plain vs oracle <= 1e-10 per spin component (the two give the
same bits by construction); FP32 vs oracle or FP64 <= 1e-4 per component and 1e-5 in <M>;
batch vs single-handle kernels <= 1e-5 (same FP32 arithmetic in two kernels); alone vs mix
and graph vs eager bit-identical (same kernels, no cross-member data)."""
import math

import numpy as np
import pytest

import pyoracle
import sllg
from sllg import batch as SB
from sllg import workloads as W

from helpers import run_pair, tensor_csr_system

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _maxdiff(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


# ------------------------------------------------------------------ config 1


@pytest.mark.timeout(600)
def test_config1_sc32_fp64_T0_10k_steps(gpu):
    """sc 32^3 (32,768 spins), J = 3e-21 J, K = 1e-23 J, lambda 0.1, dt 1e-16 s, random start
    from seed 42: 10^4 FP64 Heun steps at T = 0, <M> sampled every 100 steps (101 samples
    with the final state)."""
    sysd, s0, _ = W.sc_ferro(L=32, precision="fp64")
    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    series = []
    for c in (g, o):
        c.set_spins(s0)
        c.set_temperature(0.0)
        series.append(c.run_sampled(0, 10000, 100))
    d = _maxdiff(g.get_spins(), o.get_spins())
    dm = _maxdiff(series[0], series[1])
    dmf = _maxdiff(g.reduce(energy=False)[0], o.reduce(energy=False)[0])
    print("config 1 T=0 10^4 steps: max|dS|", d, "max|d<M>(t)|", dm, "final", dmf)
    assert series[0].shape == (100, 1, 3)
    assert d <= TOL64 and dm <= TOL64 and dmf <= TOL64


@pytest.mark.timeout(600)
def test_config1_sc32_fp64_T300_injected_10k_steps(gpu):
    """Same lattice at T = 300 K (0.96 Tc): one N(0,1) array per step fed to both through
    push_noise in chunks of 100 steps; <M> sampled every 100 steps; trajectories identical."""
    sysd, s0, _ = W.sc_ferro(L=32, precision="fp64", noise="injected")
    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    for c in (g, o):
        c.set_spins(s0)
        c.set_temperature(300.0)
    rng = np.random.default_rng(4242)
    mg, mo = [], []
    for chunk in range(100):
        xi = rng.standard_normal((100, 3, sysd.n))
        g.push_noise(xi)
        o.push_noise(xi)
        mg.append(g.run_sampled(100 * chunk, 100, 100)[0])
        mo.append(o.run_sampled(100 * chunk, 100, 100)[0])
    d = _maxdiff(g.get_spins(), o.get_spins())
    dm = _maxdiff(np.array(mg), np.array(mo))
    mz = np.array(mg)[:, 0, 2]
    print("config 1 T=300 injected 10^4 steps: max|dS|", d, "max|d<M>(t)|", dm, "<Mz> last", mz[-1])
    assert d <= TOL64 and dm <= TOL64
    # near Tc the random start stays disordered: |<M>| well below 1 on both
    assert float(np.abs(np.array(mg)).max()) < 0.9


# ------------------------------------------------------------------ config 3


@pytest.mark.parametrize("scheme", ["heun", "rk4"])
def test_config3_bct_per_row_stencil_fp64_vs_oracle(gpu, scheme):
    """bct 8 x 8 x 4 x 2 (512 spins): 3x3 J_iso = -4e-21 J plus per-bond DMI 0.3e-21 J on the
    per-row periodic stencil (one tensor per spin and neighbour), easy-plane K, Neel start with
    a 1e-3 rad tilt.  The FP64 plain path gives the oracle's bits."""
    sysd, s0 = W.bct_afm_dmi(Lx=8, Ly=8, Lz=4, precision="fp64", scheme=scheme)
    assert sysd.exch["per_row"] == 1 and sysd.exch["vals"].shape == (9, 8, 512)
    g, o = run_pair(sysd, s0, 400, T=0.0, B=(0.0, 0.0, 0.2))
    d = _maxdiff(g.get_spins(), o.get_spins())
    dh = _maxdiff(g.field(), o.field())
    print("config 3 small", scheme, "FP64 400 steps max|dS|", d, "max|dH|", dh)
    assert d <= TOL64
    assert dh <= 1e-9 * float(np.abs(o.field()).max())
    assert g.info()["exch_encoding"] == "stencil-rowvals-f64"


@pytest.mark.parametrize("scheme", ["heun", "rk4"])
def test_config3_bct_per_row_stencil_fp32_vs_oracle(gpu, scheme):
    sysd, s0 = W.bct_afm_dmi(Lx=8, Ly=8, Lz=4, precision="fp32", scheme=scheme)
    g, o = run_pair(sysd, s0, 400, T=0.0, B=(0.0, 0.0, 0.2))
    d = _maxdiff(g.get_spins(), o.get_spins())
    dm = _maxdiff(g.reduce(energy=False)[0], o.reduce(energy=False)[0])
    print("config 3 small", scheme, "FP32 400 steps max|dS|", d, "|d<M>|", dm)
    assert d <= 1e-4 and dm <= 1e-5
    assert g.info()["exch_encoding"] == "stencil-rowvals-f32"


@pytest.mark.timeout(300)
def test_config3_bct_full_size_fp32_vs_fp64_rk4(gpu):
    """The full bct 96 x 96 x 48 x 2 = 884,736-spin RK4 spin-wave run, FP32 vs FP64 plain from
    the same tilted Neel state for 1,000 steps (dt = 5e-17 s): per-spin error, sublattice <M>,
    unit norms, and a staggered magnetisation that stays near 1 (the tilt only precesses)."""
    sysd, s0 = W.bct_afm_dmi(precision="fp32")
    out = []
    for prec in ("fp32", "fp64"):
        sysd.precision = prec
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.set_temperature(0.0)
        g.step(0, 1000)
        out.append((g.get_spins(), g.reduce(energy=False)[0]))
        g.close()
    d = _maxdiff(out[0][0], out[1][0])
    dm = _maxdiff(out[0][1], out[1][1])
    nrm = float(np.abs(np.linalg.norm(out[0][0], axis=0) - 1.0).max())
    neel = float(np.abs(out[1][1][0] - out[1][1][1]).max() / 2.0)
    print("config 3 full size RK4 1000 steps: FP32 vs FP64 max|dS|", d, "|d<M>|", dm, "norm", nrm, "Neel", neel)
    assert d <= 1e-4 and dm <= 1e-5 and nrm <= 2e-6
    assert neel > 0.99


# ------------------------------------------------------------------ config 4


def _film_members(nm, refresh=100):
    out = []
    for r in range(nm):
        sysd, s0 = W.sc_film_dip(Lx=128, Ly=128, Lz=16, precision="fp32", refresh=refresh, seed=3 + r, stream=r)
        out.append((sysd, s0, 900.0 * r / (nm - 1)))
    return out


@pytest.mark.timeout(300)
def test_config4_full_size_dipolar_batch_members_equal_single(gpu):
    """Eight 128 x 128 x 16 films (262,144 spins each; dipolar on the 256 x 256 x 32 grid,
    refreshed every 100 steps) at T = 0 .. 900 K in one graph-replayed batch for 150 steps
    (refreshes at steps 0 and 100 inside the captured graphs): every member equals its own
    single-handle run."""
    mem = _film_members(8)
    B = SB.Batch(flags=SB.BATCH_GRAPH)
    ids = [B.add(s, s0, T) for s, s0, T in mem]
    B.step(100)
    B.step(50)
    assert B.info()["graphs_built"] >= 1
    for (s, s0, T), k in zip(mem, ids):
        g = sllg.Context(s)
        g.set_spins(s0)
        g.set_temperature(T)
        g.step(0, 150)
        d = _maxdiff(B.get_spins(k), g.get_spins())
        print(f"film member T={T:.0f} K: max|dS| batch vs single {d:.3e}")
        assert d <= 1e-5
        g.close()
    B.close()


def test_config4_graph_with_dipolar_refresh_equals_eager(gpu):
    """Batched dipolar films (64 x 64 x 8, refresh every 10 steps; one member every stage)
    replayed from hipGraphs that contain the refreshes give the eager launches' bits, over
    chunk sizes that put the refreshes at different places (a new graph per placement)."""
    mem = []
    for r in range(3):
        s, s0 = W.sc_film_dip(Lx=64, Ly=64, Lz=8, precision="fp32", refresh=10, seed=30 + r, stream=r,
                              every_stage=1 if r == 2 else 0)
        mem.append((s, s0, 150.0 * r))
    a, _, _ = W.sc_ferro(L=10, precision="fp32", seed=9)
    runs = []
    for flags in (SB.BATCH_GRAPH, 0):
        B = SB.Batch(flags=flags)
        ids = [B.add(s, s0, T) for s, s0, T in mem]
        ids.append(B.add(a, W.sc_ferro(L=10, precision="fp32", seed=9)[1], 200.0))
        for n in (10, 10, 7, 13, 10, 30, 30):  # 110 steps
            B.step(n)
        runs.append([B.get_spins(k) for k in ids])
        if flags:
            print("graphs built", B.info()["graphs_built"])
            assert B.info()["graphs_built"] >= 3
        B.close()
    for x, y in zip(*runs):
        assert np.array_equal(x, y)


# ------------------------------------------------------------------ config 5


def _config5_requests(nreq, seed=5):
    """Poisson arrivals at 200/s; L = round(10^U(1,2)); T ~ U(0, 1.5 Tc); steps ~ U{2000..20000};
    M(T) or field-ramp solves; a seed per request."""
    rng = np.random.default_rng(seed)
    tc = 1.44 * 3e-21 / 1.38e-23
    t = np.cumsum(rng.exponential(1 / 200.0, nreq))
    specs = []
    for k in range(nreq):
        typ = int(rng.integers(0, 2))
        specs.append(dict(type=typ, L=int(round(10 ** rng.uniform(1, 2))), steps=int(rng.integers(2000, 20001)),
                          update=100, T=float(rng.uniform(0, 1.5 * tc)), arrival_s=float(t[k]), seed=1000 + k,
                          B0=(0.0, 0.0, 0.5 if typ == 0 else -1.0), B1=(0.0, 0.0, 1.0)))
    return specs


@pytest.mark.timeout(300)
def test_config5_mix_alone_vs_in_the_mix(gpu):
    """32 requests of the serving mix run batched with graph chunks; the largest request and
    three others give the same <M> series, bit for bit, when each is served alone."""
    specs = _config5_requests(32)
    reqs = SB.make_requests(specs)
    res, st, ser = SB.serve(reqs, graph=True, want_series=True)
    print("config 5 mix of 32:", {k: st[k] for k in ("solves_per_s", "p50_latency_s", "p99_latency_s",
                                                    "spin_updates_per_s", "max_live")})
    assert st["completed"] == 32 and st["max_live"] > 1
    big = int(np.argmax([s["L"] for s in specs]))
    sample = sorted({big, 0, 7, 21})
    for k in sample:
        one = dict(specs[k], arrival_s=0.0)
        r1, _, s1 = SB.serve(SB.make_requests([one]), graph=True, want_series=True)
        print(f"request {k}: L={specs[k]['L']} steps={specs[k]['steps']} T={specs[k]['T']:.0f}",
              "final <M>", r1[0]["m_final"])
        assert np.array_equal(s1[0], ser[k])
        assert r1[0]["m_final"] == res[k]["m_final"]


# ------------------------------------------------------------------ staggered ramp


def test_staggered_ramp_on_device_fp64_bit_parity(gpu):
    """llg_scale_staggered ramps H0 on the device (CincStagFields): the FP64 path keeps the
    oracle's bits through several ramp values, and the field includes r H0."""
    sysd, s0 = tensor_csr_system(L=6, precision="fp64", ncomp=9, fmt="csr")
    h0 = np.random.default_rng(8).normal(0, 20.0, (3, sysd.n))
    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    for c in (g, o):
        c.set_spins(s0)
        c.set_temperature(0.0)
        c.set_staggered(h0)
    t = 0
    for r in (1.0, 0.5, -1.3, 0.0, 2.0):
        for c in (g, o):
            c.scale_staggered(r)
            c.step(t, 40)
        t += 40
        d = _maxdiff(g.get_spins(), o.get_spins())
        assert d <= TOL64, (r, d)
        assert _maxdiff(g.field(), o.field()) <= 1e-9 * float(np.abs(o.field()).max())
    g.set_staggered(None)
    with pytest.raises(sllg.LLGError):
        g.scale_staggered(0.5)


def test_staggered_ramp_on_device_fp32_sublattice_order(gpu):
    """FP32 path with the sublattice-major device order (fcc, 4 basis atoms): the ramped
    field equals the oracle's r H0 to FP32 rounding, and ramping equals setting r H0 directly,
    bit for bit (both round r H0 once)."""
    sysd, s0 = W.fcc_3shell(L=4, precision="fp32")
    h0 = np.random.default_rng(9).normal(0, 20.0, (3, sysd.n))
    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    g2 = sllg.Context(sysd)
    for c in (g, o, g2):
        c.set_spins(s0)
        c.set_temperature(0.0)
    g.set_staggered(h0)
    o.set_staggered(h0)
    for r in (0.7, -2.0):
        g.scale_staggered(r)
        o.scale_staggered(r)
        g2.set_staggered(r * h0)
        hf, ho = g.field(), o.field()
        assert _maxdiff(hf, ho) <= 2e-6 * float(np.abs(ho).max())
        assert np.array_equal(hf, g2.field())
    for c in (g, o, g2):
        c.step(0, 50)
    assert np.array_equal(g.get_spins(), g2.get_spins())
    assert _maxdiff(g.get_spins(), o.get_spins()) <= 1e-4


# ------------------------------------------------------------------ launch tuners


def test_autotune_restores_dipolar_field(gpu):
    """Tuning at a step that is not a refresh step leaves the cached dipolar field alone: the
    tuned handle continues exactly as an untuned one (the reader is pinned, so every candidate
    gives the same bits)."""
    out = []
    for tune in (True, False):
        sysd, s0 = W.sc_film_dip(Lx=32, Ly=32, Lz=4, precision="fp32", refresh=10, seed=12)
        sysd.flags = sllg.FLAG_TUNE_FIXED_READER
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.set_temperature(100.0)
        g.step(0, 5)
        if tune:
            r = g.autotune(4)
            assert r["candidates"] > 1
        g.step(5, 25)
        out.append(g.get_spins())
        g.close()
    assert np.array_equal(out[0], out[1])


def test_batch_autotune_restores_dipolar_field_and_steps(gpu):
    out = []
    for tune in (True, False):
        B = SB.Batch(flags=SB.BATCH_GRAPH)
        s, s0 = W.sc_film_dip(Lx=32, Ly=32, Lz=4, precision="fp32", refresh=10, seed=13)
        a, a0, _ = W.sc_ferro(L=8, precision="fp32", seed=14)
        ids = [B.add(s, s0, 150.0), B.add(a, a0, 50.0)]
        B.step(5)
        if tune:
            B.autotune(4)
            assert [B.get_step(k) for k in ids] == [5, 5]
        B.step(25)
        out.append([B.get_spins(k) for k in ids])
        B.close()
    for x, y in zip(*out):
        assert np.array_equal(x, y)


# ------------------------------------------------------------------ failure paths


def test_check_finite_single_handle(gpu):
    """A field that overflows FP32 (1e300 T is finite in FP64, +inf once rounded) turns the spins
    into NaN; with LLG_FLAG_CHECK_FINITE the step reports LLG_E_NONFINITE instead of exiting."""
    sysd, s0, _ = W.sc_ferro(L=8, precision="fp32", seed=3)
    sysd.flags = sllg.FLAG_CHECK_FINITE
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.step(0, 10)  # finite: no error
    g.set_applied((1e300, 0.0, 0.0))
    with pytest.raises(sllg.LLGError) as e:
        g.step(10, 5)
    assert e.value.code == -7


def test_nonfinite_batch_member_is_isolated(gpu):
    """One member of a batch goes non-finite: the step reports LLG_E_NONFINITE, that member's
    status says so and it leaves the launches; the other members' trajectories equal those of
    a batch that never held it, bit for bit, and they keep stepping."""
    ms = []
    for k, L in enumerate((8, 10, 12)):
        s, s0, _ = W.sc_ferro(L=L, precision="fp32", seed=40 + k)
        s.flags = sllg.FLAG_CHECK_FINITE
        ms.append((s, s0, 100.0 * k))
    bad = SB.Batch(flags=SB.BATCH_GRAPH)
    ids = [bad.add(s, s0, T) for s, s0, T in ms]
    bad.set_applied(ids[1], (1e300, 0.0, 0.0))
    with pytest.raises(sllg.LLGError) as e:
        bad.step(20)
    assert e.value.code == -7
    assert [bad.member_status(k) for k in ids] == [0, -7, 0]
    bad.step(30)  # the others go on; the failed member is not stepped again
    assert bad.get_step(ids[1]) == 20 and bad.get_step(ids[0]) == 50
    good = SB.Batch(flags=SB.BATCH_GRAPH)
    gids = [good.add(s, s0, T) for (s, s0, T) in (ms[0], ms[2])]
    good.step(20)
    good.step(30)
    for k, gk in zip((ids[0], ids[2]), gids):
        assert np.array_equal(bad.get_spins(k), good.get_spins(gk))


def test_serve_live_spin_cap(gpu):
    """A cap on the live requests' spins: four 8,000-spin requests arriving together under a
    16,000-spin cap run two at a time, first come first served, and each gives the <M> series
    it gives without the cap (bit for bit)."""
    specs = [dict(L=20, steps=300, update=100, T=100.0, arrival_s=0.0, seed=80 + k) for k in range(4)]
    res_c, st_c, ser_c = SB.serve(SB.make_requests(specs), graph=True, want_series=True, max_live_spins=16000)
    res_u, st_u, ser_u = SB.serve(SB.make_requests(specs), graph=True, want_series=True)
    assert st_c["completed"] == 4 and st_c["max_live"] == 2 and st_u["max_live"] == 4
    assert res_c[2]["admit_s"] >= min(res_c[0]["done_s"], res_c[1]["done_s"]) - 1e-9
    for k in range(4):
        assert np.array_equal(ser_c[k], ser_u[k])


def test_serve_waits_for_memory_then_admits(gpu):
    """An HBM budget that holds one request at a time: all requests arrive at once, each waits
    for the previous one to leave, and all complete."""
    sysd, _, _ = W.sc_ferro(L=20, precision="fp32")
    full = SB.estimate_bytes(sysd)
    state = SB.estimate_bytes(sllg.System(n=sysd.n, species=sysd.species, precision="fp32"))
    budget = full + state // 2
    specs = [dict(L=20, steps=300, update=100, T=100.0, arrival_s=0.0, seed=70 + k) for k in range(3)]
    res, st, _ = SB.serve(SB.make_requests(specs), budget_bytes=budget)
    assert st["completed"] == 3 and st["max_live"] == 1
    order = sorted(range(3), key=lambda k: res[k]["admit_s"])
    for a, b in zip(order, order[1:]):
        assert res[b]["admit_s"] >= res[a]["done_s"] - 1e-9
    with pytest.raises(sllg.LLGError) as e:  # does not fit even alone
        SB.serve(SB.make_requests(specs[:1]), budget_bytes=state)
    assert e.value.code == -2
