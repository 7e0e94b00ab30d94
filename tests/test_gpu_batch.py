"""GPU tests of batches (many systems per launch) and of the continuous-batching loop.

Bars: a member's trajectory is bit-identical whether it runs alone or with other members
(same kernel, no cross-member data); batch vs single-handle kernels agree to 1e-6 (FP32,
same arithmetic in two kernels); graph replay is bit-identical to eager launches."""
import math

import numpy as np
import pytest

import sllg
from sllg import batch as SB
from sllg import workloads as W

pytestmark = pytest.mark.gpu


def _members():
    a, sa, _ = W.sc_ferro(L=8, precision="fp32", seed=11)
    b, sb, _ = W.sc_ferro(L=12, precision="fp32", seed=12)
    c, sc = W.fcc_3shell(L=4, precision="fp32", seed=13)
    return [(a, sa, 300.0, (0, 0, 0.5)), (b, sb, 100.0, (0.2, 0, 0)), (c, sc, 0.0, (0, 0, 0))]


def test_batch_member_equals_single_handle(gpu):
    ms = _members()
    B = SB.Batch()
    ids = [B.add(s, s0, T, Bf) for s, s0, T, Bf in ms]
    B.step(60)
    B.step(40)
    for (s, s0, T, Bf), k in zip(ms, ids):
        g = sllg.Context(s)
        g.set_spins(s0)
        g.set_temperature(T)
        g.set_applied(np.asarray(Bf, float))
        g.step(0, 100)
        d = float(np.abs(B.get_spins(k) - g.get_spins()).max())
        print("member", k, "n", s.n, "max|dS| batch vs single", d)
        assert d <= 1e-6
        assert B.get_step(k) == 100


def test_batch_line_members_equal_single_handle(gpu):
    """Members whose couplings use the line reader (64-cell lines, 16-cell lines with packed
    boundary slices, open 16-cell lines) give the single handle's trajectory in a batch."""
    mem = []
    a, sa, _ = W.sc_ferro(precision="fp32", dims=(2, 2, 64), seed=21)
    b, sb, _ = W.sc_ferro(precision="fp32", dims=(2, 32, 16), seed=22)
    c, sc, _ = W.sc_ferro(precision="fp32", dims=(2, 32, 16), pbc=(1, 1, 0), seed=23)  # open lines
    for s, s0, T in ((a, sa, 200.0), (b, sb, 300.0), (c, sc, 100.0)):
        mem.append((s, s0, T, (0.0, 0.0, 0.3)))
    bt = SB.Batch()
    ids = [bt.add(s, s0, T, B) for s, s0, T, B in mem]
    bt.step(25)
    for k, (s, s0, T, B) in zip(ids, mem):
        g = sllg.Context(s)
        assert "line" in g.info()["exch_encoding"]
        g.set_spins(s0)
        g.set_temperature(T)
        g.set_applied(np.array(B))
        g.step(0, 25)
        assert float(np.abs(bt.get_spins(k) - g.get_spins()).max()) <= 1e-6
        g.close()
    bt.close()


def test_batch_alone_vs_mix_bitwise(gpu):
    ms = _members()
    solo = SB.Batch()
    s, s0, T, Bf = ms[0]
    k0 = solo.add(s, s0, T, Bf)
    solo.step(80)
    mix = SB.Batch()
    others = [mix.add(x, x0, xT, xB) for x, x0, xT, xB in ms[1:]]
    k1 = mix.add(s, s0, T, Bf)
    mix.step(30)
    mix.remove(others[0])          # a neighbour leaves mid-run
    mix.step(50)
    assert np.array_equal(solo.get_spins(k0), mix.get_spins(k1))


def test_batch_graph_replay_equals_eager(gpu):
    ms = _members()
    res = []
    for flags in (0, SB.BATCH_GRAPH):
        B = SB.Batch(flags=flags)
        ids = [B.add(s, s0, T, Bf) for s, s0, T, Bf in ms[:2]]
        for _ in range(3):
            B.step(25)
        late = B.add(*ms[2][:2], ms[2][2], ms[2][3], t0=0)   # joins after 75 steps of the others
        for _ in range(2):
            B.step(25)
        res.append([B.get_spins(k) for k in ids + [late]])
        if flags:
            assert B.info()["graphs_built"] >= 2
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def _persist_members():
    """Members of four reader groups: packed entries (sc 8^3, sc 12^3, fcc 4^3 x 4), 64-cell
    lines (2 x 2 x 64) and a 3x3-tensor stencil with DMI (bct 8 x 8 x 4 x 2)."""
    a, sa, _ = W.sc_ferro(precision="fp32", dims=(2, 2, 64), seed=21)
    d, sd = W.bct_afm_dmi(Lx=8, Ly=8, Lz=4, precision="fp32", scheme="heun")
    return _members() + [(a, sa, 200.0, (0.0, 0.0, 0.3)), (d, sd, 50.0, (0.0, 0.0, 0.0))]


@pytest.mark.parametrize("shape", [(256, 1, 1, 0), (128, 0, 1, 1), (512, 1, 0, 1)])
def test_batch_persistent_equals_per_half_step(gpu, shape):
    """Persistent chunks (one cooperative launch per reader group, a grid barrier between the
    half steps, each block a fixed set of tiles) give the per-half-step launches' trajectories
    bit for bit: uneven members of four reader groups, one joining late, chunks of 17, 30 and 1
    steps, for three launch shapes (threads per tile, XCD order, line reader, gather copy)."""
    res = []
    for persist in (False, True):
        B = SB.Batch()
        B.set_launch(*shape)
        if persist:
            B.set_persistent(True)
        ms = _persist_members()
        ids = [B.add(s, s0, T, Bf) for s, s0, T, Bf in ms[:-1]]
        B.step(17)
        late = B.add(*ms[-1])
        B.step(30)
        B.step(1)
        info = B.info()
        assert (info["persist_grid"] > 0) == persist
        for k in ids:
            assert B.get_step(k) == 48
        assert B.get_step(late) == 31
        res.append([B.get_spins(k) for k in ids + [late]])
        B.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_batch_persistent_dipolar_segments(gpu):
    """Dipolar members refreshing every 3 and every 5 steps (at the spins of the step's start):
    the persistent chunks split at each refresh and give the per-half-step launches' bits.  A
    batch with a member that also refreshes at the predicted spins (every stage) takes the
    per-half-step launches, with the same bits."""
    for every_stage in (0, 1):
        res = []
        for persist in (False, True):
            B = SB.Batch()
            if persist:
                B.set_persistent(True)
            s1, x1 = W.sc_film_dip(Lx=8, Ly=8, Lz=2, precision="fp32", refresh=3)
            s2, x2 = W.sc_film_dip(Lx=16, Ly=8, Lz=2, precision="fp32", refresh=5, seed=4,
                                   every_stage=every_stage)
            a, sa, _ = W.sc_ferro(L=8, precision="fp32", seed=11)
            ids = [B.add(s1, x1, 50.0), B.add(s2, x2, 100.0), B.add(a, sa, 300.0)]
            B.step(7)
            B.step(11)
            assert (B.info()["persist_grid"] > 0) == (persist and not every_stage)
            res.append([B.get_spins(k) for k in ids])
            B.close()
        for a, b in zip(*res):
            assert np.array_equal(a, b)


def test_batch_persistent_graph_capture(gpu):
    """Persistent chunks captured into hipGraphs (LLG_BATCH_GRAPH with persistent mode: barrier
    resets, cooperative launches and the dipolar refreshes between them; the residency probe
    runs before the capture) replay the per-half-step launches' bits: members of four reader
    groups and a dipolar film refreshing every 5 steps, three chunks of 20 steps and one of 7
    (a second graph).  If the runtime refuses to capture the cooperative launch, the batch
    launches it directly, with the same bits."""
    res = []
    for mode in ("eager", "graph_persist"):
        B = SB.Batch(flags=SB.BATCH_GRAPH if mode == "graph_persist" else 0)
        if mode == "graph_persist":
            B.set_persistent(True)
        ms = _persist_members()
        f, f0 = W.sc_film_dip(Lx=8, Ly=8, Lz=2, precision="fp32", refresh=5)
        ms.append((f, f0, 75.0, (0.0, 0.0, 0.0)))
        ids = [B.add(s, s0, T, Bf) for s, s0, T, Bf in ms]
        for n in (20, 20, 20, 7):
            B.step(n)
        info = B.info()
        if mode == "graph_persist":
            print("persistent graphs built:", info["graphs_built"], "grid", info["persist_grid"])
            assert info["persist_grid"] > 0
        res.append([B.get_spins(k) for k in ids])
        B.close()
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_batch_persistent_barrier_wait_is_bounded(gpu):
    """The grid barrier's failure path, forced from the host: with a poll limit of 0 the first
    block at a barrier that is not yet complete abandons the launch and every other block
    returns at its next wait, so the grid drains; llg_batch_step reports LLG_E_DEVICE and the
    step counter stays.  With the default limit the same batch then steps normally."""
    s, s0, _ = W.sc_ferro(L=32, precision="fp32", seed=5)  # 32,768 spins: 128 tiles of 256
    B = SB.Batch()
    k = B.add(s, s0, 300.0)
    B.set_persistent(True, spin_limit=0)
    with pytest.raises(sllg.LLGError) as e:
        B.step(10)
    assert e.value.code == -3
    assert B.get_step(k) == 0
    B.set_persistent(True)
    B.step(10)
    assert B.get_step(k) == 10
    # the same through a captured chunk: a new poll limit takes effect on the next chunk
    G = SB.Batch(flags=SB.BATCH_GRAPH)
    kg = G.add(s, s0, 300.0)
    G.set_persistent(True)
    G.step(10)
    G.set_persistent(True, spin_limit=0)
    with pytest.raises(sllg.LLGError) as e:
        G.step(10)
    assert e.value.code == -3
    G.set_persistent(True)
    G.step(10)
    assert G.get_step(kg) == 20
    sp = B.get_spins(k)
    assert np.isfinite(sp).all()
    assert float(np.abs(np.linalg.norm(sp, axis=0) - 1.0).max()) <= 2e-6


def test_batch_budget_admission(gpu):
    """A member is admitted when the bytes held plus its upper-bound estimate fit the budget;
    otherwise LLG_E_NOMEM and nothing is allocated.  (The bytes held can be well below the
    estimate: repeated slices of the encoding are stored once.)"""
    s, s0, _ = W.sc_ferro(L=16, precision="fp32")
    est = SB.estimate_bytes(s)
    B = SB.Batch(budget_bytes=est)
    B.add(s, s0)
    used1 = B.info()["device_bytes"]
    assert 0 < used1 <= est
    with pytest.raises(sllg.LLGError) as e:
        B.add(s, s0)
    assert e.value.code == -2
    assert B.info()["members"] == 1 and B.info()["device_bytes"] == used1
    B2 = SB.Batch(budget_bytes=used1 + est)
    B2.add(s, s0)
    B2.add(s, s0)
    with pytest.raises(sllg.LLGError) as e:
        B2.add(s, s0)
    assert e.value.code == -2
    assert B2.info()["members"] == 2


@pytest.mark.parametrize("kind", ["dipolar", "fftx", "fftx_dip"])
def test_estimate_covers_convolution_plans(gpu, kind):
    """The admission estimate bounds what a handle with a dipolar or FFT-exchange plan holds
    (plan grids, spectra, k-space kernels and FFT work areas are counted in device_bytes)."""
    if kind == "dipolar":
        s, s0 = W.sc_film_dip(Lx=16, Ly=12, Lz=4, precision="fp32", refresh=1)
    else:
        s, s0 = W.sc_film_fftx(Lx=16, Ly=12, Lz=4, precision="fp32", dipolar=kind == "fftx_dip")
    g = sllg.Context(s)
    held = g.info()["device_bytes"]
    G = 32 * 24 * 8  # padded grid points
    assert held >= 3 * G * 4  # at least one real grid of the plan is counted
    assert held <= SB.estimate_bytes(s)


def test_batch_dipolar_member(gpu):
    s, s0 = W.sc_film_dip(Lx=8, Ly=8, Lz=2, precision="fp32", refresh=5)
    B = SB.Batch()
    k = B.add(s, s0, 50.0)
    B.step(20)
    g = sllg.Context(s)
    g.set_spins(s0)
    g.set_temperature(50.0)
    g.step(0, 20)
    assert float(np.abs(B.get_spins(k) - g.get_spins()).max()) <= 1e-6


def test_serve_results_same_alone_and_in_mix(gpu):
    specs = [dict(L=8, steps=600, update=100, T=200.0, seed=5, arrival_s=0.0),
             dict(L=12, steps=400, update=100, T=400.0, seed=6, arrival_s=0.0, type=SB.REQ_RAMP,
                  B0=(0, 0, -1.0), B1=(0, 0, 1.0)),
             dict(L=10, steps=800, update=100, T=50.0, seed=7, arrival_s=0.01),
             dict(L=6, steps=300, update=100, T=600.0, seed=8, arrival_s=0.02)]
    reqs = SB.make_requests(specs)
    mix, st, ser = SB.serve(reqs, want_series=True)
    assert st["completed"] == 4 and st["max_live"] >= 2
    for k, sp in enumerate(specs):
        one = SB.make_requests([dict(sp, arrival_s=0.0)])
        alone, _, ser1 = SB.serve(one, want_series=True)
        assert np.array_equal(ser[k], ser1[0]), k
        assert alone[0]["m_final"] == mix[k]["m_final"]
    print(st)


def test_serve_graph_chunks_equal_eager(gpu):
    """The request loop with hipGraph chunks (captured again whenever members join or leave),
    and with persistent chunks (direct and graph-captured), gives the eager loop's results bit
    for bit: every sampled series and final moment."""
    specs = [dict(L=8, steps=500, update=100, T=200.0, seed=15, arrival_s=0.0),
             dict(L=10, steps=300, update=100, T=50.0, seed=16, arrival_s=0.0, type=SB.REQ_RAMP,
                  B0=(0, 0, -1.0), B1=(0, 0, 1.0)),
             dict(L=6, steps=400, update=100, T=400.0, seed=17, arrival_s=0.005),
             dict(L=8, steps=200, update=100, T=100.0, seed=18, arrival_s=0.01, type=SB.REQ_SW)]
    reqs = SB.make_requests(specs)
    out = [SB.serve(reqs, want_series=True, graph=g, persistent=p)
           for g, p in ((False, False), (True, False), (False, True), (True, True))]
    r0, s0, ser0 = out[0]
    for r1, s1, ser1 in out[1:]:
        assert s0["completed"] == s1["completed"] == len(specs)
        for k in range(len(specs)):
            assert np.array_equal(ser0[k], ser1[k]), k
            assert r0[k]["m_final"] == r1[k]["m_final"]


def test_batch_autotune_restores_and_shapes_agree(gpu):
    """The batch autotuner restores every member's spins and step counter, and every tile
    size / block order / gather copy gives bit-identical trajectories (one thread per spin,
    the same sums)."""
    res = []
    for shape in (None, (128, 0), (512, 1), (256, 1, 1, 1), (128, 0, 1, 1), "tune"):
        bt = SB.Batch(flags=SB.BATCH_GRAPH)
        ids = [bt.add(s, s0, T, B) for s, s0, T, B in _members()]
        before = [bt.get_spins(k) for k in ids]
        if shape == "tune":
            t = bt.autotune(5)
            assert t["candidates"] == 12 and t["tpb"] in (128, 256, 512) and t["gather_copy"] in (0, 1)
            for k, s0 in zip(ids, before):
                assert np.array_equal(bt.get_spins(k), s0)
                assert bt.get_step(k) == 0
        elif shape is not None:
            bt.set_launch(*shape)
        bt.step(30)
        res.append([bt.get_spins(k) for k in ids])
        bt.close()
    for r in res[1:]:
        for a, b in zip(res[0], r):
            assert np.array_equal(a, b)


def test_batch_ensemble_langevin(gpu):
    """Ensemble M(T) from one batch: replicas of n free spins in B at four temperatures
    (FP32, counter noise, hipGraph chunks).  <S_z> must equal the Langevin function
    coth x - 1/x, x = mu mu_B B / (k_B T), within five standard errors: per-spin variance
    1 - 2 L(x)/x - L(x)^2, n spins, and one independent sample per two relaxation times
    (1 + lambda^2)/(lambda gamma B) of the sampling window."""
    n, B, lam, dt = 4096, 10.0, 1.0, 2e-15
    xs = (0.5, 1.0, 2.0, 4.0)
    bt = SB.Batch(flags=SB.BATCH_GRAPH)
    ids = []
    for k, x in enumerate(xs):
        T = 2.22 * sllg.MU_B * B / (sllg.K_B * x)
        sysd = sllg.System(n=n, species=[dict(mu=2.22, lambda_=lam)], dt=dt, precision="fp32", seed=40 + k)
        s0 = np.zeros((3, n))
        s0[2] = 1.0
        ids.append(bt.add(sysd, s0, T=T, B=(0.0, 0.0, B)))
    for _ in range(60):
        bt.step(100)
    acc = []
    for _ in range(60):
        bt.step(50)
        acc.append(bt.reduce(ids)[:, 2])
    m = np.mean(np.array(acc), axis=0)
    tau = (1 + lam ** 2) / (lam * 1.76e11 * B) / dt  # relaxation time in steps
    n_t = max(1.0, 60 * 50 / (2 * tau))
    for x, mz in zip(xs, m):
        L = 1 / math.tanh(x) - 1 / x
        tol = 5 * math.sqrt(1 - 2 * L / x - L * L) / math.sqrt(n * n_t)
        print(f"x={x} <Sz>={mz:.4f} Langevin={L:.4f} tol={tol:.4f}")
        assert abs(mz - L) < tol, (x, mz, L, tol)
    bt.close()


def test_serve_spin_wave_requests(gpu):
    """Spin-wave requests sample S(k) at (q, 0, 0) beside <M>: S(0)/n is the transverse
    magnetisation of the same sample; T = 0 from the +z start gives S(k) = 0; each request's
    S(k) series is the same alone and in a mix with M(T) and ramp requests; the sequential
    FP64 plain comparator samples the same quantities."""
    specs = [dict(type=SB.REQ_SW, L=8, steps=500, update=50, T=300.0, seed=31, arrival_s=0.0),
             dict(type=SB.REQ_MT, L=10, steps=400, update=100, T=100.0, seed=32, arrival_s=0.0),
             dict(type=SB.REQ_SW, L=12, steps=300, update=100, T=0.0, seed=33, arrival_s=0.005,
                  B0=(0.0, 0.0, 0.5)),
             dict(type=SB.REQ_RAMP, L=6, steps=300, update=100, T=50.0, seed=34, arrival_s=0.01,
                  B0=(0, 0, -1.0), B1=(0, 0, 1.0))]
    reqs = SB.make_requests(specs)
    mix, st, ser = SB.serve(reqs, want_series=True)
    assert st["completed"] == 4
    sq0 = mix[0]["sq"]
    n0 = 8 ** 3
    assert sq0.shape == (10, 5)
    mt = ser[0][:, 0] + 1j * ser[0][:, 1]
    assert np.abs(sq0[:, 0] / n0 - mt).max() <= 1e-9
    assert np.abs(sq0[1:]).max() > 0  # thermal spin waves after the first sample
    assert np.abs(mix[2]["sq"]).max() == 0.0
    assert "sq" not in mix[1] and "sq" not in mix[3]
    for k in (0, 2):
        alone, _, _ = SB.serve(SB.make_requests([dict(specs[k], arrival_s=0.0)]), want_series=True)
        assert np.array_equal(alone[0]["sq"], mix[k]["sq"]), k
    seq, _, sers = SB.serve(SB.make_requests([dict(specs[0], arrival_s=0.0)]), want_series=True, sequential=True,
                            precision=sllg.FP64_PLAIN)
    assert seq[0]["sq"].shape == sq0.shape
    assert np.abs(seq[0]["sq"][:, 0] / n0 - (sers[0][:, 0] + 1j * sers[0][:, 1])).max() <= 1e-9


def test_serve_sequential_plain_matches_physics(gpu):
    """The sequential FP64 plain comparator and the batched FP32 loop give the same
    equilibrium <M> within statistical error (ordered start, T well below Tc)."""
    specs = [dict(L=10, steps=2000, update=100, T=50.0, seed=21, arrival_s=0.0)]
    reqs = SB.make_requests(specs)
    fp32, _, _ = SB.serve(reqs)
    fp64, st, _ = SB.serve(reqs, sequential=True, precision=sllg.FP64_PLAIN)
    assert st["completed"] == 1
    assert abs(np.linalg.norm(fp32[0]["m_mean"]) - np.linalg.norm(fp64[0]["m_mean"])) < 0.02
