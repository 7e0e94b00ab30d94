"""GPU tests of batches (many systems per launch) and of the continuous-batching loop.

Bars: a member's trajectory is bit-identical whether it runs alone or with other members
(same kernel, no cross-member data); batch vs single-handle kernels agree to 1e-6 (FP32,
same arithmetic in two kernels); graph replay is bit-identical to eager launches."""
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


def test_serve_sequential_plain_matches_physics(gpu):
    """The sequential FP64 plain comparator and the batched FP32 loop give the same
    equilibrium <M> within statistical error (ordered start, T well below Tc)."""
    specs = [dict(L=10, steps=2000, update=100, T=50.0, seed=21, arrival_s=0.0)]
    reqs = SB.make_requests(specs)
    fp32, _, _ = SB.serve(reqs)
    fp64, st, _ = SB.serve(reqs, sequential=True, precision=sllg.FP64_PLAIN)
    assert st["completed"] == 1
    assert abs(np.linalg.norm(fp32[0]["m_mean"]) - np.linalg.norm(fp64[0]["m_mean"])) < 0.02
