"""GPU parity: the HIP paths against the FP64 oracle, through the C ABI.

Bars (written into each assertion):
  FP64 plain path  - max |dS| <= 1e-10 per spin component (T = 0, and T > 0 with the same
                     injected noise); the two are built to give identical bits.
  FP32 path        - error against the oracle stated per test (float rounding, 2 evaluations
                     per step).
"""
import numpy as np
import pytest

import sllg
from sllg import workloads as W

from helpers import run_pair, tensor_csr_system

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _maxdiff(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def test_plain_heun_sc_T0(gpu):
    sysd, s0, _ = W.sc_ferro(L=12, precision="fp64")
    g, o = run_pair(sysd, s0, 2000, T=0.0)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("plain heun T=0 max|dS| =", d)
    assert d <= TOL64
    assert _maxdiff(g.det_field(), o.det_field()) <= 1e-6  # tesla, |H| ~ 1e3


def test_plain_heun_sc_injected_T300(gpu):
    sysd, s0, _ = W.sc_ferro(L=10, precision="fp64", noise="injected")
    rng = np.random.default_rng(1)
    xi = rng.standard_normal((300, 3, sysd.n))
    g, o = run_pair(sysd, s0, 300, T=300.0, xi=xi)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("plain heun T=300 injected max|dS| =", d)
    assert d <= TOL64


@pytest.mark.parametrize("fmt,ncomp", [("csr", 3), ("csr", 9), ("dia", 9), ("stencil", 9), ("stencil", 3)])
def test_plain_heun_formats_full_field(gpu, fmt, ncomp):
    sysd, s0 = tensor_csr_system(L=6, precision="fp64", ncomp=ncomp, fmt=fmt)
    rng = np.random.default_rng(2)
    stag = rng.normal(0, 5.0, (3, sysd.n))
    g, o = run_pair(sysd, s0, 300, T=0.0, B=(0.3, -0.2, 1.0), stag=stag)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print(fmt, ncomp, "max|dS| =", d)
    assert d <= TOL64


@pytest.mark.parametrize("scheme", ["rk4", "rk5"])
def test_plain_rk_parity(gpu, scheme):
    sysd, s0 = tensor_csr_system(L=6, precision="fp64", ncomp=9, fmt="csr", scheme=scheme)
    g, o = run_pair(sysd, s0, 300, T=0.0, B=(0.0, 0.0, 0.5))
    d = _maxdiff(g.get_spins(), o.get_spins())
    print(scheme, "max|dS| =", d)
    assert d <= TOL64


def test_fp32_heun_vs_oracle(gpu):
    sysd, s0, _ = W.sc_ferro(L=12, precision="fp32")
    g, o = run_pair(sysd, s0, 200, T=0.0)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fp32 heun 200 steps max|dS| =", d)
    assert d <= 1e-4
    assert "pack" in g.info()["exch_encoding"]


@pytest.mark.parametrize("fmt,ncomp", [("csr", 9), ("dia", 9), ("stencil", 3)])
def test_fp32_formats_full_field(gpu, fmt, ncomp):
    sysd, s0 = tensor_csr_system(L=6, precision="fp32", ncomp=ncomp, fmt=fmt)
    rng = np.random.default_rng(2)
    stag = rng.normal(0, 5.0, (3, sysd.n))
    g, o = run_pair(sysd, s0, 100, T=0.0, B=(0.3, -0.2, 1.0), stag=stag)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fp32", fmt, ncomp, "max|dS| =", d)
    assert d <= 1e-4
    h = g.field()
    ho = o.field()
    assert _maxdiff(h, ho) <= 1e-4 * np.abs(ho).max()


@pytest.mark.parametrize("scheme", ["rk4", "rk5"])
def test_fp32_rk_vs_oracle(gpu, scheme):
    sysd, s0 = tensor_csr_system(L=6, precision="fp32", ncomp=9, fmt="csr", scheme=scheme)
    g, o = run_pair(sysd, s0, 100, T=0.0, B=(0.0, 0.0, 0.5))
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fp32", scheme, "max|dS| =", d)
    assert d <= 1e-4


def test_fp32_injected_T300(gpu):
    sysd, s0, _ = W.sc_ferro(L=8, precision="fp32", noise="injected")
    xi = np.random.default_rng(3).standard_normal((50, 3, sysd.n))
    g, o = run_pair(sysd, s0, 50, T=300.0, xi=xi)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fp32 injected T=300 50 steps max|dS| =", d)
    assert d <= 1e-4


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_counter_noise_matches_oracle(gpu, prec, tol):
    sysd, s0, _ = W.sc_ferro(L=8, precision=prec, seed=1234)
    sysd.stream = 77
    import pyoracle

    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    for t in (0, 1, 123456789012):
        a, b = g.noise_sample(t), o.noise_sample(t)
        d = _maxdiff(a, b)
        print(prec, t, d)
        assert d <= tol * max(1.0, np.abs(b).max())


def test_counter_noise_trajectory_fp64(gpu):
    """Counter mode: GPU FP64 plain vs oracle at T = 300 K; the noise differs only by libm
    rounding (1e-16 relative), so the trajectories agree to 1e-9 over 100 steps."""
    sysd, s0, _ = W.sc_ferro(L=8, precision="fp64")
    g, o = run_pair(sysd, s0, 100, T=300.0)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("counter T=300 100 steps max|dS| =", d)
    assert d <= 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reduce_mag_energy(gpu, prec):
    sysd, s0 = tensor_csr_system(L=6, precision=prec, ncomp=9, fmt="csr")
    rng = np.random.default_rng(2)
    stag = rng.normal(0, 5.0, (3, sysd.n))
    g, o = run_pair(sysd, s0, 0, T=0.0, B=(0.3, -0.2, 1.0), stag=stag)
    mg, eg = g.reduce()
    mo, eo = o.reduce()
    tol = 1e-12 if prec == "fp64" else 1e-6
    print(prec, mg, mo, eg, eo)
    assert _maxdiff(mg, mo) <= tol
    assert abs(eg - eo) <= tol * abs(eo) * 10


def test_run_sampled_matches_oracle(gpu):
    sysd, s0, _ = W.sc_ferro(L=8, precision="fp64")
    import pyoracle

    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    for c in (g, o):
        c.set_spins(s0)
        c.set_temperature(0.0)
    a = g.run_sampled(0, 250, 50)
    b = o.run_sampled(0, 250, 50)
    assert a.shape == (5, 1, 3)
    assert _maxdiff(a, b) <= 1e-12


def test_dipolar_matches_bruteforce(gpu):
    for prec, tol in (("fp64", 1e-9), ("fp32", 2e-5)):
        sysd, s0 = W.sc_film_dip(Lx=8, Ly=6, Lz=4, precision=prec)
        import pyoracle

        g = sllg.Context(sysd)
        o = pyoracle.OracleContext(sysd)
        for c in (g, o):
            c.set_spins(s0)
        g.step(0, 1)  # refresh at t = 0 computes the dipolar field of s0
        o.step(0, 1)
        d = _maxdiff(g.get_spins(), o.get_spins())
        print(prec, "dipolar step max|dS| =", d)
        assert d <= (1e-10 if prec == "fp64" else 1e-5)
        # field at the current spins includes the stale dipolar field of s0 on both sides
        hg, ho = g.field(), o.field()
        rel = _maxdiff(hg, ho) / np.abs(ho).max()
        print(prec, "field rel err", rel)
        assert rel <= tol


def test_dipolar_half_grid_equals_full_fft(gpu):
    """FP32 dipolar by 2D transforms of the x < Gx/2 half of the padded grid plus the fused
    x-transform / k-space product kernel equals hipFFT's 3D transforms of the whole grid to
    FP32 rounding (x extents Gx = 16, 64, 128, 256; the last kernel block partly filled),
    over 20 thermal steps with the dipolar field refreshed every step."""
    for dims in ((8, 6, 4), (32, 12, 4), (64, 10, 2), (128, 3, 2)):
        res = []
        for flags in (0, sllg.FLAG_DIPOLAR_FULL_FFT):
            sysd, s0 = W.sc_film_dip(Lx=dims[0], Ly=dims[1], Lz=dims[2], precision="fp32", refresh=1)
            sysd.flags = flags
            g = sllg.Context(sysd)
            g.set_spins(s0)
            g.set_temperature(50.0)
            g.step(0, 20)
            res.append((g.get_spins(), g.field()))
            g.close()
        ds = _maxdiff(res[0][0], res[1][0])
        rel = _maxdiff(res[0][1], res[1][1]) / np.abs(res[1][1]).max()
        print(dims, "half grid vs full FFT: max|dS|", ds, "field rel", rel)
        assert ds <= 1e-5 and rel <= 1e-5


def test_dipolar_rejects_shared_mesh_point(gpu):
    """Two spins on one mesh point have no finite dipolar coupling: rejected at create."""
    sysd, _ = W.sc_film_dip(Lx=8, Ly=6, Lz=4, precision="fp32")
    mesh = np.array(sysd.dip["mesh"]).copy()
    mesh[3] = mesh[0]
    sysd.dip = {**sysd.dip, "mesh": mesh}
    with pytest.raises(sllg.LLGError) as e:
        sllg.Context(sysd)
    assert e.value.code == -1 and "share a mesh point" in str(e.value)


def test_invalid_descriptor_rejected(gpu):
    sysd, s0, _ = W.sc_ferro(L=4, precision="fp32")
    bad = dict(sysd.exch)
    adj = bad["adjncy"].copy()
    adj[5] = sysd.n + 3
    bad["adjncy"] = adj
    sysd.exch = bad
    with pytest.raises(sllg.LLGError) as e:
        sllg.Context(sysd)
    assert e.value.code == -1


# ----------------------------------------------------------------------------- sublattice-major order

def _fcc_multi(L=4, precision="fp32", noise="counter", dip=False):
    """fcc with 4 species (one per basis atom, 4 sublattices), every local term, optional dipolar;
    the CSR declares n_atoms = 4 so the FP32 path stores spins sublattice-major."""
    sysd, s0 = W.fcc_3shell(L=L, precision=precision, noise=noise)
    sysd.species = [dict(mu=2.22 if t != 2 else 1.5, lambda_=0.05 + 0.02 * t, gamma_rel=1.0 + 0.1 * t,
                         k1u=(1e-23 if t % 2 == 0 else -3e-24), k1u_dir=(0.0, 0.6, 0.8), sublattice=t)
                    for t in range(4)]
    sysd.spin_species = (np.arange(sysd.n) % 4).astype(np.uint32)
    sysd.n_sublattice = 4
    if dip:
        uc = sllg.unit_cell("fcc")
        sysd.dip = dict(enabled=1, mesh=sllg.mesh(uc, (L, L, L)), dim=(L, L, L), nk=(2, 2, 2),
                        abc=(3.6e-10,) * 3)
    return sysd, s0


def test_sublattice_order_is_used_and_matches_oracle(gpu):
    sysd, s0 = _fcc_multi(L=4)
    import pyoracle

    g = sllg.Context(sysd)
    assert "-t4-" in g.info()["exch_encoding"]
    rng = np.random.default_rng(4)
    stag = rng.normal(0, 3.0, (3, sysd.n))
    g2, o = run_pair(sysd, s0, 100, T=0.0, B=(0.1, 0.2, 0.7), stag=stag)
    d = _maxdiff(g2.get_spins(), o.get_spins())
    print("fcc sublattice-major 100 steps max|dS| =", d)
    assert d <= 1e-4
    assert _maxdiff(g2.field(), o.field()) <= 1e-4 * np.abs(o.field()).max()
    mg, eg = g2.reduce()
    mo, eo = o.reduce()
    assert _maxdiff(mg, mo) <= 1e-5 and abs(eg - eo) <= 1e-5 * abs(eo)
    # counter noise is keyed by the reference spin index: identical to the oracle's draws
    a, b = g2.noise_sample(17), o.noise_sample(17)
    assert _maxdiff(a, b) <= 2e-5 * np.abs(b).max()


def test_sublattice_order_equals_identity_order(gpu):
    """The device order is invisible: same results (to FP32 rounding of the same sums) with
    the permutation switched off, at T > 0 with counter noise."""
    sysd, s0 = _fcc_multi(L=4)
    res = []
    for flags in (0, sllg.FLAG_NO_SUBLATTICE_ORDER):
        sysd.flags = flags
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.set_temperature(200.0)
        g.step(0, 50)
        res.append(g.get_spins())
    assert _maxdiff(res[0], res[1]) <= 2e-5


def test_sublattice_order_injected_noise(gpu):
    sysd, s0 = _fcc_multi(L=4, noise="injected")
    xi = np.random.default_rng(9).standard_normal((30, 3, sysd.n))
    g, o = run_pair(sysd, s0, 30, T=300.0, xi=xi)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fcc sublattice-major injected max|dS| =", d)
    assert d <= 1e-4


def test_sublattice_order_dipolar(gpu):
    sysd, s0 = _fcc_multi(L=4, dip=True)
    import pyoracle

    g = sllg.Context(sysd)
    o = pyoracle.OracleContext(sysd)
    for c in (g, o):
        c.set_spins(s0)
        c.step(0, 1)
    hg, ho = g.field(), o.field()
    rel = _maxdiff(hg, ho) / np.abs(ho).max()
    print("fcc dipolar field rel err", rel)
    assert rel <= 2e-5


def test_launch_shape_does_not_change_results(gpu):
    """Each thread of the FP32 Heun kernels owns one spin: every block size and tile order
    gives identical bits.  The autotuner keeps the state it was given."""
    sysd, s0 = _fcc_multi(L=4)
    res = []
    for tpb, xcd, unr, thc, g3 in ((256, 1, 2, 1, 0), (512, 0, 4, 0, 1), (1024, 1, 4, 1, 1), (256, 0, 2, 0, 0),
                                    (256, 1, 4, 0, 1)):
        g = sllg.Context(sysd)
        g.set_launch(tpb, xcd, unr, thc, g3)
        g.set_spins(s0)
        g.set_temperature(150.0)
        g.step(0, 40)
        res.append(g.get_spins())
    for r in res[1:]:
        assert np.array_equal(r, res[0])
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(150.0)
    g.step(0, 10)
    before = g.get_spins()
    t = g.autotune(5)
    assert t["tpb"] in (256, 512, 1024) and t["xcd"] in (0, 1) and t["unroll"] in (2, 4)
    assert t["candidates"] == 48 and t["thermal_cache"] in (0, 1) and t["gather_copy"] in (0, 1)
    assert np.array_equal(g.get_spins(), before)
    g.step(10, 30)
    assert np.array_equal(g.get_spins(), res[0])
    with pytest.raises(sllg.LLGError):
        g.set_launch(384, 1)
    with pytest.raises(sllg.LLGError):
        g.set_launch(256, 1, 3)


def test_fp32_counter_noise_statistics(gpu):
    """The FP32 path's Gaussian draws (hardware log/sqrt/sin/cos in Box-Muller) over 786k
    samples: mean, variance, kurtosis within sampling error and a Kolmogorov-Smirnov test."""
    from scipy import stats
    sysd, _, _ = W.sc_ferro(L=64, precision="fp32", seed=99)
    g = sllg.Context(sysd)
    z = g.noise_sample(5).reshape(-1)
    N = z.size
    assert abs(z.mean()) < 5 / np.sqrt(N)
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0 / N)
    assert abs(stats.kurtosis(z, fisher=False) - 3.0) < 5 * np.sqrt(24.0 / N)
    assert stats.kstest(z, "norm").pvalue > 1e-3


def _slices(enc):
    # "...-slices<stored>of<total>"
    tail = enc.rsplit("-slices", 1)[1]
    a, b = tail.split("of")
    return int(a), int(b)


@pytest.mark.parametrize("kind", ["fcc", "sc"])
def test_slice_sharing_is_exact(gpu, kind):
    """Repeated 64-row slices share one stored copy of their packed entries: identical bits
    to the encoding that stores every slice, and the oracle's result to FP32 rounding."""
    if kind == "fcc":
        sysd, s0 = _fcc_multi(L=8)
    else:
        sysd, s0, _ = W.sc_ferro(L=16, precision="fp32")
    res, encs = [], []
    for flags in (0, sllg.FLAG_NO_SLICE_SHARING):
        sysd.flags = flags
        g = sllg.Context(sysd)
        encs.append(g.info()["exch_encoding"])
        g.set_spins(s0)
        g.set_temperature(100.0)
        g.step(0, 60)
        res.append(g.get_spins())
        m, e = g.reduce()
        res.append(np.array([e]))
    sysd.flags = 0
    st, tot = _slices(encs[0])
    st2, tot2 = _slices(encs[1])
    print(kind, encs)
    assert tot == tot2 and st2 == tot and st < tot
    assert np.array_equal(res[0], res[2]) and np.array_equal(res[1], res[3])
    g, o = run_pair(sysd, s0, 60, T=0.0, B=(0.0, 0.3, 0.5))
    assert _maxdiff(g.get_spins(), o.get_spins()) <= 1e-4


@pytest.mark.parametrize("cpu_form", [0, 1])
def test_fourspin_plain_bit_parity_and_fp32(gpu, cpu_form):
    """Four-spin term (plaquette quartets): FP64 plain path bit-identical to the oracle at
    T = 0 and with injected noise; FP32 path within float rounding; energies agree."""
    L = 6
    xadj, jkl = W.plaquette_quartets((L, L, L))
    for prec, noise, tol in (("fp64", "counter", TOL64), ("fp64", "injected", TOL64), ("fp32", "counter", 1e-4)):
        sysd, s0, _ = W.sc_ferro(L=L, precision=prec, noise=noise)
        sysd.four = dict(xadj=xadj, jkl=jkl, jq=25.0, cpu_form=cpu_form)
        xi = None
        T = 0.0
        if noise == "injected":
            xi = np.random.default_rng(3).standard_normal((60, 3, sysd.n))
            T = 300.0
        g, o = run_pair(sysd, s0, 60, T=T, B=(0.0, 0.2, 0.4), xi=xi)
        d = _maxdiff(g.get_spins(), o.get_spins())
        print("fourspin", cpu_form, prec, noise, d)
        assert d <= tol
        _, eg = g.reduce()
        _, eo = o.reduce()
        assert abs(eg - eo) <= (1e-12 if prec == "fp64" else 1e-5) * abs(eo)


def test_fourspin_sublattice_order(gpu):
    """Four-spin quartets on the sublattice-major FP32 path (indices permuted at the boundary)."""
    sysd, s0 = _fcc_multi(L=4)
    rng = np.random.default_rng(12)
    sets = [rng.choice(sysd.n, 4, replace=False) for _ in range(300)]
    xadj, jkl = W.quartets_from_sets(sysd.n, sets)
    sysd.four = dict(xadj=xadj, jkl=jkl, jq=30.0)
    g, o = run_pair(sysd, s0, 50, T=0.0, B=(0.1, 0.0, 0.3))
    assert "-t4-" in g.info()["exch_encoding"]
    d = _maxdiff(g.get_spins(), o.get_spins())
    print("fourspin fcc sublattice-major", d)
    assert d <= 1e-4
    assert _maxdiff(g.field(), o.field()) <= 1e-4 * np.abs(o.field()).max()


@pytest.mark.parametrize("case", ["sc", "fcc", "diag", "tensor", "sc16", "fcc16", "sc64open", "sc16open"])
def test_line_mode(gpu, case):
    """Slices made of periodic 64-cell lines (or four 16-cell lines) use the line reader
    (window loads, rotations across the wave or within 16-lane rows for the +-1
    neighbours, scalar coupling stream); 16-cell lattices keep the packed entries for the
    slices at the periodic y boundary.  Same physics as the packed reader to FP32 rounding
    (entries are summed per line, not in stored order) and as the oracle; energies agree."""
    want = "line64"
    if case == "sc64open":  # open along the lines: zero-filling shifts at the line ends
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(4, 4, 64), pbc=(1, 1, 0))
        want = "line64o"
    elif case == "sc16open":
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(2, 32, 16), pbc=(1, 1, 0))
        want = "line16o"
    elif case == "sc16":
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(2, 32, 16))
        want = "line16"
    elif case == "fcc16":
        sysd, s0 = W.fcc_3shell(precision="fp32", dims=(2, 32, 16))
        want = "line16"
    elif case == "sc":
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(4, 4, 64))
    elif case == "fcc":
        sysd, s0 = W.fcc_3shell(precision="fp32", dims=(2, 3, 64))
        sysd.species = [dict(mu=2.22 if t != 2 else 1.5, lambda_=0.05, sublattice=t) for t in range(4)]
        sysd.spin_species = (np.arange(sysd.n) % 4).astype(np.uint32)
        sysd.n_sublattice = 4
    elif case == "diag":  # three diagonal components per entry (groups of 2 lines)
        sysd, s0 = tensor_csr_system(precision="fp32", ncomp=3, dims=(2, 3, 64))
    else:
        sysd, s0 = tensor_csr_system(precision="fp32", ncomp=9, dims=(3, 2, 64))
    res, encs = [], []
    for flags in (0, sllg.FLAG_NO_LINE_MODE):
        sysd.flags = flags
        g = sllg.Context(sysd)
        encs.append(g.info()["exch_encoding"])
        g.set_spins(s0)
        g.set_temperature(150.0)
        g.set_applied(np.array([0.1, 0.0, 0.4]))
        g.step(0, 40)
        res.append((g.get_spins(), g.reduce()[1], g.field()))
    sysd.flags = 0
    # the line encoding keeps the packed entries: with the line reader off, the handle
    # reproduces the packed-only handle bit for bit
    g = sllg.Context(sysd)
    g.set_launch(256, 1, 2, 1, 1, 0)
    g.set_spins(s0)
    g.set_temperature(150.0)
    g.set_applied(np.array([0.1, 0.0, 0.4]))
    g.step(0, 40)
    assert np.array_equal(g.get_spins(), res[1][0])
    print(case, encs)
    assert want in encs[0] and "line" not in encs[1]
    if want.startswith("line16"):
        assert "-packed0" not in encs[0]  # boundary slices read from the packed entries
    assert _maxdiff(res[0][0], res[1][0]) <= 2e-5
    assert abs(res[0][1] - res[1][1]) <= 1e-5 * abs(res[1][1])
    assert _maxdiff(res[0][2], res[1][2]) <= 1e-5 * np.abs(res[1][2]).max()
    g, o = run_pair(sysd, s0, 40, T=0.0, B=(0.0, 0.2, 0.5))
    assert _maxdiff(g.get_spins(), o.get_spins()) <= 1e-4


def test_line_mode_gather_copy(gpu):
    """Line mode with and without the AoS3 gather copy gives identical bits, also across a
    set_spins between steps (the copy is rebuilt from the new state); the deterministic field
    at the predicted spins (kept only as the AoS3 copy with the gather copy on) is identical."""
    sysd, s0 = W.fcc_3shell(precision="fp32", dims=(2, 2, 64))
    res = []
    for g3 in (0, 1):
        g = sllg.Context(sysd)
        assert "line" in g.info()["exch_encoding"]
        g.set_launch(256, 1, 2, 1, g3)
        g.set_spins(s0)
        g.set_temperature(100.0)
        g.step(0, 20)
        mid = g.get_spins()
        g.set_spins(mid[:, ::-1].copy())  # new state from the host: gather copy must follow
        g.step(20, 20)
        res.append((g.get_spins(), g.det_field()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("case", ["fcc", "sc", "sc64open", "diag", "tensor", "fcc_multi_dip"])
def test_tile_reader(gpu, case):
    """The LDS-staged line reader (line_mode 2: a workgroup stages the lines its slices read,
    then sums the line records over windows read from LDS) gives the line reader's bits: same
    records, same order of the sums; with and without the thermal hand-off, across a
    set_spins between steps, with several species, staggered and dipolar fields.  Its
    trajectory matches the oracle to FP32 rounding."""
    stag = False
    if case == "fcc":
        sysd, s0 = W.fcc_3shell(precision="fp32", dims=(2, 2, 64))
    elif case == "sc":
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(4, 4, 64))
    elif case == "sc64open":
        sysd, s0, _ = W.sc_ferro(precision="fp32", dims=(4, 4, 64), pbc=(1, 1, 0))
    elif case == "diag":
        sysd, s0 = tensor_csr_system(precision="fp32", ncomp=3, dims=(2, 4, 64))
    elif case == "tensor":
        sysd, s0 = tensor_csr_system(precision="fp32", ncomp=9, dims=(4, 2, 64))
    else:
        sysd, s0 = W.fcc_3shell(precision="fp32", dims=(2, 2, 64))
        sysd.species = [dict(mu=2.22 if t != 2 else 1.5, lambda_=0.05 + 0.02 * t, k1u=1e-23 * (t % 2),
                             k1u_dir=(0.0, 0.6, 0.8), sublattice=t) for t in range(4)]
        sysd.spin_species = (np.arange(sysd.n) % 4).astype(np.uint32)
        sysd.n_sublattice = 4
        uc = sllg.unit_cell("fcc")
        sysd.dip = dict(enabled=1, mesh=sllg.mesh(uc, (2, 2, 64)), dim=(2, 2, 64), nk=(2, 2, 2),
                        abc=(3.6e-10,) * 3, refresh=5)
        stag = True
    rng = np.random.default_rng(3)
    hst = rng.normal(0.0, 0.3, (3, sysd.n))
    res = []
    for line_mode, thc in ((1, 1), (2, 1), (2, 0)):
        g = sllg.Context(sysd)
        enc = g.info()["exch_encoding"]
        assert "-tile" in enc, enc
        g.set_launch(512, 1, 2, thc, 1, line_mode)
        g.set_spins(s0)
        g.set_temperature(150.0)
        g.set_applied(np.array([0.1, 0.0, 0.4]))
        if stag:
            g.set_staggered(hst)
        g.step(0, 25)
        mid = g.get_spins()
        g.set_spins(mid[:, ::-1].copy())  # the AoS3 copy the tile reader stages must follow
        g.step(25, 25)
        res.append((g.get_spins(), g.field()))
    print(case, enc)
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0])
        assert np.array_equal(r[1], res[0][1])
    if case in ("fcc", "sc64open", "tensor"):
        g, o = run_pair(sysd, s0, 40, T=0.0, B=(0.0, 0.2, 0.5), launch=(512, 1, 2, 1, 1, 2))
        assert _maxdiff(g.get_spins(), o.get_spins()) <= 1e-4


def _basis_system(case):
    """CSR systems of other bases for the tile-stream shapes: bcc (two sublattices, two
    shells), a three-atom basis on the cube diagonal (nearest neighbours along it), and fcc
    with three lines per sublattice."""
    if case == "bcc":
        uc, dims, J = sllg.unit_cell("bcc"), (2, 2, 64), [4e-21, 1e-21]
    elif case == "three_atoms":
        xred = np.array([[0, 0, 0], [1 / 3, 1 / 3, 1 / 3], [2 / 3, 2 / 3, 2 / 3]], float)
        uc = sllg.UnitCellSpec(xred, np.eye(3), (1, 1, 1), np.zeros((3, 3), np.uint32))
        dims, J = (2, 4, 64), [4e-21, 1e-21]
    else:
        uc, dims, J = sllg.unit_cell("fcc"), (3, 1, 64), [5e-21]
    na = len(uc.xred)
    ints, _ = sllg.shells(uc, J)
    xadj, adj, vals = sllg.build_csr(uc, dims, ints, [2.22] * na)
    n = int(np.prod(dims)) * na
    sysd = sllg.System(n=n, species=[dict(mu=2.22, lambda_=0.1)], anis=dict(kc=2e-24),
                       exch=dict(format=sllg.EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals, n_atoms=na),
                       dt=1e-16, precision="fp32")
    return sysd, sllg.random_spins(n, 5, 105)


@pytest.mark.parametrize("case", ["bcc", "three_atoms", "fcc_3_lines"])
def test_tile_reader_shapes(gpu, case):
    """Tile streams where the workgroup shape fits (bcc: 8 slices = 4 lines of each of the 2
    sublattices) and none where it does not (3 sublattices: 8 and 16 slices are no whole number
    of lines per sublattice; fcc with 3 lines per sublattice and 12 slices: no workgroup of 8
    or 16 divides them); line_mode 2 then runs the line reader.  Either way the bits are the
    line reader's, and the trajectory matches the oracle to FP32 rounding."""
    sysd, s0 = _basis_system(case)
    g = sllg.Context(sysd)
    enc = g.info()["exch_encoding"]
    print(case, enc)
    assert "line64" in enc
    assert ("-tile8x" in enc) == (case == "bcc")
    res = []
    for line_mode in (1, 2):
        g.set_launch(512, 1, 2, 0, 1, line_mode)
        g.set_spins(s0)
        g.set_temperature(200.0)
        g.step(0, 40)
        res.append(g.get_spins())
    g.close()
    assert np.array_equal(res[0], res[1])
    g, o = run_pair(sysd, s0, 40, T=0.0, B=(0.0, 0.1, 0.3), launch=(512, 1, 2, 0, 1, 2))
    assert _maxdiff(g.get_spins(), o.get_spins()) <= 1e-4


def _fftx_system(prec, dip, seed=7, dims=(6, 5, 4)):
    """Open sc lattice, two species on a checkerboard, random anisotropic/antisymmetric 3x3
    couplings on the 6 nearest and 12 next-nearest mesh vectors; the same couplings as an
    FFT-exchange list and as an explicit CSR (ncomp 9) assembled neighbour by neighbour."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = dims
    n = Lx * Ly * Lz
    idx = np.arange(n)
    x, y, z = idx // (Ly * Lz), (idx // Lz) % Ly, idx % Lz
    spec = ((x + y + z) % 2).astype(np.uint32)
    mus = [2.22, 1.5]
    vecs = [v for v in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
    vecs += [(a, b, 0) for a in (1, -1) for b in (1,)] + [(a, 0, b) for a in (1, -1) for b in (1,)] + \
            [(0, a, b) for a in (1, -1) for b in (1,)]
    ints = []
    for s1 in range(2):
        for s2 in range(2):
            for v in vecs:
                J = rng.normal(0, 1e-21, (3, 3)) + np.eye(3) * 3e-21
                ints.append((s1, s2, v, J))
                ints.append((s2, s1, tuple(-c for c in v), J.T.copy()))  # J_ji = J_ij^T
    mesh = np.stack([x, y, z], axis=1).astype(np.uint32)
    a = 3e-10
    species = [dict(mu=mus[t], lambda_=0.1) for t in range(2)]
    fft = sllg.System(n=n, species=species, spin_species=spec, precision=prec, dt=1e-16, seed=seed,
                      fftx=ints, dip=dict(enabled=1 if dip else 0, dim=dims, nk=(1, 1, 1), abc=(a, a, a), mesh=mesh))
    # the same couplings, CSR ncomp 9 in tesla (J / (mu_i mu_B)), interaction order per row
    xadj, adj, vals = [0], [], []
    pos = {(int(x[i]), int(y[i]), int(z[i])): i for i in range(n)}
    for i in range(n):
        for s1, s2, v, J in ints:
            if s1 != spec[i]:
                continue
            j = pos.get((int(x[i]) + v[0], int(y[i]) + v[1], int(z[i]) + v[2]))
            if j is None or spec[j] != s2:
                continue
            adj.append(j)
            vals.append(J.reshape(9) / (mus[s1] * sllg.MU_B))
        xadj.append(len(adj))
    vals = np.ascontiguousarray(np.array(vals).T)
    csr = sllg.System(n=n, species=species, spin_species=spec, precision=prec, dt=1e-16, seed=seed,
                      exch=dict(format=sllg.EXCH_CSR, xadj=np.array(xadj, np.uint32), adjncy=np.array(adj, np.uint32),
                                vals=vals, ncomp=9),
                      dip=dict(enabled=1 if dip else 0, dim=dims, nk=(1, 1, 1), abc=(a, a, a), mesh=mesh,
                               every_stage=1, refresh=1))
    return fft, csr, sllg.random_spins(n, seed, seed + 100)


@pytest.mark.parametrize("prec,dip,dims", [("fp64", False, (6, 5, 4)), ("fp64", True, (6, 5, 4)),
                                            ("fp32", False, (6, 5, 4)), ("fp32", True, (6, 5, 4)),
                                            ("fp32", False, (8, 5, 4)), ("fp32", True, (8, 6, 2))])
def test_fft_exchange_matches_direct_sum(gpu, prec, dip, dims):
    """FFT exchange (species-resolved 3x3 convolution on the zero-padded mesh, with the
    dipolar terms in the same kernel when enabled) equals the direct sum of the same
    couplings: fields at the same spins (exchange only), energies, and 30 Heun steps at
    T = 0 against the oracle running the CSR (plus its brute-force dipolar sum, refreshed
    every stage).  Gx = 12 transforms the whole padded grid; Gx = 16 (FP32) runs the
    half-grid path with two species."""
    import pyoracle

    fft, csr, s0 = _fftx_system(prec, dip, dims=dims)
    g = sllg.Context(fft)
    h = sllg.Context(csr)
    o = pyoracle.OracleContext(csr)
    for c in (g, h, o):
        c.set_spins(s0)
        c.set_applied(np.array([0.0, 0.1, 0.3]))
    tol = 1e-10 if prec == "fp64" else 2e-5
    if not dip:  # a dipolar field keeps its refresh schedule (zero before the first step)
        hf, hc = g.field(), h.field()
        rel = _maxdiff(hf, hc) / np.abs(hc).max()
        print(prec, dip, "field rel", rel)
        assert rel <= tol
    eg, eh = g.reduce()[1], h.reduce()[1]  # both evaluate every term at the current spins
    assert abs(eg - eh) <= (1e-10 if prec == "fp64" else 1e-5) * abs(eh)
    for c in (g, o):
        c.step(0, 30)
    d = _maxdiff(g.get_spins(), o.get_spins())
    print(prec, dip, "30 steps max|dS|", d)
    assert d <= (1e-10 if prec == "fp64" else 1e-5)


def test_fft_exchange_half_grid_equals_full_fft(gpu):
    """One-species FFT exchange on an open film: the half-grid path (2D plane transforms and
    the x kernel with the complex product) equals the whole-grid 3D transforms to FP32
    rounding over 20 thermal steps, with and without the dipolar terms."""
    for dip in (False, True):
        res = []
        for flags in (0, sllg.FLAG_DIPOLAR_FULL_FFT):
            sysd, s0 = W.sc_film_fftx(Lx=32, Ly=12, Lz=4, precision="fp32", dipolar=dip)
            sysd.flags = flags
            g = sllg.Context(sysd)
            g.set_spins(s0)
            g.set_temperature(50.0)
            g.step(0, 20)
            res.append((g.get_spins(), g.field()))
            g.close()
        ds = _maxdiff(res[0][0], res[1][0])
        rel = _maxdiff(res[0][1], res[1][1]) / np.abs(res[1][1]).max()
        print("dip", dip, "half grid vs full FFT: max|dS|", ds, "field rel", rel)
        assert ds <= 1e-5 and rel <= 1e-5


def test_fft_exchange_rejects_bad_descriptors(gpu):
    """FFT exchange needs exch.format NONE, species in range and one spin per species and
    mesh point."""
    import dataclasses

    fft, csr, _ = _fftx_system("fp32", False)
    with pytest.raises(sllg.LLGError) as e:
        sllg.Context(dataclasses.replace(fft, exch=csr.exch))
    assert "LLG_EXCH_NONE" in str(e.value)
    ints = list(fft.fftx)
    ints[0] = (5, 0, ints[0][2], ints[0][3])
    with pytest.raises(sllg.LLGError):
        sllg.Context(dataclasses.replace(fft, fftx=ints))
    mesh = np.array(fft.dip["mesh"]).copy()
    mesh[2] = mesh[0]  # spins 0 and 2 share species 0 and now a mesh point
    with pytest.raises(sllg.LLGError) as e:
        sllg.Context(dataclasses.replace(fft, dip={**fft.dip, "mesh": mesh}))
    assert "share a mesh point" in str(e.value)
