"""Pins of the FP64 oracle (CPU): reference golden vectors, published known answers,
closed forms and invariants, each chosen so a plausible slip (sign, index, dropped term,
transposed tensor) fails.  None of these retypes the oracle's formula."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import sllg
from sllg import workloads as W

from helpers import full_anis, tensor_csr_system

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------- random numbers

def test_mwc1616_matches_reference_golden():
    """MWC1616 + Marsaglia polar normal against draws of the reference's own src/random.cpp."""
    with open(os.path.join(HERE, "golden", "mwc1616_golden.json")) as f:
        g = json.load(f)
    for s in g["seeds"]:
        r = pyoracle.MWC(s["a"], s["b"])
        u = [r.uniform() for _ in range(g["count"])]
        assert u == [float(x) for x in s["uniform"]]
        r = pyoracle.MWC(s["a"], s["b"])
        nv = [r.normal() for _ in range(g["count"])]
        assert nv == [float(x) for x in s["normal"]]


@pytest.mark.skipif(not os.path.exists(pyoracle.REF_RANDOM), reason="reference RNG build (oracle/_ref) absent")
def test_mwc1616_matches_reference_build_live():
    out = subprocess.run([pyoracle.REF_RANDOM, "7", "11", "500"], check=True, capture_output=True, text=True).stdout.split()
    vals = [float(x) for x in out[1::2]]
    r = pyoracle.MWC(7, 11)
    assert [r.uniform() for _ in range(500)] == vals[:500]
    r = pyoracle.MWC(7, 11)
    assert [r.normal() for _ in range(500)] == vals[500:]


def test_random_spins_product_equals_oracle():
    """setSpinsRandom restated twice (product host builder, oracle); unit length."""
    a = pyoracle.random_spins(42, 142, 2000)
    b = sllg.random_spins(2000, 42, 142)
    assert np.array_equal(a, b)
    assert np.abs(np.linalg.norm(a, axis=0) - 1).max() < 1e-15
    # uniform on the sphere: mean ~ 0, <z^2> ~ 1/3
    assert np.abs(a.mean(axis=1)).max() < 0.05
    assert abs((a[2] ** 2).mean() - 1 / 3) < 0.03


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123 (Salmon et al. SC'11)."""
    assert pyoracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert pyoracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert pyoracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_counter_normals_statistics():
    """Moments, lag-1 correlation and stream independence of the counter-based normals."""
    sysd, _, _ = W.sc_ferro(L=40, precision="fp64", seed=99)
    o = pyoracle.OracleContext(sysd)
    z = np.concatenate([o.noise_sample(t).ravel() for t in range(4)])
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1) < 0.01
    assert abs((z ** 4).mean() - 3) < 0.05
    assert abs(np.corrcoef(z[:-1], z[1:])[0, 1]) < 0.01
    sysd.stream = 1
    o2 = pyoracle.OracleContext(sysd)
    z2 = o2.noise_sample(0).ravel()
    assert abs(np.corrcoef(z[:z2.size], z2)[0, 1]) < 0.01
    assert np.array_equal(o.noise_sample(3), o.noise_sample(3))


# ----------------------------------------------------------------------------- integrators

def _macrospin(scheme, dt, steps, B=100.0, lam=0.1, theta0=1.0, phi0=0.3):
    sysd = sllg.System(n=1, species=[dict(mu=2.22, lambda_=lam)], dt=dt, scheme=scheme, precision="fp64")
    o = pyoracle.OracleContext(sysd)
    s0 = np.array([[math.sin(theta0) * math.cos(phi0)], [math.sin(theta0) * math.sin(phi0)], [math.cos(theta0)]])
    o.set_spins(s0)
    o.set_applied([0.0, 0.0, B])
    o.step(0, steps)
    tp = sllg.GAMMA_E * dt * steps
    # dS/dt' = -1/(1+l^2) [S x H + l S x (S x H)], H = B z:
    # tan(theta/2) = tan(theta0/2) exp(-l B t'/(1+l^2)), phi = phi0 + B t'/(1+l^2)
    th = 2 * math.atan(math.tan(theta0 / 2) * math.exp(-lam * B * tp / (1 + lam ** 2)))
    ph = phi0 + B * tp / (1 + lam ** 2)
    exact = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    return float(np.abs(o.get_spins()[:, 0] - exact).max())


@pytest.mark.parametrize("scheme", ["heun", "rk4", "rk5"])
def test_macrospin_matches_closed_form(scheme):
    assert _macrospin(scheme, 1e-16, 20000, B=10.0) < 1e-7


@pytest.mark.parametrize("scheme,order", [("heun", 2), ("rk4", 4), ("rk5", 5)])
def test_convergence_order(scheme, order):
    """Global error ~ dt^p on the analytic macrospin (B rdt = 0.35 .. 0.09 rad per step)."""
    T = 2e-12
    errs = [_macrospin(scheme, T / m, m, B=100.0) for m in (100, 200, 400)]
    p1 = math.log2(errs[0] / errs[1])
    p2 = math.log2(errs[1] / errs[2])
    assert abs(p2 - order) < 0.35, (errs, p1, p2)


def test_unit_length_and_energy_descent():
    """|S| = 1 after every step; energy non-increasing at T = 0 with damping (all terms)."""
    sysd, s0 = tensor_csr_system(L=5, precision="fp64", ncomp=9)
    o = pyoracle.OracleContext(sysd)
    o.set_spins(s0)
    o.set_applied([0.2, 0.0, 0.5])
    E = [o.reduce()[1]]
    for _ in range(30):
        o.step(0, 20)
        E.append(o.reduce()[1])
        assert np.abs(np.linalg.norm(o.get_spins(), axis=0) - 1).max() < 1e-14
    assert all(b <= a + 1e-12 * abs(a) for a, b in zip(E, E[1:]))


def test_energy_conserved_without_damping():
    """lambda = 0: the energy drift of RK4 shrinks as dt^4 (conservative dynamics)."""
    drift = []
    for dt in (2e-16, 1e-16):
        sysd, s0 = tensor_csr_system(L=4, precision="fp64", ncomp=9, scheme="rk4")
        for sp in sysd.species:
            sp["lambda_"] = 0.0
        o = pyoracle.OracleContext(sysd)
        o.set_spins(s0)
        sysd.dt = dt
        o2 = pyoracle.OracleContext(sysd)
        o2.set_spins(s0)
        E0 = o2.reduce()[1]
        o2.step(0, int(round(4e-14 / dt)))
        drift.append(abs(o2.reduce()[1] - E0) / abs(E0))
    assert drift[0] < 1e-6
    assert drift[1] < drift[0] / 8


def test_uniform_ferromagnet_has_no_exchange_torque():
    sysd, _, _ = W.sc_ferro(L=6, precision="fp64", K=0.0)
    o = pyoracle.OracleContext(sysd)
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    s = np.repeat(u[:, None], sysd.n, axis=1)
    o.set_spins(s)
    o.step(0, 50)
    assert np.abs(o.get_spins() - s).max() < 1e-14


def test_compat_llgcpu_quirks():
    """OR_COMPAT_LLGCPU reproduces llgCPU: the corrector sees H(S_n) + H(S~) (A1) and the
    z noise reuses the x draw (A2); without it the GPU semantics hold."""
    sysd, s0, _ = W.sc_ferro(L=4, precision="fp64", K=0.0, noise="injected")
    xi = np.random.default_rng(0).standard_normal((1, 3, sysd.n))
    res = {}
    for opt in (0, pyoracle.COMPAT_LLGCPU):
        o = pyoracle.OracleContext(sysd, opt)
        o.set_spins(s0)
        o.set_temperature(100.0)
        o.push_noise(xi)
        o.step(0, 1)
        res[opt] = o.get_spins()
    # manual restatement of one step of each semantics with numpy
    J = sysd.exch["vals"][0][0]
    xadj, adj = sysd.exch["xadj"], sysd.exch["adjncy"]
    n = sysd.n

    def hex_(S):
        return np.stack([np.array([J * S[c][adj[xadj[i]:xadj[i + 1]]].sum() for i in range(n)]) for c in range(3)])

    lam, mu = 0.1, 2.22
    sigma = math.sqrt(2 * sllg.K_B * lam / (mu * sllg.MU_B * 1e-16 * sllg.GAMMA_E))
    pf = -1 / (1 + lam * lam)
    rdt = 1e-16 * sllg.GAMMA_E

    def f(S, H):
        sxh = np.cross(S.T, H.T)
        return (pf * (sxh + lam * np.cross(S.T, sxh))).T

    for opt, compat in ((0, False), (pyoracle.COMPAT_LLGCPU, True)):
        th = xi[0] * math.sqrt(100.0) * sigma
        if compat:
            th = th.copy()
            th[2] = th[0]
        H1 = hex_(s0)
        fn = f(s0, H1 + th)
        Y = s0 + rdt * fn
        Y /= np.linalg.norm(Y, axis=0)
        H2 = (H1 + hex_(Y)) if compat else hex_(Y)
        S1 = s0 + 0.5 * (fn + f(Y, H2 + th)) * rdt
        S1 /= np.linalg.norm(S1, axis=0)
        assert np.abs(S1 - res[opt]).max() < 1e-12
    assert np.abs(res[0] - res[pyoracle.COMPAT_LLGCPU]).max() > 1e-6


# ----------------------------------------------------------------------------- field terms

def test_exchange_field_dense_bruteforce():
    """Oracle field (CSR, DIA and stencil) = brute-force dense sum over neighbours found by
    distance, with tensors J_ij applied as J[a][b] S_b and J_ji = J_ij^T."""
    L = 5
    for fmt in ("csr", "dia", "stencil"):
        sysd, s0 = tensor_csr_system(L=L, precision="fp64", ncomp=9, fmt=fmt)
        sysd.anis = {}
        for sp in sysd.species:
            sp["k1u"] = 0.0
        o = pyoracle.OracleContext(sysd)
        o.set_spins(s0)
        h = o.field()
        # brute force from the CSR tensors of the reference-order builder, by explicit loops
        sc, _ = tensor_csr_system(L=L, precision="fp64", ncomp=9, fmt="csr")
        xadj, adj, vals = sc.exch["xadj"], sc.exch["adjncy"], sc.exch["vals"]
        n = sc.n
        dense = np.zeros((3 * n, 3 * n))
        pos = np.stack(np.unravel_index(np.arange(n), (L, L, L)), axis=1)
        for i in range(n):
            for q in range(xadj[i], xadj[i + 1]):
                j = adj[q]
                d = (pos[j] - pos[i] + L // 2) % L - L // 2
                r = np.linalg.norm(d)
                assert r in (1.0, math.sqrt(2.0))  # first two sc shells only
                dense[3 * i:3 * i + 3, 3 * j:3 * j + 3] += vals[:, q].reshape(3, 3)
        # J_ji = J_ij^T for every pair
        for i in range(0, n, 7):
            for j in range(n):
                assert np.allclose(dense[3 * i:3 * i + 3, 3 * j:3 * j + 3], dense[3 * j:3 * j + 3, 3 * i:3 * i + 3].T)
        Sv = s0.T.reshape(-1)
        hb = (dense @ Sv).reshape(n, 3).T
        assert np.abs(h - hb).max() < 1e-9 * np.abs(hb).max(), fmt


def test_field_is_minus_energy_gradient():
    """H_i = -dE/dS_i / (mu_i mu_B) by central differences of the oracle energy: pins the
    exchange tensors (incl. DMI), uniaxial, K2, K4, cubic, Zeeman and staggered terms."""
    sysd, s0 = tensor_csr_system(L=3, precision="fp64", ncomp=9)
    o = pyoracle.OracleContext(sysd)
    o.set_spins(s0)
    B = np.array([0.4, -0.3, 0.9])
    o.set_applied(B)
    stag = np.random.default_rng(8).normal(0, 3.0, (3, sysd.n))
    o.set_staggered(stag)
    h = o.field()
    mu = 2.22 * sllg.MU_B
    eps = 1e-6
    for i in (0, 4, 13, 26):
        for a in range(3):
            s = s0.copy()
            s[a, i] += eps
            o.set_spins(s)
            Ep = o.reduce()[1]
            s[a, i] -= 2 * eps
            o.set_spins(s)
            Em = o.reduce()[1]
            g = -(Ep - Em) / (2 * eps) / mu
            assert abs(g - h[a, i]) < 1e-6 * max(1.0, abs(h[a, i])), (i, a, g, h[a, i])


def test_dipole_pair_and_monolayer_closed_forms():
    """Brute-force dipolar field: two moments along z on the x axis give
    H = -1e-7 mu mu_B / r^3 z; a square monolayer along z gives at its centre
    H_z -> -1e-7 mu mu_B / a^3 * 9.0336 (square-lattice sum of r^-3)."""
    a = 3e-10
    sysd, _ = W.sc_film_dip(Lx=2, Ly=1, Lz=1, precision="fp64", a=a)
    o = pyoracle.OracleContext(sysd)
    o.set_spins(np.array([[0, 0], [0, 0], [1.0, 1.0]]))
    h = o.dipolar()
    expect = -1e-7 * 2.22 * sllg.MU_B / a ** 3
    assert np.allclose(h[2], expect, rtol=1e-12) and np.abs(h[:2]).max() < 1e-12 * abs(expect)
    L = 61
    sysd, _ = W.sc_film_dip(Lx=L, Ly=L, Lz=1, precision="fp64", a=a)
    o = pyoracle.OracleContext(sysd)
    s = np.zeros((3, sysd.n))
    s[2] = 1.0
    o.set_spins(s)
    hz = o.dipolar()[2].reshape(L, L)[L // 2, L // 2]
    lattice_sum = 9.0336
    tail = 2 * math.pi / (L // 2 + 0.5)  # missing sum beyond the film edge ~ integral 2 pi r dr / r^3
    assert abs(hz / (-1e-7 * 2.22 * sllg.MU_B / a ** 3) - (lattice_sum - tail)) < 0.05


def test_langevin_paramagnet():
    """Non-interacting spins in B: <S_z> = coth x - 1/x, x = mu mu_B B / (k_B T); pins sigma
    (the thermal-field amplitude) and the Stratonovich Heun with one draw per step."""
    B, x = 10.0, 2.0
    T = 2.22 * sllg.MU_B * B / (sllg.K_B * x)
    n = 2000
    # relaxation rate lambda B / (1 + lambda^2) = 5 per reduced time unit; 6000 steps of
    # rdt = 3.5e-4 equilibrate for ~10 relaxation times before sampling
    sysd = sllg.System(n=n, species=[dict(mu=2.22, lambda_=1.0)], dt=2e-15, precision="fp64", seed=5)
    o = pyoracle.OracleContext(sysd)
    s = np.zeros((3, n))
    s[2] = 1.0
    o.set_spins(s)
    o.set_temperature(T)
    o.set_applied([0.0, 0.0, B])
    o.step(0, 6000)
    acc = []
    for k in range(60):
        o.step(6000 + 50 * k, 50)
        acc.append(o.reduce(energy=False)[0][0, 2])
    m = float(np.mean(acc))
    lang = 1 / math.tanh(x) - 1 / x
    assert abs(m - lang) < 0.01, (m, lang)


def test_spin_wave_frequency_matches_lswt():
    """Small-amplitude spin wave on an sc ferromagnet, lambda = 0, RK4: the precession
    frequency equals linear spin-wave theory omega'(k) = B + J~ sum_R (1 - cos k.R)
    (Examples/Spinwaves/LSWTCode/lswt_2spec.cpp:112-141)."""
    L, m = 16, 3
    J = 3e-21
    sysd, _, _ = W.sc_ferro(L=L, precision="fp64", K=0.0, lam=0.0, dt=2e-17, scheme="rk4", J=J)
    o = pyoracle.OracleContext(sysd)
    n = sysd.n
    x = np.arange(n) // (L * L)
    k = 2 * math.pi * m / L
    eps = 1e-3
    s = np.stack([eps * np.cos(k * x), eps * np.sin(k * x), np.ones(n)])
    s /= np.linalg.norm(s, axis=0)
    o.set_spins(s)
    B = 2.0
    o.set_applied([0, 0, B])
    steps = 2000
    o.step(0, steps)
    sf = o.get_spins()
    ph0 = np.arctan2(s[1, 0], s[0, 0])
    ph1 = np.arctan2(sf[1, 0], sf[0, 0])
    tp = sllg.GAMMA_E * 2e-17 * steps
    Jt = J / (2.22 * sllg.MU_B)
    omega = B + Jt * 2 * (1 - math.cos(k))
    dph = (ph1 - ph0) % (2 * math.pi)
    expect = (omega * tp) % (2 * math.pi)
    assert abs(dph - expect) < 1e-3 * omega * tp


# ----------------------------------------------------------------------------- four-spin term

def _fourspin_system(L=3, cpu_form=0, jq=40.0):
    sysd, s0, _ = W.sc_ferro(L=L, precision="fp64")
    xadj, jkl = W.plaquette_quartets((L, L, L))
    sysd.four = dict(xadj=xadj, jkl=jkl, jq=jq, cpu_form=cpu_form)
    return sysd, s0


def test_fourspin_field_is_minus_energy_gradient():
    """Plaquette quartets listed from every member: the four-spin field equals
    -dE/dS / (mu mu_B) of the oracle energy (a wrong neighbour index, a dropped pairing or
    a sign slip breaks the identity)."""
    sysd, s0 = _fourspin_system(L=3)
    base, _ = _fourspin_system(L=3)
    base.four = {}
    o = pyoracle.OracleContext(sysd)
    ob = pyoracle.OracleContext(base)
    o.set_spins(s0)
    ob.set_spins(s0)
    h4 = o.field() - ob.field()
    assert np.abs(h4).max() > 1.0
    mu = 2.22 * sllg.MU_B
    eps = 1e-6
    for i in (0, 5, 13, 26):
        for a in range(3):
            s = s0.copy()
            s[a, i] += eps
            o.set_spins(s)
            ob.set_spins(s)
            Ep = o.reduce()[1] - ob.reduce()[1]
            s[a, i] -= 2 * eps
            o.set_spins(s)
            ob.set_spins(s)
            Em = o.reduce()[1] - ob.reduce()[1]
            g = -(Ep - Em) / (2 * eps) / mu
            assert abs(g - h4[a, i]) < 1e-6 * max(1.0, abs(h4[a, i])), (i, a, g, h4[a, i])


def test_fourspin_uniform_state_closed_form():
    """Parallel unit spins: each quartet contributes 3 S, so H4 = -jq nq_i S with the GPU
    kernel's /3 and -3 jq nq_i S in the CPU form (src/cufields.cu:76-78 vs src/llgCPU.cpp:163-165)."""
    for cpu_form, f in ((0, 1.0), (1, 3.0)):
        sysd, s0 = _fourspin_system(L=4, cpu_form=cpu_form, jq=7.0)
        base, _ = _fourspin_system(L=4)
        base.four = {}
        u = np.array([0.36, 0.48, 0.8])
        s = np.tile(u[:, None], (1, sysd.n))
        o, ob = pyoracle.OracleContext(sysd), pyoracle.OracleContext(base)
        o.set_spins(s)
        ob.set_spins(s)
        h4 = o.field() - ob.field()
        nq = np.diff(sysd.four["xadj"]).astype(float)
        assert np.all(nq == 12)  # 12 plaquettes touch every site of a periodic sc lattice
        assert np.allclose(h4, -f * 7.0 * nq[None, :] * u[:, None], rtol=0, atol=1e-12)
