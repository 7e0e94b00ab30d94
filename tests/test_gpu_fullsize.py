"""GPU checks at the benchmark sizes, through properties that do not depend on size.

The oracle takes minutes per step at these sizes, so the reference here is the FP64 plain
path, which the smaller parity tests pin bit for bit to the oracle (test_gpu_parity.py).
Tolerances: FP32 vs FP64 trajectories at T = 0 within 1e-4 per spin component and 1e-5 in
<M>; unit spin norms to FP32 rounding; the damped energy at T = 0 never rises."""
import numpy as np
import pytest

import sllg
from sllg import workloads as W

pytestmark = pytest.mark.gpu


def _fcc(prec, **kw):
    return W.fcc_3shell(L=64, precision=prec, **kw)  # 64^3 x 4 = 1,048,576 spins, 42 nnz/row


def test_fcc_1m_fp32_tracks_fp64_plain(gpu):
    """The 1M-spin fcc system of the headline benchmark (line-mode encoding, tuned launch):
    200 Heun steps at T = 0 from the same random state, FP32 fused path vs FP64 plain path."""
    sf, s0 = _fcc("fp32")
    sd, _ = _fcc("fp64")
    out = []
    for sysd in (sf, sd):
        g = sllg.Context(sysd)
        if sysd is sf:
            assert "line64" in g.info()["exch_encoding"]
            g.autotune(5)
        g.set_spins(s0)
        g.set_temperature(0.0)
        g.step(0, 200)
        out.append((g.get_spins(), g.reduce(energy=False)[0].mean(axis=0)))
        g.close()
    d = float(np.abs(out[0][0] - out[1][0]).max())
    dm = float(np.abs(out[0][1] - out[1][1]).max())
    print("1M fcc, 200 steps T=0: max|dS| FP32 vs FP64", d, "|d<M>|", dm)
    assert d <= 1e-4 and dm <= 1e-5


def test_fcc_1m_tile_reader_equals_line_reader(gpu):
    """The headline lattice's reader choices at full size: the LDS-staged line reader (tile
    streams, 8 slices per workgroup) gives the line reader's bits over 60 thermal steps at
    300 K, with and without the thermal hand-off and across a set_spins; both stay within FP32
    rounding of the packed entries."""
    sf, s0 = _fcc("fp32")
    g = sllg.Context(sf)
    enc = g.info()["exch_encoding"]
    assert "line64" in enc and "-tile8x" in enc, enc
    res = {}
    for name, launch in (("line", (256, 1, 2, 0, 1, 1)), ("tile", (512, 1, 2, 0, 1, 2)),
                         ("tile_h", (512, 0, 2, 1, 1, 2)), ("packed", (256, 1, 2, 0, 1, 0))):
        g.set_launch(*launch)
        g.set_spins(s0)
        g.set_temperature(300.0)
        g.step(0, 30)
        g.set_spins(g.get_spins()[:, ::-1].copy())
        g.step(30, 30)
        res[name] = g.get_spins()
    g.close()
    assert np.array_equal(res["tile"], res["line"])
    assert np.array_equal(res["tile_h"], res["line"])
    d = float(np.abs(res["tile"] - res["packed"]).max())
    print("1M fcc tile vs packed, 60 steps at 300 K: max|dS|", d)
    assert d <= 2e-5


@pytest.mark.timeout(900)
def test_fcc_1m_fp32_ordered_start_10k_steps(gpu):
    """The accuracy target of the FP32 form at full size: 10^4 Heun steps at T = 0 of the 1M-spin
    fcc system from an ordered start (+z with a Gaussian tilt of 0.05 rad per transverse
    component: damped spin waves), FP32 fused path (tuned launch, line reader) vs FP64 plain
    path: <= 1e-4 per spin component and <= 1e-5 in <M>.  (The uniformly random start, a
    chaotic quench at T = 0, is measured by bench.py --mode accuracy beside FP64 controls.)"""
    sf, s0 = _fcc("fp32")
    sd, _ = _fcc("fp64")
    rng = np.random.default_rng(11)
    so = np.zeros_like(s0)
    so[0] = 0.05 * rng.standard_normal(sf.n)
    so[1] = 0.05 * rng.standard_normal(sf.n)
    so[2] = 1.0
    so /= np.linalg.norm(so, axis=0)
    out = []
    for sysd in (sf, sd):
        g = sllg.Context(sysd)
        if sysd is sf:
            g.autotune(5)
        g.set_spins(so)
        g.set_temperature(0.0)
        for k in range(10):
            g.step(1000 * k, 1000)
            print(f"{sysd.precision} step {1000 * (k + 1)}", flush=True)
        out.append((g.get_spins(), g.reduce(energy=False)[0].mean(axis=0)))
        g.close()
    d = float(np.abs(out[0][0] - out[1][0]).max())
    dm = float(np.abs(out[0][1] - out[1][1]).max())
    print("1M fcc ordered start, 10^4 steps T=0: max|dS| FP32 vs FP64", d, "|d<M>|", dm)
    assert d <= 1e-4 and dm <= 1e-5


def test_fcc_1m_norms_energy_and_reproducibility(gpu):
    """At full size: spin norms stay 1 to FP32 rounding, the energy at T = 0 (damped, no
    field) does not rise over 4 blocks of 50 steps, and a thermal run repeated from the same
    state and seed is bit-identical; the line reader and the packed entries agree to FP32
    rounding after 40 thermal steps."""
    sysd, s0 = _fcc("fp32")
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(0.0)
    e_prev = g.reduce()[1]
    for b in range(4):
        g.step(50 * b, 50)
        e = g.reduce()[1]
        assert e <= e_prev + 1e-7 * abs(e_prev), (b, e, e_prev)
        e_prev = e
    s = g.get_spins()
    nrm = np.linalg.norm(s, axis=0)
    assert float(np.abs(nrm - 1.0).max()) <= 2e-6
    runs = []
    for line in (1, 1, 0):
        g.set_launch(512, 0, 2, 0, 1, line)
        g.set_spins(s0)
        g.set_temperature(300.0)
        g.step(0, 40)
        runs.append(g.get_spins())
    g.close()
    assert np.array_equal(runs[0], runs[1])
    d = float(np.abs(runs[0] - runs[2]).max())
    print("1M fcc, 40 thermal steps: line reader vs packed entries max|dS|", d)
    assert d <= 2e-5


def test_film_full_size_dipolar_fp32_tracks_fp64(gpu):
    """One member of the batched benchmark at its real shape (128 x 128 x 16 sc film, NN
    exchange, dipolar on the 256 x 256 x 32 padded grid, refreshed every step): the FP32
    half-grid convolution against the FP64 whole-grid path over 50 steps at T = 0."""
    out = []
    for prec in ("fp32", "fp64"):
        sysd, s0 = W.sc_film_dip(Lx=128, Ly=128, Lz=16, precision=prec, refresh=1)
        g = sllg.Context(sysd)
        g.set_spins(s0)
        g.set_temperature(0.0)
        g.step(0, 50)
        out.append(g.get_spins())
        g.close()
    d = float(np.abs(out[0] - out[1]).max())
    print("film 128x128x16 with dipolar, 50 steps T=0: max|dS| FP32 vs FP64", d)
    assert d <= 1e-5
