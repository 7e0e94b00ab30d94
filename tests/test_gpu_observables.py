"""Device binned magnetisation sums (llg_set_bins / llg_reduce_bins) for every calc_mag
method with a bin form, against numpy sums of the spins read back, on both paths."""
import numpy as np
import pytest

import sllg
from sllg import observables as O
from sllg import workloads as W

pytestmark = pytest.mark.gpu


def _system(prec):
    L = 6
    sysd, s0 = W.fcc_3shell(L=L, precision=prec)
    sysd.species = [dict(mu=2.22 if t != 2 else 1.5, lambda_=0.05, sublattice=t % 2) for t in range(4)]
    sysd.spin_species = (np.arange(sysd.n) % 4).astype(np.uint32)
    sysd.n_sublattice = 2
    mesh = sllg.mesh(sllg.unit_cell("fcc"), (L, L, L))
    atom = np.arange(sysd.n) % 4
    sub = atom % 2
    mu = np.where(atom == 2, 1.5, 2.22)
    return sysd, s0, mesh, sub, atom, mu


def _numpy_method(method, S, mesh, sub, atom, mu, nsub, disc):
    nb, b, w, scale, shape = O.mag_bins(method, mesh, sub, atom, mu, nsub, disc)
    ww = np.ones(S.shape[1]) if w is None else w
    out = np.zeros((nb, 3))
    for c in range(3):
        out[:, c] = np.bincount(b, weights=ww * S[c], minlength=nb)
    if scale is not None:
        out *= scale[:, None]
    return out.reshape(shape + (3,))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_binned_sums_all_methods(gpu, prec):
    sysd, s0, mesh, sub, atom, mu = _system(prec)
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(200.0)
    g.step(0, 30)
    S = g.get_spins()
    for method in (1, 2, 3, 5, 6, 7, 8, 9, 10):
        dev = O.reduce_method(g, method, mesh, sub, atom, mu, 2, disc_size=(3, 2, 5))
        ref = _numpy_method(method, S, mesh, sub, atom, mu, 2, (3, 2, 5))
        assert dev.shape == ref.shape
        err = np.abs(dev - ref).max() / max(1.0, np.abs(ref).max())
        print(prec, method, dev.shape, err)
        assert err <= 1e-13, (method, err)
    # method 10 sums equal the per-atom sums of llg_reduce's per-sublattice means x counts
    m, _ = g.reduce(energy=False)
    per_atom = O.reduce_method(g, 10, mesh, sub, atom, mu, 2)
    n_per = sysd.n // 4
    for s in range(2):
        tot = per_atom[s] + per_atom[s + 2]
        assert np.allclose(tot / (2 * n_per), m[s], rtol=0, atol=1e-14)


def test_bins_skip_and_state(gpu):
    sysd, s0, mesh, sub, atom, mu = _system("fp32")
    g = sllg.Context(sysd)
    g.set_spins(s0)
    with pytest.raises(sllg.LLGError):
        g._nbins = 1
        g.reduce_bins()
    b = np.where(np.arange(sysd.n) % 3 == 0, 0, 7).astype(np.uint32)  # 7 >= nbins: left out
    g.set_bins(1, b)
    r = g.reduce_bins()
    assert np.allclose(r[0], s0[:, ::3].sum(axis=1), atol=1e-12)


@pytest.mark.parametrize("prec,axis", [("fp64", 2), ("fp32", 0)])
def test_structure_factor_matches_fft(gpu, prec, axis):
    """Device S(k) at chosen k equals numpy's FFT of the transverse spin grid read there
    (the reference's FFTW_FORWARD plan over the dim x Nk mesh)."""
    sysd, s0, mesh, sub, atom, mu = _system(prec)
    grid = tuple(int(x) for x in mesh.max(axis=0) + 1)
    g = sllg.Context(sysd)
    g.set_spins(s0)
    g.set_temperature(100.0)
    g.step(0, 20)
    S = g.get_spins()
    rng = np.random.default_rng(5)
    kpts = np.stack([rng.integers(0, grid[d], 40) for d in range(3)], axis=1)
    kpts[0] = 0
    ua = np.where(sub == 0, 1.0, -1.0)
    ub = np.where(atom == 3, -1.0, 1.0)
    a, b = {0: (1, 2), 1: (0, 2), 2: (0, 1)}[axis]
    # total and per sublattice
    g.set_sq(grid, mesh, kpts, axis=axis, ua=ua, ub=ub)
    tot = g.sample_sq()[0]
    g.set_sq(grid, mesh, kpts, axis=axis, ua=ua, ub=ub, ngroups=2, group=sub)
    per = g.sample_sq()
    for grp, dev in ((None, tot), (0, per[0]), (1, per[1])):
        sel = np.ones(sysd.n, bool) if grp is None else (sub == grp)
        s3d = np.zeros(grid, complex)
        m = mesh[sel]
        s3d[m[:, 0], m[:, 1], m[:, 2]] = S[a][sel] * ua[sel] + 1j * S[b][sel] * ub[sel]
        ref = np.fft.fftn(s3d)[kpts[:, 0], kpts[:, 1], kpts[:, 2]]
        err = np.abs(dev - ref).max() / max(1.0, np.abs(ref).max())
        print(prec, axis, grp, err)
        assert err <= 1e-11
    assert np.allclose(per.sum(axis=0), tot, rtol=0, atol=1e-9)
    with pytest.raises(sllg.LLGError):
        g.set_sq(grid, mesh, np.array([[grid[0], 0, 0]]), axis=axis)


def test_mvt_driver_on_device(tmp_path):
    """The M(T) driver on an 8^3 sc ferromagnet (FP32, lambda = 0.5, ordered start): one
    MvT.dat line per temperature and |m| falling with T (Tc ~ 313 K for J = 3e-21 J)."""
    from sllg import drivers

    sysd, _, _ = W.sc_ferro(L=8, precision="fp32", lam=0.5, dt=1e-15)
    g = sllg.Context(sysd)
    s = np.zeros((3, sysd.n))
    s[2] = 1.0
    g.set_spins(s)
    out = tmp_path / "MvT.dat"
    rows = drivers.mvt(g, 0.0, 400.0, 100.0, dt=1e-15, update=50, min_run_time=2e-12, max_run_time=8e-12,
                       mean_tol=1e-4, var_tol=1e-5, path=str(out))
    lines = out.read_text().splitlines()
    assert lines[0] == "#Temperature\tMean" and len(lines) == 5
    m = [r["modm"][0] for r in rows]
    print(rows)
    assert abs(m[0] - 1.0) < 1e-5  # T = 0 stays ordered (no noise, anisotropy along z)
    assert m[0] > m[1] > m[2] > m[3]
    assert all(r["steps"] >= int(2e-12 / 1e-15) for r in rows)  # 1999: int() truncates as the reference does
