"""Bin maps of the calc_mag methods (sllg.observables.mag_bins) against direct loops that
follow src/util_mag.cpp:44-245 on a small lattice.  CPU only."""
import numpy as np
import pytest

import sllg
from sllg import observables as O


def _case():
    L = (3, 2, 4)
    mesh = sllg.mesh(sllg.unit_cell("bcc"), L)
    n = mesh.shape[0]
    atom = np.arange(n) % 2
    sub = atom.copy()
    mu = np.where(atom == 0, 2.0, 3.0)
    S = np.random.default_rng(3).normal(size=(3, n))
    return mesh, sub, atom, mu, S


def _apply(method, mesh, sub, atom, mu, S, disc=None):
    nb, b, w, scale, shape = O.mag_bins(method, mesh, sub, atom, mu, 2, disc)
    out = np.zeros((nb, 3))
    for i in range(S.shape[1]):
        out[b[i]] += (1.0 if w is None else w[i]) * S[:, i]
    if scale is not None:
        out *= scale[:, None]
    return out.reshape(shape + (3,))


@pytest.mark.parametrize("method", [1, 2, 3])
def test_layer_sums_weighted_by_mu(method):
    mesh, sub, atom, mu, S = _case()
    got = _apply(method, mesh, sub, atom, mu, S)
    ref = np.zeros((mesh[:, method - 1].max() + 1, 3))
    for i in range(S.shape[1]):
        ref[mesh[i, method - 1]] += S[:, i] * mu[i]
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("method", [5, 6, 7])
def test_layer_species_sums(method):
    mesh, sub, atom, mu, S = _case()
    got = _apply(method, mesh, sub, atom, mu, S)
    ref = np.zeros((mesh[:, method - 5].max() + 1, 2, 3))
    for i in range(S.shape[1]):
        ref[mesh[i, method - 5], sub[i]] += S[:, i]
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


def test_chequerboard_disc_and_atoms():
    mesh, sub, atom, mu, S = _case()
    n = S.shape[1]
    got = _apply(8, mesh, sub, atom, mu, S)
    l = np.zeros((2, 3))
    for i in range(n):
        l[(mesh[i].sum()) % 2] += S[:, i]
    assert np.allclose(got, l / (n * 0.5), atol=1e-12)
    disc = (2, 3, 3)
    got = _apply(9, mesh, sub, atom, mu, S, disc)
    cells = {}
    for i in range(n):
        c = tuple(int(float(mesh[i, a]) / float(disc[a])) for a in range(3))
        cells.setdefault(c, []).append(S[:, i])
    for c, v in cells.items():
        assert np.allclose(got[c], np.mean(v, axis=0), atol=1e-12)
    got = _apply(10, mesh, sub, atom, mu, S)
    assert np.allclose(got[1], S[:, 1::2].sum(axis=1), atol=1e-12)
    with pytest.raises(ValueError):
        O.mag_bins(4, mesh, sub, atom, mu, 2)
