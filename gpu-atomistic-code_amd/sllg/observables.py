"""Magnetisation observables of the reference's util::calc_mag (src/util_mag.cpp:44-245) as
bin maps for the device reduction llg_set_bins / llg_reduce_bins.

`mag_bins(method, ...)` returns (nbins, bin_of_spin, weight, scale, shape): the device sums
w_i S_i per bin, the host multiplies by `scale` (per bin, or None) and reshapes to
`shape + (3,)`, giving the reference's array for that method:

  1, 2, 3   magx/magy/magz(plane, c): sum of mu_i S_i over the mesh plane lu(i, 0/1/2)
  5, 6, 7   mag_species_x/y/z(plane, sublattice, c): sum of S_i (no weight)
  8         (l1, l2): chequerboard by parity of lu(i,0)+lu(i,1)+lu(i,2), each / (n/2)
  9         magdisc(cx, cy, cz, c): cell = int(lu / NoMagDisc), mean over the cell's spins
  10        magEachAtUC(atom, c): sum of S_i over spins of each unit-cell atom
Method 0 is llg_reduce (per-sublattice mean); method 4 computes nothing.
"""
import numpy as np


def mag_bins(method, mesh, sublattice, atom_of_spin, mu, n_sublattice, disc_size=None):
    mesh = np.asarray(mesh, dtype=np.int64)
    n = mesh.shape[0]
    sub = np.asarray(sublattice, dtype=np.int64)
    if method in (1, 2, 3):
        k = mesh[:, method - 1]
        nb = int(k.max()) + 1
        return nb, k.astype(np.uint32), np.asarray(mu, float), None, (nb,)
    if method in (5, 6, 7):
        k = mesh[:, method - 5]
        npl = int(k.max()) + 1
        return npl * n_sublattice, (k * n_sublattice + sub).astype(np.uint32), None, None, (npl, n_sublattice)
    if method == 8:
        b = (mesh.sum(axis=1) % 2).astype(np.uint32)
        return 2, b, None, np.full(2, 1.0 / (n * 0.5)), (2,)
    if method == 9:
        d = np.asarray(disc_size, dtype=np.float64)
        cell = (mesh / d).astype(np.int64)  # truncation, as static_cast<int>(double)
        mx = cell.max(axis=0) + 1
        b = (cell[:, 0] * mx[1] + cell[:, 1]) * mx[2] + cell[:, 2]
        nb = int(np.prod(mx))
        cnt = np.bincount(b, minlength=nb).astype(float)
        scale = np.where(cnt > 0, 1.0 / np.maximum(cnt, 1), 0.0)
        return nb, b.astype(np.uint32), None, scale, tuple(int(x) for x in mx)
    if method == 10:
        a = np.asarray(atom_of_spin, dtype=np.int64)
        nb = int(a.max()) + 1
        return nb, a.astype(np.uint32), None, None, (nb,)
    raise ValueError(f"calc_mag method {method} has no bin form (0 is llg_reduce, 4 computes nothing)")


def reduce_method(ctx, method, mesh, sublattice, atom_of_spin, mu, n_sublattice, disc_size=None):
    """Set the bins of `method` on ctx (once) and return the reference-shaped array."""
    nb, b, w, scale, shape = mag_bins(method, mesh, sublattice, atom_of_spin, mu, n_sublattice, disc_size)
    ctx.set_bins(nb, b, w)
    out = ctx.reduce_bins()
    if scale is not None:
        out = out * scale[:, None]
    return out.reshape(shape + (3,))
