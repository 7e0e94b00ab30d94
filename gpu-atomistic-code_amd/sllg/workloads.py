"""Synthetic workloads of the stochastic-LLG hot path, built through the C lattice builders.

Each builder returns (System, initial spins[3, n]).  Sizes are parameters so tests can
run the same construction at small sizes.

  sc_ferro      simple-cubic, nearest-neighbour J = 3.0e-21 J (6 nnz/row CSR), uniaxial K = 1e-23 J
                along z, mu = 2.22 mu_B, lambda = 0.1, dt = 1e-16 s (the 32^3 FP64 plain case).
  fcc_3shell    fcc conventional cells x 4 atoms, exchange to the 3rd shell (12 + 6 + 24 = 42 nnz/row,
                J1 = 5e-21, J2 = 1e-21, J3 = -4e-22 J), cubic anisotropy Kc = 5e-24 J, lambda = 0.05,
                mu = 2.22 mu_B (not given by the case; stated here), dt = 1e-16 s (the 1,048,576-spin case).
  bct_afm_dmi   bct (c/a = 1.2) x 2 sites, 8 body-diagonal neighbours, 3x3 tensors J_iso = -4e-21 J plus
                antisymmetric DMI |D| = 0.3e-21 J along a per-bond random in-plane axis (J_ji = J_ij^T),
                easy-plane anisotropy K = -5e-24 J along z, periodic stencil with per-row tensors,
                Neel state + Gaussian tilt of 1e-3 rad, dt = 5e-17 s, RK4.
  sc_film_dip   simple-cubic film (PBC in-plane, open in z), NN J = 3e-21 J, dipolar on the zero-padded
                (2 Lx, 2 Ly, 2 Lz) grid, a = 3e-10 m, lambda = 0.1, dt = 1e-16 s.
"""
from __future__ import annotations

import numpy as np

from . import (EXCH_CSR, EXCH_STENCIL, System, build_csr, build_stencil, interactions_from, mesh, random_spins,
               shells, unit_cell)

MU = 2.22


def _species(mu=MU, lam=0.1, k1u=0.0, k1u_dir=(0, 0, 1), sub=0):
    return dict(mu=mu, lambda_=lam, gamma_rel=1.0, k1u=k1u, k1u_dir=k1u_dir, sublattice=sub)


def sc_ferro(L=32, precision="fp64", T=300.0, J=3.0e-21, K=1.0e-23, lam=0.1, dt=1e-16, seed=42, update=100,
             noise="counter", scheme="heun", ncomp=1, pbc=(1, 1, 1), dims=None):
    dims = tuple(dims) if dims is not None else (L, L, L)
    uc = unit_cell("sc")
    ints, _ = shells(uc, [J])
    xadj, adj, vals = build_csr(uc, dims, ints, [MU], ncomp=ncomp, pbc=pbc)
    n = int(np.prod(dims))
    sysd = System(n=n, species=[_species(lam=lam, k1u=K)], exch=dict(format=EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals),
                  dt=dt, scheme=scheme, precision=precision, noise=noise, update=update, seed=seed)
    return sysd, random_spins(n, seed, seed + 100), T


def fcc_3shell(L=64, precision="fp32", lam=0.05, dt=1e-16, Kc=5e-24, seed=42, update=100, noise="counter",
               scheme="heun", ncomp=1, J=(5e-21, 1e-21, -4e-22), dims=None):
    uc = unit_cell("fcc")
    ints, _ = shells(uc, list(J))
    dims = tuple(dims) if dims is not None else (L, L, L)
    xadj, adj, vals = build_csr(uc, dims, ints, [MU] * 4, ncomp=ncomp)
    n = int(np.prod(dims)) * 4
    sysd = System(n=n, species=[_species(lam=lam)], anis=dict(kc=Kc),
                  exch=dict(format=EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals, n_atoms=4), dt=dt, scheme=scheme,
                  precision=precision, noise=noise, update=update, seed=seed)
    return sysd, random_spins(n, seed, seed + 100)


def dmi_tensor(D):
    """Antisymmetric 3x3 with (A S)_a = (D x S)_a."""
    Dx, Dy, Dz = D
    return np.array([[0.0, -Dz, Dy], [Dz, 0.0, -Dx], [-Dy, Dx, 0.0]])


def bct_afm_dmi(Lx=96, Ly=96, Lz=48, precision="fp32", c_over_a=1.2, Jiso=-4e-21, Dmag=0.3e-21, K=-5e-24,
                lam=0.1, dt=5e-17, seed=7, tilt=1e-3, scheme="rk4", update=100):
    """Two-sublattice antiferromagnet with per-bond DMI as a periodic stencil with per-row tensors."""
    uc = unit_cell("bct", c_over_a)
    ints, _ = shells(uc, [1.0])  # 8 body-diagonal neighbours per site (geometry only)
    count, ent, _ = build_stencil(uc, ints, [MU, MU], ncomp=9)
    dims = (Lx, Ly, Lz)
    na = 2
    n = Lx * Ly * Lz * na
    k = ent.shape[1]
    idx = np.arange(n)
    t = idx % na
    r = idx // na
    z = r % Lz
    r //= Lz
    y = r % Ly
    x = r // Ly
    vals = np.zeros((9, k, n))
    rng = np.random.default_rng(seed)
    # per-bond random in-plane DMI axis keyed by the unordered bond: theta(bond) from a hash
    # of (min(i,j), max(i,j)); D_ij = sign * |D| (cos th, sin th, 0) with sign +1 for i<j, -1 else,
    # so J_ji = J_ij^T.
    base = rng.integers(1, 2**31 - 1)
    den = MU * 9.27e-24
    for q in range(k):
        e = ent[t, q]  # (n, 4)
        nx = (x + e[:, 0]) % Lx
        ny = (y + e[:, 1]) % Ly
        nz = (z + e[:, 2]) % Lz
        j = ((nx * Ly + ny) * Lz + nz) * na + e[:, 3]
        lo = np.minimum(idx, j).astype(np.uint64)
        hi = np.maximum(idx, j).astype(np.uint64)
        h = (lo * np.uint64(0x9E3779B97F4A7C15) ^ (hi + np.uint64(base)) * np.uint64(0xBF58476D1CE4E5B9))
        h ^= h >> np.uint64(31)
        th = (h % np.uint64(1 << 24)).astype(np.float64) / float(1 << 24) * 2 * np.pi
        sgn = np.where(idx < j, 1.0, -1.0)
        Dx = sgn * Dmag * np.cos(th)
        Dy = sgn * Dmag * np.sin(th)
        Dz = np.zeros(n)
        T = np.zeros((n, 3, 3))
        T[:, 0, 0] = T[:, 1, 1] = T[:, 2, 2] = Jiso
        T[:, 0, 1] = -Dz; T[:, 0, 2] = Dy
        T[:, 1, 0] = Dz; T[:, 1, 2] = -Dx
        T[:, 2, 0] = -Dy; T[:, 2, 1] = Dx
        vals[:, q, :] = (T.reshape(n, 9) / den).T
    sp = [_species(lam=lam, k1u=K, sub=0), _species(lam=lam, k1u=K, sub=1)]
    sysd = System(n=n, species=sp, spin_species=(idx % na).astype(np.uint32), n_sublattice=2,
                  exch=dict(format=EXCH_STENCIL, count=count, entries=ent, vals=vals, dims=dims, pbc=(1, 1, 1),
                            per_row=1),
                  dt=dt, scheme=scheme, precision=precision, noise="counter", update=update, seed=seed)
    # Neel state along x plus a Gaussian tilt of `tilt` rad
    s = np.zeros((3, n))
    sign = np.where(t == 0, 1.0, -1.0)
    g = np.random.default_rng(seed + 1)
    a1 = g.normal(0.0, tilt, n)
    a2 = g.normal(0.0, tilt, n)
    s[0] = sign * np.cos(a1) * np.cos(a2)
    s[1] = np.sin(a1) * np.cos(a2)
    s[2] = np.sin(a2)
    s /= np.linalg.norm(s, axis=0)
    return sysd, s


def sc_film_dip(Lx=128, Ly=128, Lz=16, precision="fp32", J=3e-21, a=3e-10, lam=0.1, dt=1e-16, seed=3, update=100,
                every_stage=0, refresh=0, scheme="heun", stream=0):
    uc = unit_cell("sc")
    ints, _ = shells(uc, [J])
    dims = (Lx, Ly, Lz)
    xadj, adj, vals = build_csr(uc, dims, ints, [MU], pbc=(1, 1, 0))
    n = Lx * Ly * Lz
    m = mesh(uc, dims)
    sysd = System(n=n, species=[_species(lam=lam)], exch=dict(format=EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals),
                  dip=dict(enabled=1, mesh=m, dim=dims, nk=(1, 1, 1), abc=(a, a, a), every_stage=every_stage,
                           refresh=refresh),
                  dt=dt, scheme=scheme, precision=precision, noise="counter", update=update, seed=seed,
                  stream=stream)
    return sysd, random_spins(n, seed, seed + 100)


def sc_film_fftx(Lx=128, Ly=128, Lz=16, precision="fp32", J=3e-21, a=3e-10, lam=0.1, dt=1e-16, seed=3, update=100,
                 dipolar=True, as_csr=False):
    """Open sc film with nearest-neighbour exchange J as FFT exchange (the reference's
    Exchange_method "fft": open boundaries, convolution every stage; dipolar terms in the
    same kernel when `dipolar`), or with as_csr the same physics as CSR exchange plus the
    dipolar convolution refreshed every stage."""
    uc = unit_cell("sc")
    dims = (Lx, Ly, Lz)
    n = Lx * Ly * Lz
    m = mesh(uc, dims)
    dip = dict(enabled=1 if dipolar else 0, mesh=m, dim=dims, nk=(1, 1, 1), abc=(a, a, a), every_stage=1, refresh=1)
    common = dict(n=n, species=[_species(lam=lam)], dip=dip, dt=dt, precision=precision, noise="counter",
                  update=update, seed=seed)
    if as_csr:
        ints, _ = shells(uc, [J])
        xadj, adj, vals = build_csr(uc, dims, ints, [MU], pbc=(0, 0, 0))
        sysd = System(exch=dict(format=EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals), **common)
    else:
        Jt = np.eye(3) * J
        vecs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
        sysd = System(fftx=[(0, 0, v, Jt) for v in vecs], **common)
    return sysd, random_spins(n, seed, seed + 100)


def plaquette_quartets(dims, pbc=(1, 1, 1)):
    """Four-spin quartets of a simple-cubic lattice (spin index (x*d1 + y)*d2 + z): every
    elementary square plaquette of the xy, yz and zx planes, listed once from the point of
    view of each of its four spins (j, k, l = the other three in cyclic order).  Returns
    (xadj [n+1], jkl [nq, 3]) for System.four."""
    d0, d1, d2 = dims
    n = d0 * d1 * d2

    def idx(x, y, z):
        return (x * d1 + y) * d2 + z

    rows = [[] for _ in range(n)]
    planes = ((0, 1), (1, 2), (2, 0))
    for x in range(d0):
        for y in range(d1):
            for z in range(d2):
                for (a, b) in planes:
                    corners = []
                    ok = True
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        c = [x, y, z]
                        c[a] += da
                        c[b] += db
                        for ax in range(3):
                            if c[ax] >= dims[ax]:
                                if not pbc[ax]:
                                    ok = False
                                c[ax] %= dims[ax]
                        corners.append(idx(*c))
                    if not ok or len(set(corners)) < 4:
                        continue
                    for m in range(4):
                        rows[corners[m]].append([corners[(m + 1) % 4], corners[(m + 2) % 4], corners[(m + 3) % 4]])
    xadj = np.zeros(n + 1, np.uint32)
    xadj[1:] = np.cumsum([len(r) for r in rows])
    jkl = np.array([q for r in rows for q in r], np.uint32).reshape(-1, 3)
    return xadj, jkl


def quartets_from_sets(n, sets):
    """xadj/jkl of the four-spin term for arbitrary 4-spin sets, each listed from every member."""
    rows = [[] for _ in range(n)]
    for q in sets:
        q = [int(v) for v in q]
        for m in range(4):
            rows[q[m]].append([q[(m + 1) % 4], q[(m + 2) % 4], q[(m + 3) % 4]])
    xadj = np.zeros(n + 1, np.uint32)
    xadj[1:] = np.cumsum([len(r) for r in rows])
    jkl = np.array([v for r in rows for v in r], np.uint32).reshape(-1, 3)
    return xadj, jkl
