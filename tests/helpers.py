"""Shared builders for the tests: small systems exercising every term of the field."""
import numpy as np

import sllg
from sllg import workloads as W


def full_anis():
    """Every anisotropy term of the reference (K2 perp/par, K4 perp/par on 3 axes) plus cubic."""
    ax = np.eye(3)
    rot = np.array([[np.cos(0.3), -np.sin(0.3), 0], [np.sin(0.3), np.cos(0.3), 0], [0, 0, 1]])
    return dict(k2perp=2e-24, k2par=-1.5e-24, k4perp=0.7e-24, k4par=-0.4e-24, kc=1.2e-24,
                k2perp_dir=(1.0, 0.0, 0.0), k2par_dir=(0.0, 0.0, 1.0), k4perp_dirs=(rot @ ax).T,
                k4par_dirs=ax)


def tensor_csr_system(L=6, precision="fp64", ncomp=9, seed=5, fmt="csr", scheme="heun", noise="counter", dims=None):
    """sc lattice, NN + NNN exchange with random anisotropic/antisymmetric 3x3 tensors per
    interaction (J_ji = J_ij^T), two species on a checkerboard, all anisotropy terms."""
    rng = np.random.default_rng(seed)
    uc = sllg.unit_cell("sc")
    base, _ = sllg.shells(uc, [3e-21, 1e-21])
    dicts = []
    seen = {}
    for k in range(len(base)):
        dc = tuple(base[k].dc)
        key = tuple(-x for x in dc)
        if key in seen:
            J = seen[key].T
        else:
            J = np.eye(3) * base[k].J[0] + rng.normal(0, 0.3e-21, (3, 3))
            seen[dc] = J
        if ncomp == 3:
            J = np.diag(np.diag(J))
        dicts.append(dict(atom_i=0, atom_j=0, dc=dc, J=J))
    ints = sllg.interactions_from(dicts)
    dims = tuple(dims) if dims is not None else (L, L, L)
    xadj, adj, vals = sllg.build_csr(uc, dims, ints, [2.22], ncomp=ncomp)
    n = int(np.prod(dims))
    idx = np.arange(n)
    x, y, z = idx // (dims[1] * dims[2]), (idx // dims[2]) % dims[1], idx % dims[2]
    spec = ((x + y + z) % 2).astype(np.uint32)
    species = [dict(mu=2.22, lambda_=0.1, gamma_rel=1.0, k1u=1e-23, k1u_dir=(0, 0, 1), sublattice=0),
               dict(mu=2.22, lambda_=0.2, gamma_rel=1.3, k1u=-4e-24, k1u_dir=(0.6, 0.0, 0.8), sublattice=1)]
    if fmt == "csr":
        exch = dict(format=sllg.EXCH_CSR, xadj=xadj, adjncy=adj, vals=vals)
    elif fmt == "dia":
        offs, dv = sllg.csr_to_dia(n, xadj, adj, vals)
        exch = dict(format=sllg.EXCH_DIA, offsets=offs, vals=dv)
    else:
        cnt, ent, sv = sllg.build_stencil(uc, ints, [2.22], ncomp=ncomp)
        exch = dict(format=sllg.EXCH_STENCIL, count=cnt, entries=ent, vals=sv, dims=dims, pbc=(1, 1, 1), per_row=0)
    sysd = sllg.System(n=n, species=species, spin_species=spec, n_sublattice=2, anis=full_anis(), exch=exch,
                       dt=1e-16, scheme=scheme, precision=precision, noise=noise, update=10, seed=seed)
    return sysd, sllg.random_spins(n, seed, seed + 100)


def run_pair(sysd, s0, nsteps, T=0.0, B=(0.0, 0.0, 0.0), stag=None, xi=None, oracle_opts=0, launch=None):
    """Run the GPU handle (launch shape `launch` if given) and the oracle from the same
    state; return both contexts."""
    import pyoracle

    g = sllg.Context(sysd)
    if launch is not None:
        g.set_launch(*launch)
    o = pyoracle.OracleContext(sysd, oracle_opts)
    for c in (g, o):
        c.set_spins(s0)
        c.set_temperature(T)
        c.set_applied(np.asarray(B, float))
        if stag is not None:
            c.set_staggered(stag)
        if xi is not None:
            c.push_noise(xi)
        c.step(0, nsteps)
    return g, o
