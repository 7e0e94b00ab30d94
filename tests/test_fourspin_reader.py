"""Reader of the reference's four-spin quartets (llg_four_spin_quartets; exch::read4spin and
exch::setup4SpinCSR, src/exch_read_4_spin.cpp:29-258) and of 4spin_csr_exch_mat.dat
(:260-381).  CPU only.

Pins: the square plaquettes of a simple-cubic lattice written as FourSpinExchange_0
permutations give exactly the plaquette quartets a direct construction lists (periodic and
open boundaries); JQ is converted to tesla as J / (mu mu_B); the reference's rounding of
negative offsets is reproduced when asked for; an interface species admits quartets
through the other species; the oracle's four-spin field from the read quartets equals the
field from the constructed ones."""
import numpy as np
import pytest

import pyoracle
import sllg
from sllg import refinput as R
from sllg import workloads as W

MUB = 9.27e-24


def _plaquette_perms():
    """From one corner of each square plaquette (xy, yz, zx planes) the other three corners
    in cyclic order: 4 corners x 3 planes = 12 permutations."""
    sq = [(0, 0), (1, 0), (1, 1), (0, 1)]
    perms = []
    for a, b in ((0, 1), (1, 2), (2, 0)):
        for m in range(4):
            trip = []
            for d in (1, 2, 3):
                c = sq[(m + d) % 4]
                v = [0, 0, 0]
                v[a] = c[0] - sq[m][0]
                v[b] = c[1] - sq[m][1]
                trip.append(v)
            perms.append(trip)
    return perms


def _exch_text(perms, jq=1.0e-22, iface=0, nspec=1, maxq=None, scale=None):
    t = "exchange: { MaxInteractions = 0; FourSpin = true; MaxQuartetPermutations = %d;%s };\n" % (
        maxq if maxq is not None else len(perms), "" if scale is None else " Scale = [%g, %g, %g];" % tuple(scale))
    for s in range(nspec):
        t += "FourSpinExchange_%d: {\n NumPermutations = %d;\n JQ = %r;\n InterfaceSpecies = %d;\n" % (
            s, len(perms), jq, iface if s == 0 else s)
        for q, trip in enumerate(perms, 1):
            for m, v in enumerate(trip, 1):
                t += " Permute%d_%d = [%s];\n" % (q, m, ", ".join("%.1f" % x for x in v))
        t += "};\n"
    return t


def _rows(xadj, jkl):
    return [sorted(map(tuple, jkl[xadj[i]:xadj[i + 1]].tolist())) for i in range(len(xadj) - 1)]


def _sc_mesh(dims):
    uc = sllg.unit_cell("sc")
    return sllg.mesh(uc, dims)


@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 1, 0)])
def test_plaquettes_equal_direct_construction(pbc):
    dims = (4, 5, 3)
    ex = R.Config(text=_exch_text(_plaquette_perms()))
    n = int(np.prod(dims))
    xadj, jkl, jq = R.four_spin_quartets(ex, 1, [2.22], _sc_mesh(dims), np.zeros(n), dims, (1, 1, 1), pbc)
    rx, rj = W.plaquette_quartets(dims, pbc)
    assert np.array_equal(xadj, rx)
    assert _rows(xadj, jkl) == _rows(rx, rj)
    assert jq[0] == 1.0e-22 / (2.22 * MUB)


def test_first_quartet_is_the_first_permutation():
    dims = (4, 4, 4)
    ex = R.Config(text=_exch_text(_plaquette_perms()))
    xadj, jkl, _ = R.four_spin_quartets(ex, 1, [2.22], _sc_mesh(dims), np.zeros(64), dims, (1, 1, 1), (1, 1, 1))
    idx = lambda x, y, z: ((x % 4) * 4 + (y % 4)) * 4 + (z % 4)
    # spin (1, 2, 3): permutation 1 = xy plaquette, corners (+1,0), (+1,+1), (0,+1)
    i = idx(1, 2, 3)
    assert tuple(jkl[xadj[i]]) == (idx(2, 2, 3), idx(2, 3, 3), idx(1, 3, 3))
    # permutation 2 = the same plane, the spin as corner 1: (0,+1), (-1,+1), (-1,0)
    assert tuple(jkl[xadj[i] + 1]) == (idx(1, 3, 3), idx(0, 3, 3), idx(0, 2, 3))
    assert xadj[i + 1] - xadj[i] == 12


def test_reference_rounding_of_negative_offsets():
    """static_cast<int>(x + 0.5): -1 -> 0, so a quartet can name the spin itself."""
    dims = (4, 4, 4)
    ex = R.Config(text=_exch_text(_plaquette_perms()))
    args = (ex, 1, [2.22], _sc_mesh(dims), np.zeros(64), dims, (1, 1, 1), (1, 1, 1))
    xadj, jkl, _ = R.four_spin_quartets(*args, rounding="reference")
    i = 21
    # permutation 2 offsets (0,1,0), (-1,1,0), (-1,0,0) become (0,1,0), (0,1,0), (0,0,0)
    assert tuple(jkl[xadj[i] + 1]) == (i + 4, i + 4, i)
    # with a mesh of 2 points per cell a negative unit offset (-2 points) becomes -1
    ex2 = R.Config(text=_exch_text([[[-1, 0, 0], [1, 0, 0], [0, 0, 1]]]))
    grid = (8, 8, 8)
    mesh = np.array([[x, y, z] for x in range(0, 8, 2) for y in range(0, 8, 2) for z in range(0, 8, 2)], np.uint32)
    for rnd, ok in (("nearest", True), ("reference", False)):
        xa, jk, _ = R.four_spin_quartets(ex2, 1, [2.0], mesh, np.zeros(64), grid, (2, 2, 2), (1, 1, 1), rnd)
        # nearest: every spin finds its quartet; reference: -2 -> -1 lands between spins
        assert (xa[-1] == 64) == ok and (xa[-1] == 0) == (not ok)


def test_interface_species_and_scale():
    """Two species on alternating z planes (CsCl-like sc): a species-0 spin's plaquettes in
    the yz and zx planes include species-1 spins and are kept only with InterfaceSpecies = 1;
    Scale multiplies the offsets (0.5 x 2 = one cell)."""
    dims = (4, 4, 4)
    mesh = _sc_mesh(dims)
    sub = (mesh[:, 2] % 2).astype(np.uint32)
    half = [[[0.5 * c for c in v] for v in trip] for trip in _plaquette_perms()]
    for iface, want in ((0, 4), (1, 12)):
        ex = R.Config(text=_exch_text(half, iface=iface, nspec=2, scale=(2, 2, 2)))
        xadj, _, jq = R.four_spin_quartets(ex, 2, [2.0, 3.0], mesh, sub, dims, (1, 1, 1), (1, 1, 1))
        cnt = np.diff(xadj)
        assert np.all(cnt[sub == 0] == want)
        assert np.all(cnt[sub == 1] == 4)  # species 1 names itself as interface: only its own xy squares
        assert np.allclose(jq, [1e-22 / (2.0 * MUB), 1e-22 / (3.0 * MUB)], rtol=0, atol=0)


def test_errors():
    dims = (3, 3, 3)
    m, z = _sc_mesh(dims), np.zeros(27)
    bad = R.Config(text=_exch_text(_plaquette_perms(), maxq=4))
    with pytest.raises(sllg.LLGError):
        R.four_spin_quartets(bad, 1, [2.2], m, z, dims, (1, 1, 1), (1, 1, 1))
    missing = R.Config(text="exchange: { FourSpin = true; MaxQuartetPermutations = 2; };\n")
    with pytest.raises(sllg.LLGError):
        R.four_spin_quartets(missing, 1, [2.2], m, z, dims, (1, 1, 1), (1, 1, 1))
    ok = R.Config(text=_exch_text(_plaquette_perms()))
    with pytest.raises(sllg.LLGError):  # two spins on one mesh point
        R.four_spin_quartets(ok, 1, [2.2], np.zeros((27, 3)), z, dims, (1, 1, 1), (1, 1, 1))


def test_four_csr_file_round_trip(tmp_path):
    xadj, jkl = W.plaquette_quartets((3, 3, 3))
    p = str(tmp_path / "4spin_csr_exch_mat.dat")
    R.four_csr_write(p, xadj, jkl)
    x2, j2 = R.four_csr_read(p, 27)
    assert np.array_equal(x2, xadj) and np.array_equal(j2, jkl)
    head = open(p).read().split("\n")
    assert head[0] == "27" and head[1] == str(len(jkl))
    with pytest.raises(sllg.LLGError):
        R.four_csr_read(p, 26)


def test_oracle_field_from_read_quartets():
    """The oracle's four-spin field (GPU form, /3) from the read quartets equals the one from
    the direct construction (same quartets, summed in another order)."""
    dims = (4, 4, 4)
    sysd, s0, _ = W.sc_ferro(L=4, precision="fp64")
    ex = R.Config(text=_exch_text(_plaquette_perms(), jq=2e-22))
    xadj, jkl, jq = R.four_spin_quartets(ex, 1, [2.22], _sc_mesh(dims), np.zeros(64), dims, (1, 1, 1), (1, 1, 1))
    out = []
    for fx, fj in ((xadj, jkl), W.plaquette_quartets(dims)):
        sysd.four = dict(xadj=fx, jkl=fj, jq=float(jq[0]), cpu_form=0)
        o = pyoracle.OracleContext(sysd)
        o.set_spins(s0)
        out.append(o.field())
        o.close()
    assert np.allclose(out[0], out[1], rtol=0, atol=1e-12 * np.abs(out[1]).max())


def test_load_reference_input_reads_four_spin(tmp_path):
    (tmp_path / "sc.ucf").write_text("1\n1\n0 0.0 0.0 0.0 2.22 0.1 1.0 Fe 0.0 0.0 1.0 1e-23 0 0 1\n")
    ex = "exchange: { MaxInteractions = 6; FourSpin = true; MaxQuartetPermutations = 12; };\n"
    ex += "exchange_0_0: {\n Num_Interactions = 6;\n units = \"J\";\n"
    for k, v in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], 1):
        ex += " UnitCell%d = [%d.0, %d.0, %d.0];\n" % ((k,) + v)
        for r in range(3):
            row = [0.0, 0.0, 0.0]
            row[r] = 3.0e-21
            ex += " J%d_%d = [%r, %r, %r];\n" % (k, r + 1, *row)
    ex += "};\n" + _exch_text(_plaquette_perms(), jq=5e-22).split("\n", 1)[1]
    (tmp_path / "sc.exch").write_text(ex)
    (tmp_path / "cff").write_text("""Seed = 42;
system: { dim = [4, 4, 4]; Unitcellfile = "sc.ucf"; abc = [3e-10, 3e-10, 3e-10];
          Periodic_Boundaries = [TRUE, TRUE, TRUE]; Include_dipole = FALSE; };
exchange: { Read_exch_method = "unitcell"; Exchfile = "sc.exch"; ReadExchangeMatrix = FALSE; };
llg: { dt = 1e-16; intscheme = "heun"; update = 100; SpinConfigMethod = "random"; };
""")
    ri = R.load_reference_input(str(tmp_path / "cff"), precision="fp64")
    f = ri.system.four
    rx, rj = W.plaquette_quartets((4, 4, 4))
    assert _rows(f["xadj"], f["jkl"]) == _rows(rx, rj)
    assert f["jq"] == 5e-22 / (2.22 * MUB) and f["cpu_form"] == 0
    # ReadExchangeMatrix4Spin: the quartets come from 4spin_csr_exch_mat.dat instead
    xs = np.zeros(65, np.uint32)
    xs[1:] = 1
    R.four_csr_write(str(tmp_path / "4spin_csr_exch_mat.dat"), xs, [[1, 2, 3]])
    (tmp_path / "cff").write_text(open(tmp_path / "cff").read().replace(
        "ReadExchangeMatrix = FALSE;", "ReadExchangeMatrix = FALSE; ReadExchangeMatrix4Spin = TRUE;"))
    ri2 = R.load_reference_input(str(tmp_path / "cff"), precision="fp64")
    assert np.array_equal(ri2.system.four["xadj"], xs) and ri2.system.four["jkl"].tolist() == [[1, 2, 3]]
