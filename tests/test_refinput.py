"""Readers of the reference's input files (include/llg_io.h): libconfig-format parser,
unit-cell file, "unitcell" exchange file with energy units, csr_exch_mat.dat, and the
configuration -> System mapping.  CPU only."""
import os

import numpy as np
import pytest

import sllg
from sllg import refinput as R
from sllg import workloads as W

MUB = 9.27e-24
REF = "/root/reference"


def test_cfg_grammar():
    text = r'''
    # comment
    Seed = 123456;          // trailing comment
    /* block
       comment */
    system:
    {
        dim = [ 3 , 3 , 3 ];
        abc = [3.328e-10, 3.328e-10, 8.579e-10];
        Periodic_Boundaries = [ TRUE , true , False ];
        name = "a\tb" " c";
        big = 0x1FL;
        neg = -1.33e-22;
        plus = +2;
        nested : { x = 1.5; };
        lst = ( 1, "two", [3, 4], { y = 5; } );
    };
    '''
    c = R.Config(text=text)
    assert c.get("Seed") == 123456 and c.type("Seed") == R.CFG_INT
    assert c.get("system.dim") == [3, 3, 3]
    assert c.get("system.abc") == [3.328e-10, 3.328e-10, 8.579e-10]
    assert c.get("system.Periodic_Boundaries") == [1.0, 1.0, 0.0]
    assert c.type("system.Periodic_Boundaries") == R.CFG_ARRAY
    assert c.get("system.name") == "a\tb c"
    assert c.get("system.big") == 31
    assert c.get("system.neg") == -1.33e-22 and c.type("system.neg") == R.CFG_FLOAT
    assert c.get("system.plus") == 2
    assert c.get("system.nested.x") == 1.5
    assert c.type("system.lst") == R.CFG_LIST and c.length("system.lst") == 4
    assert c.number("system.lst", 0) == 1.0
    assert c.children("system")[:3] == ["dim", "abc", "Periodic_Boundaries"]
    assert c.children("") == ["Seed", "system"]
    assert c.get("system.missing", 7) == 7 and not c.exists("nosuch.path")
    with pytest.raises(sllg.LLGError):
        c.number("system.name")
    with pytest.raises(sllg.LLGError):
        c.number("system.dim", 3)


@pytest.mark.parametrize("bad,line", [("a = [1, 2;\n", 1), ("x = 1;\ny = ;\n", 2), ("g : { a = 1;\n", 2),
                                      ("a = 1.2.3;", 1), ("a = \"open", 1)])
def test_cfg_syntax_errors_name_the_line(bad, line):
    with pytest.raises(sllg.LLGError) as e:
        R.Config(text=bad)
    assert f"line {line}" in str(e.value)


UCF_PLAIN = """2
2
0   0.0   0.0   0.0   7.63 0.01 1.0 Gd  0.0 0.0 1 1e-23 0 0 1
1   0.5   0.5   0.5   1.92 0.02 2.0 Fe  1.0 0.0 0 0 1 0 0
"""
UCF_MESH = """2
2
0 0.0 0.0 0.0 0 0 0 7.63 0.01 1.0 Gd 0.0 0.0 1 1e-23 0 0 1
1 0.5 0.5 0.5 1 1 1 1.92 0.02 2.0 Fe 1.0 0.0 0 0 1 0 0
"""


def test_read_ucf(tmp_path):
    p = tmp_path / "a.ucf"
    p.write_text(UCF_PLAIN)
    ns, atoms = R.read_ucf(str(p))
    assert ns == 2 and len(atoms) == 2
    a, b = atoms
    assert a.sublattice == 0 and list(a.xred) == [0, 0, 0] and a.mu == 7.63 and a.lambda_ == 0.01
    assert a.element == b"Gd" and list(a.spin0) == [0, 0, 1] and a.k1u == 1e-23 and list(a.k1u_dir) == [0, 0, 1]
    assert b.sublattice == 1 and list(b.xred) == [.5, .5, .5] and b.gamma_rel == 2.0 and list(b.k1u_dir) == [1, 0, 0]
    p2 = tmp_path / "m.ucf"
    p2.write_text(UCF_MESH)
    _, atoms = R.read_ucf(str(p2), with_mesh=True)
    assert list(atoms[1].mp) == [1, 1, 1] and atoms[1].mu == 1.92
    # the plain file read as a meshed one is rejected (columns shift: mu becomes 'Gd')
    with pytest.raises(sllg.LLGError):
        R.read_ucf(str(p), with_mesh=True)
    p3 = tmp_path / "bad.ucf"
    p3.write_text(UCF_PLAIN.replace("7.63", "1e-7"))
    with pytest.raises(sllg.LLGError) as e:
        R.read_ucf(str(p3))
    assert "magnetic moment" in str(e.value)


EXCH = """
exchange: { MaxInteractions = 3; Scale = [1.0, 1.0, 1.0]; FourSpin = FALSE; };
exchange_0_0:
{
    Num_Interactions = 2;
    units = "mRy";
    UnitCell1 = [ 0.5 , -0.5 , 0.49 ];
    J1_1 = [ 1.0 , 0.1 , 0.0 ];
    J1_2 = [ -0.1 , 1.0 , 0.0 ];
    J1_3 = [ 0.0 , 0.0 , 2.0 ];
    UnitCell2 = [ -1.0 , 0.0 , 0.0 ];
    J2_1 = [ 3.0 , 0.0 , 0.0 ];
    J2_2 = [ 0.0 , 3.0 , 0.0 ];
    J2_3 = [ 0.0 , 0.0 , 3.0 ];
};
exchange_1_0:
{
    Num_Interactions = 1;
    units = "meV";
    UnitCell1 = [ 0.0 , 0.0 , 1.0 ];
    J1_1 = [ 1.0 , 0.0 , 0.0 ];
    J1_2 = [ 0.0 , 1.0 , 0.0 ];
    J1_3 = [ 0.0 , 0.0 , 1.0 ];
};
exchange_0_1:
{
    Num_Interactions = 1;
    UnitCell1 = [ 0.0 , 1.0 , 0.0 ];
    J1_1 = [ 1e-21 , 0.0 , 0.0 ];
    J1_2 = [ 0.0 , 1e-21 , 0.0 ];
    J1_3 = [ 0.0 , 0.0 , 1e-21 ];
};
"""


def test_read_exch_unitcell_units_order_rounding():
    ints = R.read_exch_unitcell(R.Config(text=EXCH), 2)
    assert len(ints) == 4
    key = [(i.atom_i, i.atom_j, i.shell) for i in ints]
    assert key == [(0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)]  # atom_i, atom_j, file order
    assert list(ints[0].dc) == [1, -1, 0]  # half away from zero; 0.49 -> 0
    assert list(ints[1].dc) == [-1, 0, 0]
    mry = 1.0 * 2.179872172e-18 * 1.0e-3
    J0 = np.array(ints[0].J).reshape(3, 3)
    assert J0[0, 0] == mry and J0[0, 1] == 0.1 * 2.179872172e-18 * 1.0e-3 and J0[1, 0] == -0.1 * 2.179872172e-18 * 1e-3
    assert J0[2, 2] == 2.0 * 2.179872172e-18 * 1.0e-3
    assert ints[2].J[0] == 1e-21  # no units: joules
    assert ints[3].J[4] == 1.602176565e-22
    bad = EXCH.replace('units = "meV"', 'units = "kcal"')
    with pytest.raises(sllg.LLGError):
        R.read_exch_unitcell(R.Config(text=bad), 2)
    with pytest.raises(sllg.LLGError):
        R.read_exch_unitcell(R.Config(text=EXCH.replace("J2_3", "J2_9")), 2)


def test_csr_file_round_trip_and_reference_layout(tmp_path):
    sysd, _, _ = W.sc_ferro(L=4, precision="fp64", ncomp=3)
    e = sysd.exch
    p = tmp_path / "csr_exch_mat.dat"
    R.csr_write(str(p), e["xadj"], e["adjncy"], e["vals"])
    x2, a2, v2 = R.csr_read(str(p), sysd.n)
    assert np.array_equal(x2, e["xadj"]) and np.array_equal(a2, e["adjncy"]) and np.array_equal(v2, e["vals"])
    with pytest.raises(sllg.LLGError):
        R.csr_read(str(p), sysd.n + 1)
    # the reference's own writer: tab-separated, default 6-digit values (src/exch_unitcell.cpp:45-58)
    q = tmp_path / "ref.dat"
    q.write_text("3\n4\n0\t1\t3\t4\n1\t145.778\t145.778\t145.778\t0\t1\t2\t3\t2\t-1e-05\t0\t0\t"
                 "1\t4\t5\t6\t")
    x, a, v = R.csr_read(str(q), 3)
    assert list(x) == [0, 1, 3, 4] and list(a) == [1, 0, 2, 1]
    assert v[0, 0] == 145.778 and v[:, 1].tolist() == [1.0, 2.0, 3.0] and v[0, 2] == -1e-05
    assert v[:, 3].tolist() == [4.0, 5.0, 6.0]
    q.write_text("3\n4\n0\t1\t3\t5\n1\t1\t1\t1\t")
    with pytest.raises(sllg.LLGError):
        R.csr_read(str(q), 3)


def _write_case(tmp_path, cff_extra=""):
    (tmp_path / "sc.ucf").write_text("1\n1\n0 0.0 0.0 0.0 2.22 0.1 1.0 Fe 0.0 0.0 1.0 1e-23 0 0 1\n")
    ex = "exchange: { MaxInteractions = 6; };\nexchange_0_0: {\n Num_Interactions = 6;\n units = \"J\";\n"
    vecs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for k, v in enumerate(vecs, 1):
        ex += f" UnitCell{k} = [{v[0]}.0, {v[1]}.0, {v[2]}.0];\n"
        for r in range(3):
            row = [0.0, 0.0, 0.0]
            row[r] = 3.0e-21
            ex += f" J{k}_{r + 1} = [{row[0]}, {row[1]}, {row[2]}];\n"
    ex += "};\n"
    (tmp_path / "sc.exch").write_text(ex)
    cff = f"""Seed = 42;
system: {{ dim = [4, 4, 4]; Unitcellfile = "sc.ucf"; abc = [3e-10, 3e-10, 3e-10];
           Periodic_Boundaries = [TRUE, TRUE, TRUE]; Include_dipole = FALSE; Exchange_method = "CSR"; }};
exchange: {{ Read_exch_method = "unitcell"; Exchfile = "sc.exch"; ReadExchangeMatrix = FALSE; }};
llg: {{ dt = 1e-16; intscheme = "RK4"; update = 100; applied = [0.0, 0.0, 0.5]; SpinConfigMethod = "random"; }};
anis: {{ k2perp = -1e-24; k2perpdir = [0.0, 0.0, 1.0]; }};
sim: {{ sim_type = "simple"; }};
simple: {{ Temperature = 300.0; RunTime = 1e-12; }};
{cff_extra}
"""
    (tmp_path / "cff").write_text(cff)
    return str(tmp_path / "cff")


def test_load_reference_input_matches_builder(tmp_path):
    ri = R.load_reference_input(_write_case(tmp_path), precision="fp64")
    s = ri.system
    ref, s0, _ = W.sc_ferro(L=4, precision="fp64")
    assert s.n == 64 and ri.dims == (4, 4, 4) and ri.n_atoms == 1
    assert s.scheme == "rk4" and s.update == 100 and s.dt == 1e-16 and s.seed == 42
    assert ri.applied == (0.0, 0.0, 0.5) and ri.staggered is None
    assert ri.sim == {"sim_type": "simple", "Temperature": 300.0, "RunTime": 1e-12}
    assert s.anis["k2perp"] == -1e-24 and tuple(s.anis["k2perp_dir"]) == (0.0, 0.0, 1.0)
    sp = s.species[0]
    assert sp["mu"] == 2.22 and sp["lambda_"] == 0.1 and sp["k1u"] == 1e-23
    # same couplings per row as the distance-shell builder (order may differ): J/(mu mu_B)
    e, r = s.exch, ref.exch
    assert np.array_equal(e["xadj"], r["xadj"])
    for i in range(s.n):
        a = sorted(zip(e["adjncy"][e["xadj"][i]:e["xadj"][i + 1]], e["vals"][0][e["xadj"][i]:e["xadj"][i + 1]]))
        b = sorted(zip(r["adjncy"][r["xadj"][i]:r["xadj"][i + 1]], r["vals"][0][r["xadj"][i]:r["xadj"][i + 1]]))
        assert a == b
    assert e["vals"][0][0] == 3.0e-21 / (2.22 * MUB)
    # reference-compatible random spins (MWC1616 seeded Seed, Seed + 100)
    assert np.array_equal(ri.spins, s0)
    # the matrix written by the reference's "OutputExchangeMatrix" path is read back instead
    p = tmp_path / "csr_exch_mat.dat"
    R.csr_write(str(p), e["xadj"], e["adjncy"], np.vstack([e["vals"]] * 3))
    cff = _write_case(tmp_path).replace("ReadExchangeMatrix = FALSE;",
                                        "ReadExchangeMatrix = TRUE; ExchangeMatrixFilename = \"csr_exch_mat.dat\";")
    (tmp_path / "cff2").write_text(open(cff).read().replace("ReadExchangeMatrix = FALSE;",
                                   "ReadExchangeMatrix = TRUE; ExchangeMatrixFilename = \"csr_exch_mat.dat\";"))
    ri2 = R.load_reference_input(str(tmp_path / "cff2"), precision="fp64")
    assert np.array_equal(ri2.system.exch["adjncy"], e["adjncy"])
    assert np.array_equal(ri2.system.exch["vals"], e["vals"])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "cff")), reason="reference example inputs not present")
def test_reference_example_inputs_parse():
    """The reference's own example inputs (Mn2Au: 4-atom cell on a mesh, mRy tensors)."""
    c = R.Config(os.path.join(REF, "cff"))
    assert c.get("system.dim") == [3, 3, 3] and c.get("system.Nm") == [2, 2, 6]
    assert c.get("exchange.Read_exch_method") == "unitcell"
    assert c.get("llg.staggered_field.sapplied2") == [0.3, 0.0, 0.0]
    ns, atoms = R.read_ucf(os.path.join(REF, "Mn2Au_ua.ucf"), with_mesh=True)
    assert ns == 4 and len(atoms) == 4 and list(atoms[2].mp) == [0, 0, 4] and atoms[0].element == b"MnA"
    ints = R.read_exch_unitcell(R.Config(os.path.join(REF, "Mn2Au_ua.exch")), 4)
    assert len(ints) > 0 and all(i.atom_i < 4 and i.atom_j < 4 for i in ints)
    assert ints[0].J[0] == 0.6484 * 2.179872172e-18 * 1.0e-3


def test_spin_files_round_trip_and_reference_layout(tmp_path):
    n = 8 * 4
    mesh = sllg.mesh(sllg.unit_cell("bcc"), (2, 2, 4))
    s = np.random.default_rng(2).normal(size=(3, n))
    s /= np.linalg.norm(s, axis=0)
    p = tmp_path / "restart.txt"
    R.write_spins_text(str(p), s)
    assert np.array_equal(R.read_spins_text(str(p), n), s / np.linalg.norm(s, axis=0))
    # the reference writes 6 significant digits and normalises on reading
    q = tmp_path / "r6.txt"
    q.write_text("2\n0 0 2\n3\t4\t0\n")
    r = R.read_spins_text(str(q), 2)
    assert np.allclose(r[:, 0], [0, 0, 1]) and np.allclose(r[:, 1], [0.6, 0.8, 0])
    q.write_text("2\n0 0 0\n1 0 0\n")
    with pytest.raises(sllg.LLGError):
        R.read_spins_text(str(q), 2)
    with pytest.raises(sllg.LLGError):
        R.read_spins_text(str(p), n + 1)
    # grid: written in spin order, read back in any line order
    g = tmp_path / "dw_grid.out"
    R.write_grid(str(g), mesh, s)
    lines = g.read_text().splitlines()
    body = lines[1:]
    np.random.default_rng(4).shuffle(body)
    g.write_text("\n".join([lines[0]] + body) + "\n")
    assert np.allclose(R.read_grid(str(g), mesh), s, rtol=0, atol=1e-15)
    g.write_text(f"{n}\n" + "\n".join(body[:-1] + ["99 99 99 1 0 0"]) + "\n")
    with pytest.raises(sllg.LLGError):
        R.read_grid(str(g), mesh)


def test_vtu_is_well_formed(tmp_path):
    import xml.etree.ElementTree as ET
    n = 10
    pos = np.arange(3 * n, dtype=float).reshape(n, 3)
    s = np.tile(np.array([[0.0], [0.6], [0.8]]), (1, n))
    p = tmp_path / "spinmap_0.vtu"
    R.write_vtu(str(p), pos, s)
    root = ET.parse(str(p)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == str(n)
    spin = np.array(piece.find("PointData/DataArray").text.split(), float).reshape(n, 3)
    pts = np.array(piece.find("Points/DataArray").text.split(), float).reshape(n, 3)
    assert np.allclose(spin, s.T) and np.array_equal(pts, pos)
