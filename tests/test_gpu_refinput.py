"""Reference-format inputs through the HIP path: configuration, unit-cell and exchange files
written here in the reference's formats are read by load_reference_input and stepped on the
GPU through the C ABI, against the FP64 oracle on the same System.

  bcc NN        two-atom bcc cell, 8 nearest neighbours in the "unitcell" exchange file
                (joules), uniaxial anisotropy from the unit-cell file, applied field, random
                start, Heun (the shape of Examples/MvsT/bccFeNN)
  Mn2Au-like    four atoms on a 2 x 2 x 6 mesh, exchange in mRy (in-plane ferromagnetic,
                interlayer antiferromagnetic), K2 perp/par and K4 terms, a staggered field per
                sublattice, the unit cell's own spins as the start (the shape of the
                reference's cff / Mn2Au_ua.ucf / Mn2Au_ua.exch), Heun and RK4
  four-spin     simple cubic with the square plaquettes as FourSpinExchange_0 quartets

Bars: FP64 plain = oracle to 1e-10 per spin component (the two give the same bits), FP32
within 1e-4 per component and 1e-5 in <M>."""
import numpy as np
import pytest

import pyoracle
import sllg
from sllg import refinput as R

from test_fourspin_reader import _exch_text, _plaquette_perms

pytestmark = pytest.mark.gpu


def _tensor_block(k, J):
    t = ""
    for r in range(3):
        row = [0.0, 0.0, 0.0]
        row[r] = J
        t += f" J{k}_{r + 1} = [{row[0]!r}, {row[1]!r}, {row[2]!r}];\n"
    return t


def _write_bcc(d):
    (d / "bcc.ucf").write_text("2\n2\n0 0.0 0.0 0.0 2.0 0.2 1.0 Fe 0.0 0.0 1.0 1e-24 0 0 1\n"
                               "1 0.5 0.5 0.5 2.0 0.2 1.0 Co 0.0 0.0 1.0 1e-24 0 0 1\n")
    ex = "exchange: { MaxInteractions = 8; };\nexchange_0_0: { Num_Interactions = 0; units = \"joules\"; };\n"
    ex += "exchange_1_1: { Num_Interactions = 0; units = \"joules\"; };\n"
    for (a, b, sgn) in ((0, 1, -1), (1, 0, 1)):
        ex += f"exchange_{a}_{b}:\n{{\n Num_Interactions = 8;\n units = \"joules\";\n"
        k = 0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    k += 1
                    ex += f" UnitCell{k} = [{float(sgn * dx)}, {float(sgn * dy)}, {float(sgn * dz)}];\n"
                    ex += _tensor_block(k, 7.0e-21)
        ex += "};\n"
    (d / "bcc.exch").write_text(ex)
    (d / "cff").write_text("""Seed = 123123123;
system: { dim = [6, 6, 6]; Unitcellfile = "bcc.ucf"; abc = [3e-10, 3e-10, 3e-10];
          rprim1 = [1.0, 0.0, 0.0]; rprim2 = [0.0, 1.0, 0.0]; rprim3 = [0.0, 0.0, 1.0];
          Include_dipole = FALSE; Exchange_method = "CSR"; Periodic_Boundaries = [TRUE, TRUE, TRUE]; };
exchange: { Read_exch_method = "unitcell"; Exchfile = "bcc.exch"; OutputExchange = TRUE; };
llg: { dt = 1e-16; intscheme = "heun"; update = 100; applied = [0.0, 0.1, 0.5]; SpinConfigMethod = "random"; };
sim: { sim_type = "MvT"; };
mvt: { Lower_temp = 100.0; Upper_temp = 1200.0; Temp_step = 100.0; MinRunTime = 1e-12; MaxRunTime = 2e-12; };
""")
    return str(d / "cff")


def _write_mn2au(d):
    (d / "m.ucf").write_text("4\n4\n"
                             "0 0.5 0.5 0.166 1 1 1 4.0 0.05 1.0 MnA -1.0 0.0 0.1 5e-29 0.0 0.0 1.0\n"
                             "1 0.0 0.0 0.333 0 0 2 4.0 0.05 1.0 MnB 1.0 0.0 0.1 5e-29 0.0 0.0 1.0\n"
                             "2 0.0 0.0 0.666 0 0 4 4.0 0.05 1.0 MnA -1.0 0.1 0.0 5e-29 0.0 0.0 1.0\n"
                             "3 0.5 0.5 0.833 1 1 5 4.0 0.05 1.0 MnB 1.0 -0.1 0.0 5e-29 0.0 0.0 1.0\n")
    ex = "exchange: { MaxInteractions = 4; Scale = [1.0, 1.0, 1.0]; FourSpin = FALSE; };\n"
    plane = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]
    for a in range(4):
        ex += f"exchange_{a}_{a}: {{\n Num_Interactions = 4;\n units = \"mRy\";\n"
        for k, v in enumerate(plane, 1):
            ex += f" UnitCell{k} = [{float(v[0])}, {float(v[1])}, {float(v[2])}];\n" + _tensor_block(k, 0.6484)
        ex += "};\n"
    # interlayer: atoms 0-1 and 2-3 (antiferromagnetic), 1-2 (ferromagnetic, weaker)
    pairs = {(0, 1): ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], -5.3422),
             (1, 0): ([(0, 0, 0), (-1, 0, 0), (-1, -1, 0), (0, -1, 0)], -5.3422),
             (2, 3): ([(0, 0, 0), (-1, 0, 0), (-1, -1, 0), (0, -1, 0)], -5.3422),
             (3, 2): ([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], -5.3422),
             (1, 2): ([(0, 0, 0)], 1.2), (2, 1): ([(0, 0, 0)], 1.2)}
    for (a, b), (vecs, J) in pairs.items():
        ex += f"exchange_{a}_{b}: {{\n Num_Interactions = {len(vecs)};\n units = \"mRy\";\n"
        for k, v in enumerate(vecs, 1):
            ex += f" UnitCell{k} = [{float(v[0])}, {float(v[1])}, {float(v[2])}];\n" + _tensor_block(k, J)
        ex += "};\n"
    (d / "m.exch").write_text(ex)
    (d / "cff").write_text("""Seed = 123456;
system: { dim = [4, 4, 3]; Unitcellfile = "m.ucf"; abc = [3.328e-10, 3.328e-10, 8.579e-10]; Nm = [2, 2, 6];
          Include_dipole = FALSE; Exchange_method = "CSR"; Periodic_Boundaries = [TRUE, TRUE, TRUE]; };
exchange: { Read_exch_method = "unitcell"; Exchfile = "m.exch"; ReadExchangeMatrix = FALSE; };
llg: { dt = 1e-16; intscheme = "heun"; update = 10;
       staggered_field: { sapplied1 = [-0.3, 0.0, 0.0]; sapplied2 = [0.3, 0.0, 0.0];
                          sapplied3 = [-0.3, 0.0, 0.0]; sapplied4 = [0.3, 0.0, 0.0]; }; };
anis: { k2perp = -1.33e-22; k2perpdir = [0.0, 0.0, 1.0]; k2par = 13.02e-25; k2pardir = [0.0, 1.0, 0.0];
        k4par = 1.86e-25; k4pardir1 = [1.0, 0.0, 0.0]; k4pardir2 = [0.0, 1.0, 0.0]; k4pardir3 = [0.0, 0.0, 0.0];
        k4perp = 3.72e-25; k4perpdir1 = [0.0, 0.0, 1.0]; k4perpdir2 = [0.0, 0.0, 0.0]; k4perpdir3 = [0.0, 0.0, 0.0]; };
sim: { sim_type = "simple"; };
simple: { EquilibrationTime = 0e-12; RunTime = 10e-12; Temperature = 1e-6; };
""")
    return str(d / "cff")


def _write_fourspin(d):
    (d / "sc.ucf").write_text("1\n1\n0 0.0 0.0 0.0 2.22 0.1 1.0 Fe 0.0 0.0 1.0 1e-23 0 0 1\n")
    ex = "exchange: { MaxInteractions = 6; FourSpin = true; MaxQuartetPermutations = 12; };\n"
    ex += "exchange_0_0: {\n Num_Interactions = 6;\n units = \"J\";\n"
    for k, v in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], 1):
        ex += f" UnitCell{k} = [{float(v[0])}, {float(v[1])}, {float(v[2])}];\n" + _tensor_block(k, 3.0e-21)
    ex += "};\n" + _exch_text(_plaquette_perms(), jq=4e-22).split("\n", 1)[1]
    (d / "sc.exch").write_text(ex)
    (d / "cff").write_text("""Seed = 42;
system: { dim = [6, 6, 6]; Unitcellfile = "sc.ucf"; abc = [3e-10, 3e-10, 3e-10];
          Periodic_Boundaries = [TRUE, TRUE, TRUE]; Include_dipole = FALSE; };
exchange: { Read_exch_method = "unitcell"; Exchfile = "sc.exch"; ReadExchangeMatrix = FALSE; };
llg: { dt = 1e-16; intscheme = "heun"; update = 100; applied = [0.0, 0.0, 0.2]; SpinConfigMethod = "random"; };
""")
    return str(d / "cff")


def _run(ri, steps, T=0.0, oracle=True):
    sysd = ri.system
    out = []
    for mk in ((lambda: sllg.Context(sysd)), (lambda: pyoracle.OracleContext(sysd))) if oracle else \
            ((lambda: sllg.Context(sysd)),):
        c = mk()
        c.set_spins(ri.spins)
        c.set_temperature(T)
        c.set_applied(ri.applied)
        if ri.staggered is not None:
            c.set_staggered(ri.staggered)
        c.step(0, steps)
        out.append((c.get_spins(), c.reduce(energy=False)[0], c.field()))
        c.close()
    return out


def _maxdiff(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


WRITERS = {"bcc": _write_bcc, "mn2au": _write_mn2au, "fourspin": _write_fourspin}


@pytest.mark.parametrize("case", ["bcc", "mn2au", "fourspin"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reference_input_on_gpu_vs_oracle(gpu, tmp_path, case, prec):
    ri = R.load_reference_input(WRITERS[case](tmp_path), precision=prec)
    if case == "fourspin":
        assert ri.system.four and len(ri.system.four["jkl"]) == 12 * ri.system.n
    if case == "mn2au":
        assert ri.n_atoms == 4 and ri.staggered is not None and ri.system.exch["n_atoms"] == 4
    (sg, mg, hg), (so, mo, ho) = _run(ri, 300)
    d, dm = _maxdiff(sg, so), _maxdiff(mg, mo)
    print(f"{case} {prec}: 300 Heun steps, n={ri.system.n}: max|dS| {d:.3e} max|d<M>| {dm:.3e}")
    if prec == "fp64":
        assert d <= 1e-10 and dm <= 1e-10
        assert _maxdiff(hg, ho) <= 1e-9 * float(np.abs(ho).max())
    else:
        assert d <= 1e-4 and dm <= 1e-5


def test_reference_input_rk4_and_thermal(gpu, tmp_path):
    """The Mn2Au-like input with RK4 (FP64 vs oracle), and at 50 K with the same injected
    noise on both (Heun, FP64)."""
    cff = _write_mn2au(tmp_path)
    ri = R.load_reference_input(cff, precision="fp64", scheme="rk4")
    (sg, _, _), (so, _, _) = _run(ri, 200)
    assert _maxdiff(sg, so) <= 1e-10
    ri = R.load_reference_input(cff, precision="fp64", noise="injected")
    sysd = ri.system
    xi = np.random.default_rng(3).standard_normal((100, 3, sysd.n))
    res = []
    for c in (sllg.Context(sysd), pyoracle.OracleContext(sysd)):
        c.set_spins(ri.spins)
        c.set_temperature(50.0)
        c.set_staggered(ri.staggered)
        c.push_noise(xi)
        c.step(0, 100)
        res.append(c.get_spins())
        c.close()
    assert _maxdiff(res[0], res[1]) <= 1e-10
