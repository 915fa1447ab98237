"""CHARMM input files (CPU): the reference's own water216 test deck
(test/water216: minimized.psf, minimized.crd and
test/toppar_water_ions/par_water_ions.prm, kept as fixtures under
tests/golden/water216) read through blade_read_parameter_file /
blade_read_psf / blade_read_crd.

The system read from files must equal the one built by direct blade_*
calls from an independent parse of the same files, term for term.
"""
import os

import numpy as np

from blade_amd import system
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
W216 = os.path.join(HERE, "golden", "water216")
PRM = os.path.join(W216, "par_water_ions.prm")
PSF = os.path.join(W216, "minimized.psf")
CRD = os.path.join(W216, "minimized.crd")
BOX = (18.856, 18.856, 18.856)


def _from_files():
    s = system.load_charmm(PRM, PSF, CRD, BOX, beta=1 / 3.125, rcut=9.0, rswitch=8.0)
    s.prepare()
    return s


def _parse_psf(path):
    lines = open(path).read().split("\n")
    k = 2
    ntitle = int(lines[k].split()[0])
    k += ntitle + 2
    natom = int(lines[k].split()[0])
    atoms = [lines[k + 1 + i].split() for i in range(natom)]
    k += natom + 2

    def section(k, width):
        n = int(lines[k].split()[0])
        vals = []
        k += 1
        while len(vals) < width * n:
            vals += [int(v) for v in lines[k].split()]
            k += 1
        return k + 1, [tuple(vals[width * i:width * i + width]) for i in range(n)]
    k, bonds = section(k, 2)
    k, angles = section(k, 3)
    return atoms, bonds, angles


def _from_api_calls(prepare=True):
    atoms, bonds, angles = _parse_psf(PSF)
    s = system.BladeSystem((0,))
    s.init_parameters()
    s.add_parameter_atoms("OT", 15.99940)
    s.add_parameter_atoms("HT", 1.00800)
    s.add_parameter_bonds("HT", "HT", 0.0, 1.5139)
    s.add_parameter_bonds("HT", "OT", 450.0, 0.9572)
    s.add_parameter_angles("HT", "OT", "HT", 55.0, 104.52, 0.0, 0.0)
    s.add_parameter_nbonds("HT", -0.046, 0.2245, -0.046, 0.2245, 1.0, 0)
    s.add_parameter_nbonds("OT", -0.1521, 1.7682, -0.1521, 1.7682, 1.0, 0)
    s.init_structure()
    for a in atoms:
        s.add_atom(int(a[0]), a[1], a[2], a[3], a[4], a[5], float(a[6]), float(a[7]))
    for i, j in bonds:
        s.add_bond(i, j)
    for i, j, k in angles:
        s.add_angle(i, j, k)
    s.add_shake(1)
    crd = [l.split() for l in open(CRD).read().split("\n") if l and not l.startswith("*")][1:]
    s.init_coordinates(len(atoms))
    for r in crd:
        s.add_coordinates_position(int(r[0]), float(r[4]), float(r[5]), float(r[6]))
    s.add_coordinates_box(0, *BOX)
    s.init_msld(0)
    s.init_run()
    from blade_amd.synthetic import PICOSECOND
    s.add_run_flags(1.0 / PICOSECOND, 1 / 3.125, 9.0, 8.0, 0, 1, 1.0, -1, -1, -1, 6, 2e-7)
    s.add_run_dynopts(0, 0, 0, 0.002 * PICOSECOND, 300.0)
    s.n = len(atoms)
    s.nblocks = 1
    if prepare:
        s.prepare()
    return s


def test_water216_files_equal_api_calls():
    a, b = _from_files(), _from_api_calls()
    ta, tb = a.topology().contents, b.topology().contents
    assert ta.n_atoms == tb.n_atoms == 648
    for name, cnt, w in (("charge", "n_atoms", 1), ("mass", "n_atoms", 1), ("vdw_c12", None, 1),
                         ("vdw_c6", None, 1), ("bond_idx", "n_bonds", 2), ("angle_idx", "n_angles", 3),
                         ("excl_idx", "n_excl", 2), ("nbex_idx", "n_nbex", 2), ("tri_idx", "n_tri", 3),
                         ("tri_b0", "n_tri", 3)):
        if cnt is None:
            n = ta.n_types * ta.n_types
            assert ta.n_types == tb.n_types
        else:
            n = getattr(ta, cnt)
            assert n == getattr(tb, cnt), name
        assert np.array_equal(ta.array(name, n, w), tb.array(name, n, w)), name
    assert ta.n_tri == 216 and ta.n_br1 == ta.n_br2 == ta.n_br3 == 0
    assert np.array_equal(a.positions(), b.positions())


def test_water216_minimized_energy():
    s = _from_files()
    top, rp = s.topology(), s.runparams()
    f, _, e, pairs = orc.forces(top, rp, s.positions(), np.ones(1))
    per_water = e[20] / 216
    # a minimized TIP3P box: about -9 to -11 kcal/mol per water
    assert -12.0 < per_water < -8.0, per_water
    assert rp.contents.grid[0] == 20      # L / gridSpace = 18.9 -> 20


# A five-atom chain whose two dihedrals (1-2-3-4, 2-3-4-5) carry a CMAP
# cross term: the PRM's CMAP section (eight types, the grid size, then
# ngrid*ngrid values phi-major over as many lines as they take, with
# comment lines between rows, as CHARMM36 writes them) read by
# blade_read_parameter_file, the PSF's NCRTERM section by blade_read_psf.
_CHAIN = ("CA1", "CA2", "CA3", "CA4", "CA5")
_NG = 8


def _cmap_values():
    P, Q = np.meshgrid(np.arange(_NG), np.arange(_NG), indexing="ij")
    return np.round(0.37 * P - 0.11 * Q + 0.05 * P * Q - 1.25, 4)   # distinct, phi rows


def _write_chain_files(d):
    g = _cmap_values()
    prm = ["* chain with a CMAP term", "*", "", "ATOMS"]
    prm += ["MASS  -1  %-6s 12.01100" % t for t in _CHAIN]
    prm += ["", "BONDS"] + ["%s %s 300.0 1.53" % (_CHAIN[k], _CHAIN[k + 1]) for k in range(4)]
    prm += ["", "ANGLES"] + ["%s %s %s 50.0 112.0" % _CHAIN[k:k + 3] for k in range(3)]
    prm += ["", "CMAP", "", "%s %s %s %s %s %s %s %s %d" % (_CHAIN[:4] + _CHAIN[1:] + (_NG,)), ""]
    for i in range(_NG):
        prm.append("!phi = %.1f" % (-180.0 + i * 360.0 / _NG))
        row = ["%.4f" % v for v in g[i]]
        prm += [" ".join(row[:5]), " ".join(row[5:]), ""]
    prm += ["NONBONDED nbxmod  5 atom cdiel fshift vatom vdistance vfswitch -",
            "cutnb 14.0 ctofnb 12.0 ctonnb 10.0 eps 1.0 e14fac 1.0 wmin 1.5", ""]
    prm += ["%s 0.0 -0.0800 2.0000" % t for t in _CHAIN]
    prm += ["", "END", ""]
    psf = ["PSF EXT CMAP XPLOR", "", "         1 !NTITLE", "* chain", "", "         5 !NATOM"]
    q = (0.1, -0.2, 0.15, -0.1, 0.05)
    for k in range(5):
        psf.append("%10d CH       1        CHN      A%-7d %-6s %14.6f %14.4f           0"
                   % (k + 1, k + 1, _CHAIN[k], q[k], 12.011))
    psf += ["", "         4 !NBOND: bonds", "         1         2         2         3         3         4"
            "         4         5", ""]
    psf += ["         3 !NTHETA: angles", "         1         2         3         2         3         4"
            "         3         4         5", ""]
    psf += ["         0 !NPHI: dihedrals", "", "", "         0 !NIMPHI: impropers", "", "",
            "         0 !NDON: donors", "", "", "         0 !NACC: acceptors", "", "", "         0 !NNB", "",
            "         0         0         0         0         0", "", "         1         0 !NGRP NST2",
            "         0         0         0", "", "         0         0 !NUMLP NUMLPH", "",
            "         1 !NCRTERM: cross-terms",
            "         1         2         3         4         2         3         4         5", ""]
    open(os.path.join(d, "chain.prm"), "w").write("\n".join(prm))
    open(os.path.join(d, "chain.psf"), "w").write("\n".join(psf))


def _chain_finish(s):
    xyz = [(0.0, 0.0, 0.0), (1.53, 0.0, 0.0), (2.1, 1.4, 0.0), (3.6, 1.5, 0.3), (4.2, 2.8, 1.1)]
    s.add_shake(0)
    s.init_coordinates(5)
    for k, (x, y, z) in enumerate(xyz):
        s.add_coordinates_position(k + 1, x + 10.0, y + 10.0, z + 10.0)
    s.add_coordinates_box(0, 30.0, 30.0, 30.0)
    s.init_msld(0)
    s.init_run()
    from blade_amd.synthetic import PICOSECOND
    s.add_run_flags(1.0 / PICOSECOND, 1 / 3.125, 9.0, 8.0, 0, 1, 1.0, -1, -1, -1, 6, 2e-7)
    s.add_run_dynopts(0, 0, 0, 0.002 * PICOSECOND, 300.0)
    s.n, s.nblocks = 5, 1
    s.prepare()
    return s


def test_cmap_section_of_parameter_file(tmp_path):
    _write_chain_files(str(tmp_path))
    a = system.BladeSystem((0,))
    a.init_parameters()
    a.read_parameter_file(str(tmp_path / "chain.prm"))
    a.read_psf(str(tmp_path / "chain.psf"))
    a = _chain_finish(a)

    g = _cmap_values()
    b = system.BladeSystem((0,))
    b.init_parameters()
    for t in _CHAIN:
        b.add_parameter_atoms(t, 12.011)
        b.add_parameter_nbonds(t, -0.08, 2.0, -0.08, 2.0, 1.0, 0)
    for k in range(4):
        b.add_parameter_bonds(_CHAIN[k], _CHAIN[k + 1], 300.0, 1.53)
    for k in range(3):
        b.add_parameter_angles(*_CHAIN[k:k + 3], 50.0, 112.0, 0.0, 0.0)
    key = [t.encode() for t in _CHAIN[:4] + _CHAIN[1:]]
    b.L.blade_add_parameter_cmaps(b.h, *key, _NG)
    for i in range(_NG):
        for j in range(_NG):
            b.L.blade_add_parameter_cmaps_fill(b.h, *key, i + 1, j + 1, float(g[i, j]))
    b.init_structure()
    q = (0.1, -0.2, 0.15, -0.1, 0.05)
    for k in range(5):
        b.add_atom(k + 1, "CH", "1", "CHN", "A%d" % (k + 1), _CHAIN[k], q[k], 12.011)
    for k in range(4):
        b.add_bond(k + 1, k + 2)
    for k in range(3):
        b.add_angle(k + 1, k + 2, k + 3)
    b.L.blade_add_cmap(b.h, 1, 2, 3, 4, 2, 3, 4, 5)
    b = _chain_finish(b)

    ta, tb = a.topology().contents, b.topology().contents
    assert ta.n_cmaps == tb.n_cmaps == 1 and ta.n_cmap_types == 1
    assert ta.cmap_ngrid[0] == _NG
    grid = np.array([ta.cmap_grid[ta.cmap_off[0] + k] for k in range(_NG * _NG)]).reshape(_NG, _NG)
    assert np.array_equal(grid, g)
    assert [ta.cmap_idx[k] for k in range(8)] == [tb.cmap_idx[k] for k in range(8)] == [0, 1, 2, 3, 1, 2, 3, 4]
    ea = orc.forces(a.topology(), a.runparams(), a.positions(), np.ones(1))[2]
    eb = orc.forces(b.topology(), b.runparams(), b.positions(), np.ones(1))[2]
    assert ea[5] != 0.0 and ea[5] == eb[5]
    assert np.array_equal(ea, eb)
