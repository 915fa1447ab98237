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


def _from_api_calls():
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
