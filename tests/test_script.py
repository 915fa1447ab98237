"""The input-script interpreter, blade_interpretter (io/io.h:69-70), on the
CPU: the language (variables, prefix calculations, nested braces,
if/elseif/else, while, stream), selections, and the reference's own
test/water216/input deck up to its dynamics, against the same system built
by blade_* calls.  The deck runs whole on the GPU in
tests/test_gpu_script.py."""
import os
import shutil

import numpy as np

from blade_amd import system
from tests import test_files as tf

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def deck_tree(tmp):
    """the reference's test/ layout (decks name ../test/<deck>/...), run
    from a sibling directory"""
    t = tmp / "test"
    (t / "water216").mkdir(parents=True)
    for f in ("minimized.psf", "minimized.crd", "input"):
        shutil.copy(os.path.join(GOLD, "water216", f), t / "water216" / f)
    (t / "toppar_water_ions").mkdir()
    shutil.copy(os.path.join(GOLD, "water216", "par_water_ions.prm"), t / "toppar_water_ions")
    (t / "toppar").mkdir()
    shutil.copy(os.path.join(GOLD, "toppar", "par_all36_prot.prm"), t / "toppar")
    run = tmp / "run"
    run.mkdir()
    return run


def dhfr_tree(tmp):
    """the reference's test/dhfr deck (23,558 atoms: DHFR in water, CHARMM22,
    kept gzip-compressed as fixtures) unpacked in the test/ layout; returns
    the directory to run from"""
    import gzip
    t = tmp / "test" / "dhfr"
    t.mkdir(parents=True)
    for f in ("5dfr.psf", "5dfr.crd", "par_all22_prot.inp"):
        with gzip.open(os.path.join(GOLD, "dhfr", f + ".gz"), "rb") as src, open(t / f, "wb") as dst:
            shutil.copyfileobj(src, dst)
    shutil.copy(os.path.join(GOLD, "dhfr", "input"), t / "input")
    run = tmp / "run"
    run.mkdir()
    return run


def interpret(s, path):
    s.L.blade_interpretter(str(path).encode(), s.h)
    return s.L.blade_system_error(s.h).decode()


def test_language_variables_control_flow_stream(tmp_path, capfd):
    (tmp_path / "sub.inp").write_text("variables calculate fromsub real * {acc} 0.5\n")
    (tmp_path / "main.inp").write_text("""! comment line
variables set n 3
variables set s1seq1 nat
variables set s1seq2 cys   ! trailing comment
variables set ii 1
variables set acc 0
while <= {ii} {n}
   if == {ii} 2
      variables calculate acc int + {acc} 10
   elseif == {ii} 3
      variables calculate acc int + {acc} 100
      if > {acc} 1000
         variables set never 1
      endif
   else
      variables calculate acc int + {acc} 1
   endif
   variables set jj 1
   while < {jj} 3
      variables calculate jj int + {jj} 1
   endwhile
   variables calculate ii int + {ii} 1
endwhile
variables set jj 2
variables set name {s1seq{jj}}
variables setupper up {name}
variables calculate expr real + sqrt 16 * 2 floor 2.7
if eq {name} cys
   variables set matched yes
else
   variables set matched no
endif
stream {here}/sub.inp
variables print
""".replace("{here}", str(tmp_path)))
    s = system.BladeSystem((0,))
    assert interpret(s, tmp_path / "main.inp") == ""
    out = capfd.readouterr().err
    got = dict(l.split("variables print> ")[1].split(" = ") for l in out.splitlines() if "variables print>" in l)
    assert got["acc"] == "111" and got["ii"] == "4" and got["jj"] == "2"
    assert got["name"] == "cys" and got["up"] == "CYS" and got["matched"] == "yes"
    assert float(got["expr"]) == 8.0 and float(got["fromsub"]) == 55.5
    assert "never" not in got


def test_errors_name_the_problem(tmp_path):
    for text, what in (("variables set a {b}\n", "{b}"), ("while < 1 2\n", "while"),
                       ("frobnicate\n", "frobnicate"), ("msld call 1 x\n", "selection x")):
        (tmp_path / "bad.inp").write_text(text)
        s = system.BladeSystem((0,))
        err = interpret(s, tmp_path / "bad.inp")
        assert what in err, (text, err)


def test_water216_deck_builds_the_same_system(tmp_path, monkeypatch):
    run = deck_tree(tmp_path)
    monkeypatch.chdir(run)
    deck = (tmp_path / "test" / "water216" / "input").read_text().splitlines()
    assert deck[-1].strip() == "run dynamics"
    (run / "setup.inp").write_text("\n".join(deck[:-1]) + "\n")
    a = system.BladeSystem((0,))
    assert interpret(a, run / "setup.inp") == ""
    a.n, a.nblocks = 648, 1
    a.prepare()
    b = tf._from_api_calls()
    ta, tb = a.topology().contents, b.topology().contents
    for name, cnt, w in (("charge", "n_atoms", 1), ("mass", "n_atoms", 1), ("bond_idx", "n_bonds", 2),
                         ("angle_idx", "n_angles", 3), ("excl_idx", "n_excl", 2), ("tri_idx", "n_tri", 3),
                         ("tri_b0", "n_tri", 3)):
        n = getattr(ta, cnt)
        assert n == getattr(tb, cnt), name
        assert np.array_equal(ta.array(name, n, w), tb.array(name, n, w)), name
    assert np.array_equal(a.positions(), b.positions())
    assert tuple(a.runparams().contents.box) == tf.BOX


def test_selections_and_msld_commands(tmp_path, capfd):
    w = os.path.join(GOLD, "water216")
    (tmp_path / "m.inp").write_text(f"""
parameters file {w}/par_water_ions.prm
structure file psf {w}/minimized.psf
structure shake hbond
coordinates file crd {w}/minimized.crd
coordinates box cubi 18.856 18.856 18.856 90 90 90
selection limit 20
selection define w1 resid 1
selection define w2 and segid WT00 resid 2
selection define hyd and hydrogen selection w1
selection count hyd nh
selection define oxy atomnames OH2
selection count oxy no
selection define some or resids 3 4 5 atom WT00 7 H1
selection count some ns
selection define pat atomsearch WT00 1% H%
selection count pat np
selection define rng residrange 10 19
selection count rng nr
variables print
msld nblocks 2
msld call 1 w1
msld call 2 w2
msld initialize 0 0 0 0 5 0 0
msld initialize 1 1 1 0 5 0.5 0
msld initialize 2 1 -1 0 5 0 0
msld bias 1 2 6 0.0 10.0 0
msld softcore on
msld fnex 5.5
msld removescaling bond angle
""")
    a = system.BladeSystem((0,))
    assert interpret(a, tmp_path / "m.inp") == ""
    out = capfd.readouterr().err
    got = dict(l.split("variables print> ")[1].split(" = ") for l in out.splitlines() if "variables print>" in l)
    # 3 atoms per water; H1 and H2 of residue 1; "1%" matches the two-digit
    # resids 10..19 (20 hydrogens), residrange 10 19 their 30 atoms
    assert got["nh"] == "2" and got["no"] == "216" and got["ns"] == "10"
    assert got["np"] == "20" and got["nr"] == "30"
    a.n, a.nblocks = 648, 3
    a.prepare()
    t = a.topology().contents
    assert t.n_blocks == 3 and t.n_varbias == 1
    assert t.vb_type[0] == 6 and t.vb_i[0] == 1 and t.vb_j[0] == 2 and t.vb_k[0] == 10.0
    sb = np.array([t.site_block[i] for i in range(648)])
    assert list(sb[:3] & 0xFFFF) == [1, 1, 1] and list(sb[3:6] & 0xFFFF) == [2, 2, 2]
    assert list(sb[:6] >> 16) == [1] * 6 and not sb[6:].any()
    assert a.runparams().contents.soft_core == 1

    # the same set-up by blade_* calls
    b = tf._from_api_calls(prepare=False)
    b.L.blade_init_msld(b.h, 2)
    for i in range(3):
        b.L.blade_add_msld_atomassignment(b.h, i + 1, 2)
        b.L.blade_add_msld_atomassignment(b.h, i + 4, 3)
    for blk, site, th, fb in ((1, 1, 0.0, 0.0), (2, 2, 1.0, 0.5), (3, 2, -1.0, 0.0)):
        b.L.blade_add_msld_initialconditions(b.h, blk, site, th, 0.0, 5.0, fb, 0.0)
    b.L.blade_add_msld_bias(b.h, 2, 3, 6, 0.0, 10.0, 0)
    b.L.blade_add_msld_flags(b.h, 1.0 / 20.454828284385908, 5.5, 1, 0, 2, 59.2, 0.1, 2.0, 1.0, 0)
    b.L.blade_add_msld_termscaling(b.h, 0, 1, 0, 1, 1, 1)
    b.L.blade_set_msld_piecewise_constraint(b.h, 0, 1.0, 100.0)
    b.nblocks = 3
    b.prepare()
    tb = b.topology().contents
    for name, cnt, w in (("site_block", "n_atoms", 1), ("bond_sb", "n_bonds", 2), ("angle_sb", "n_angles", 2),
                         ("charge", "n_atoms", 1)):
        if not hasattr(t, name):
            continue
        n = getattr(t, cnt)
        assert np.array_equal(t.array(name, n, w), tb.array(name, n, w)), name


def test_dhfr_deck_sets_up(tmp_path, monkeypatch):
    """the reference's DHFR benchmark deck up to its dynamics: CHARMM22
    protein parameters, the 23,558-atom PSF and CRD, SHAKE on bonds to
    hydrogen (7,023 waters as rigid triangles)"""
    run = dhfr_tree(tmp_path)
    monkeypatch.chdir(run)
    deck = (tmp_path / "test" / "dhfr" / "input").read_text().splitlines()
    k = [i for i, l in enumerate(deck) if l.strip().startswith("run dynamics")][0]
    (run / "setup.inp").write_text("\n".join(deck[:k]) + "\n")
    s = system.BladeSystem((0,))
    assert interpret(s, run / "setup.inp") == ""
    s.n, s.nblocks = 23558, 1
    s.prepare()
    t = s.topology().contents
    assert t.n_atoms == 23558 and t.n_tri == 7023
    assert t.n_br1 + t.n_br2 + t.n_br3 > 0 and t.n_dihes > 0 and t.n_imprs > 0
    assert tuple(s.runparams().contents.box) == (62.23, 62.23, 62.23)
