"""The reference's test/water216/input deck run verbatim through
blade_interpretter on the device: 10,000 steps of 216 TIP3P waters at 300 K
with SETTLE, PME and the deck's outputs (default.nrg every 100 steps, test.lmd
every 10, test.xtc every 10,000, the checkpoint default.cpt).  Checked: the
output files, the XTC frame's header and coordinates
(read back by the reference's xdrfile reader where it is built), and the
ensemble the run samples (mean kinetic energy against equipartition with
the SETTLE constraints removed, mean potential per water of a TIP3P
liquid)."""
import os
import struct

import numpy as np
import pytest

from blade_amd import system
from tests.test_script import deck_tree, interpret

pytestmark = pytest.mark.gpu

KB = 0.001987191


def test_water216_deck_verbatim(tmp_path, monkeypatch):
    run = deck_tree(tmp_path)
    monkeypatch.chdir(run)
    s = system.BladeSystem((0,))
    print("\n[water216 deck] running ../test/water216/input", flush=True)
    err = interpret(s, tmp_path / "test" / "water216" / "input")
    assert err == ""
    nrg = np.loadtxt(run / "default.nrg")   # the deck names no .nrg file
    assert nrg.shape == (100, 24)
    assert np.array_equal(nrg[:, 0], np.arange(0, 10000, 100))
    lmd = [l.split() for l in open(run / "test.lmd")]
    assert len(lmd) == 1000 and [int(l[0]) for l in lmd[:3]] == [0, 10, 20]
    assert os.path.getsize(run / "default.cpt") > 648 * 48
    # one XTC frame (step 0): magic, atoms, step, time, box in nm
    raw = open(run / "test.xtc", "rb").read()
    magic, natoms, step = struct.unpack(">iii", raw[:12])
    box = struct.unpack(">9f", raw[16:52])
    assert (magic, natoms, step) == (1995, 648, 0)
    assert abs(box[0] - 1.8856) < 1e-6 and abs(box[4] - 1.8856) < 1e-6 and box[1] == 0
    ref = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "libxdrfile.so")
    if os.path.exists(ref):
        from tests.test_xtc import _read_all
        frames = _read_all(str(run / "test.xtc"), 648)
        assert len(frames) == 1 and np.isfinite(frames[0][3]).all()
    # the ensemble: equipartition over 3N - 3 * 216 constrained degrees of
    # freedom at 300 K (second half of the run), and a TIP3P liquid's
    # potential energy per water
    ke = nrg[50:, 1 + 21]
    dof = 3 * 648 - 3 * 216
    ke_exp = 0.5 * dof * KB * 300.0
    epot = nrg[50:, 1 + 20] / 216
    print("   <KE> %.1f kcal/mol (equipartition %.1f), <Epot>/water %.2f kcal/mol"
          % (ke.mean(), ke_exp, epot.mean()), flush=True)
    assert abs(ke.mean() / ke_exp - 1) < 0.05
    assert -11.0 < epot.mean() < -8.5
    s.close()


def test_dhfr_deck_parity_and_dynamics(tmp_path, monkeypatch):
    """The reference's DHFR benchmark deck (test/dhfr/input: 23,558 atoms,
    CHARMM22, rcut 9, PME order 4 on 64^3, gamma 5/ps, SHAKE tolerance
    1e-8) through the interpreter.  Its set-up lines run as written; the
    forces and energy of the rectified start are checked against the FP64
    oracle; then the deck's run runs with its step count cut from 500,000
    to 2,000 (10,000 atoms' worth of protein and 7,023 waters stay
    together, the energy stays finite)."""
    from oracle import oracle as orc
    from tests.helpers import rel_rms
    from tests.test_script import dhfr_tree
    run = dhfr_tree(tmp_path)
    monkeypatch.chdir(run)
    deck = (tmp_path / "test" / "dhfr" / "input").read_text()
    assert "run setvariable nsteps 500000" in deck
    short = deck.replace("run setvariable nsteps 500000", "run setvariable nsteps 2000")
    k = short.index("run dynamics")
    (run / "setup.inp").write_text(short[:k] + "run energy\n")
    s = system.BladeSystem((0,))
    print("\n[dhfr deck] set-up and run energy", flush=True)
    assert interpret(s, run / "setup.inp") == ""
    s.n, s.nblocks = 23558, 1
    top, rp = s.topology(), s.runparams()
    x = s.positions()
    fo, do, eo, pairs = orc.forces(top, rp, x, np.ones(1))
    fg, eg = s.forces(), s.energy()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    print("   pairs %d force relRMS %.3e Epot %.4f (oracle %.4f) rel %.3e" % (pairs, frms, eg[20], eo[20], erel),
          flush=True)
    assert frms <= 1e-5 and erel <= 1e-6
    (run / "dyn.inp").write_text(short[k:])
    print("[dhfr deck] 2,000 steps", flush=True)
    assert interpret(s, run / "dyn.inp") == ""
    x1 = s.positions().reshape(3, 23558).T
    assert np.isfinite(x1).all()
    # the energy of the final state (the deck's .nrg cadence, 50,000 steps,
    # wrote only step 0)
    (run / "final.inp").write_text("run energy\n")
    assert interpret(s, run / "final.inp") == ""
    e = s.energy()
    assert np.isfinite(e).all()
    ke_exp = 0.5 * (3 * 23558 - 3 * 7023 - (458 + 2 * 233 + 3 * 99)) * KB * 298.0
    print("   after 2,000 steps: Epot %.1f (start %.1f) KE %.1f (equipartition %.1f)"
          % (e[20], eg[20], e[21], ke_exp), flush=True)
    assert abs(e[20] - eg[20]) < 0.05 * abs(eg[20])
    if e[21] != 0:
        assert abs(e[21] / ke_exp - 1) < 0.1
    s.close()
