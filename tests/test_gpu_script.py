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
