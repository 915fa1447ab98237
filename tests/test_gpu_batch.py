"""Batched systems on one GPU (run on an MI355X): every member advanced in
a batch ends bit-identical to the same member advanced alone from the same
state, step numbers and Philox stream.  A batch step runs one launch per
kernel type over all members (non-energy steps, with and without a list
rebuild) or every member's own step as a parallel graph branch (energy
steps); either way the schedule changes, never the arithmetic.  Energies
every 7 steps and rebuilds every 10 cover all four step variants."""
import copy

import numpy as np
import pytest

from blade_amd import synthetic, system
from tests.helpers import relaxed

pytestmark = pytest.mark.gpu

STEPS, FREQ_NRG = 30, 7


def _members():
    return [relaxed("msld_small"), relaxed("config1"),
            system.relax(synthetic.batch_member(11, 3300, 6))]


def test_batch_members_equal_solo_runs():
    """members with different tile chunks (4, default 16, 8) each keep their
    own in the batch; the second member runs its PME FFT on the matrix cores
    (32^3), outside the fused FP32 FFT launches"""
    specs = [copy.copy(sp) for sp in _members()]
    specs[0].tile_chunk, specs[2].tile_chunk = 4, 8
    specs[1].fft_mode, specs[1].grid = 1, (32, 32, 32)
    solo = []
    for sp in specs:
        s = system.build(sp)
        s.run_energy()                   # device system and first list
        assert s.L.bm_dynamics(s.device(), 1, STEPS, FREQ_NRG, 1) == 0
        assert s.L.bm_sync(s.device()) == 0
        solo.append(system.state(s, sp.n, sp.nblocks))
        s.close()
    systems = []
    for sp in specs:
        s = system.build(sp)
        s.run_energy()
        systems.append(s)
    b = system.Batch(systems)
    b.dynamics(1, STEPS, FREQ_NRG)
    b.sync()
    for k, (sp, s) in enumerate(zip(specs, systems)):
        x, v, th, tv = system.state(s, sp.n, sp.nblocks)
        xs, vs, ths, tvs = solo[k]
        print("\n[member %d] N=%d max|dx| %.3e" % (k, sp.n, float(np.max(np.abs(x - xs)))), flush=True)
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, xs) and np.array_equal(v, vs)
        assert np.array_equal(th, ths)
    b.close()
    for s in systems:
        s.close()


def test_batch_autotune_sets_a_measured_choice():
    """the batch tuner's cadence and tile chunk are the fastest of its
    trials, every trial ran, and every member's run parameters carry them"""
    specs = _members()
    systems = [system.build(sp) for sp in specs]
    r = system.autotune_batch(systems, 30)
    print("\n[batch autotune] %s" % r, flush=True)
    trials = r["trials"]
    assert len(trials) >= 4
    assert all(t["status"] == 0 and t["ms_per_step"] > 0 for t in trials)
    assert r["ms_per_step"] == min(t["ms_per_step"] for t in trials)
    for s in systems:
        s.run_energy()
        rp = s.runparams().contents
        assert (rp.freq_domdec, rp.tile_chunk) == (r["freq_domdec"], r["tile_chunk"])
    b = system.Batch(systems)
    b.dynamics(1, 2 * r["freq_domdec"], FREQ_NRG)
    b.sync()
    for sp, s in zip(specs, systems):
        assert np.all(np.isfinite(system.state(s, sp.n, sp.nblocks)[0]))
    b.close()
    for s in systems:
        s.close()
