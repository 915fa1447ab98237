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


def test_replica_exchange_swaps_and_metropolis_energy():
    """Neighbour replica exchange inside a batch: three replicas of one box
    that differ only in per-block fixed biases, so the exchange energy has a
    closed form, dW = (b_k - b_k+1) . (lambda(x_k+1) - lambda(x_k)) / kT.
    Replicas 0 and 1 share their biases (dW = 0: always accepted, so their
    positions, theta and velocities trade places exactly); replica 2 sits
    out the parity-0 round unchanged; the parity-1 round pairs 1 and 2 and
    leaves both in a state that is either swapped or restored, bit for bit,
    as its acceptance says."""
    base = relaxed("msld_small")
    specs = []
    for k, bias in enumerate(([0.0, 1.0, -0.5, 0.3], [0.0, 1.0, -0.5, 0.3], [0.0, -2.0, 3.0, 1.0])):
        sp = copy.copy(base)
        sp.fixed_bias = list(bias)
        sp.vel = synthetic._maxwell(np.random.default_rng(70 + k), [a[6] for a in sp.atoms], sp.temperature)
        if k == 2:                  # its own lambdas, so the parity-1 dW is not small
            sp.theta0 = [t if j == 0 else t + 0.8 * (j % 2 * 2 - 1) for j, t in enumerate(sp.theta0)]
            sp.theta_v = None
        specs.append(sp)
    systems = [system.build(sp) for sp in specs]
    for s in systems:
        s.run_energy()
    b = system.Batch(systems)
    b.dynamics(1, 10, 0)
    b.sync()
    n, nb = specs[0].n, specs[0].nblocks
    kT = synthetic.KBOLTZ * specs[0].temperature

    L = np.array(specs[0].box, float)        # positions come back as (N, 3)

    def wrap(x):
        # the list rebuild's wrap into [0, L) (x - floor(x / L) L), which the
        # exchange's energy evaluations apply
        return x - np.floor(x * (1.0 / L)) * L

    def states():
        st = []
        for s in systems:
            x, v, th, tv = system.state(s, n, nb)
            st.append((wrap(x), v, th, tv))
        return st, [s.lambdas() for s in systems]

    before, lam = states()
    rep, acc, dw = b.exchange(10, 0, 1234, [0, 1, 2])
    after, _ = states()
    print("\n[rex parity 0] accepted %s dW %s" % (acc[:2], dw[:2]), flush=True)
    assert acc[0] == 1 and abs(dw[0]) < 1e-9 and list(rep) == [1, 0, 2]
    for q in range(4):
        assert np.array_equal(after[0][q], before[1][q]) and np.array_equal(after[1][q], before[0][q])
        assert np.array_equal(after[2][q], before[2][q])
    before, lam = after, [s.lambdas() for s in systems]
    rep2, acc2, dw2 = b.exchange(10, 1, 1234, rep)
    after, _ = states()
    bias1, bias2 = np.array(specs[1].fixed_bias), np.array(specs[2].fixed_bias)
    want = float((bias1 - bias2) @ (lam[2] - lam[1])) / kT
    print("[rex parity 1] accepted %d dW %.6f closed form %.6f" % (acc2[1], dw2[1], want), flush=True)
    assert abs(dw2[1] - want) <= 1e-6 * max(1.0, abs(want))
    src = {1: 2, 2: 1} if acc2[1] else {1: 1, 2: 2}
    for k in (1, 2):
        for q in range(3):          # positions, velocities (swapped only if accepted), theta
            if q == 1:
                want_v = before[src[k]][1]
                assert np.array_equal(after[k][1], want_v)
            else:
                assert np.array_equal(after[k][q], before[src[k]][q])
    assert list(rep2) == ([1, 2, 0] if acc2[1] else [1, 0, 2])
    b.dynamics(11, 10, 0)           # the exchanged batch keeps stepping
    b.sync()
    for s in systems:
        assert np.all(np.isfinite(system.state(s, n, nb)[0]))
    b.close()
    for s in systems:
        s.close()


def _member_spec(args):
    natoms, nsubs, seed = args
    return synthetic.batch_member(seed, natoms, nsubs)


def batch32_specs():
    """the 32 batched-workload members (atom counts uniform in 3,000-25,000,
    2-6 substituents, seeded), generated on a process pool (spawned workers
    run numpy only)"""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    args = synthetic.batch_set(32)
    with cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1),
                                mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_member_spec, args))


def test_batch_32_members_equal_solo_and_oracle():
    """The 32-member batched workload (387,615 atoms): 21 steps with list
    rebuilds every 10 and energies every 7, from the generated (unrelaxed)
    states with the relaxation's first stage (0.1 fs, gamma 1000/ps).
    Every member ends bit-identical to its solo run; the smallest, the
    median and the largest member then meet the path's force, energy and
    dU/dlambda bars against the FP64 oracle at their final positions."""
    from oracle import oracle as orc
    from tests.helpers import rel_rms
    specs = batch32_specs()
    for sp in specs:
        sp.dt = 0.1e-3 * synthetic.PICOSECOND
        sp.gamma = 1000.0 / synthetic.PICOSECOND
    print("\n[batch32] %d members, %d atoms" % (len(specs), sum(sp.n for sp in specs)), flush=True)
    solo = []
    for sp in specs:
        s = system.build(sp)
        s.run_energy()
        assert s.L.bm_dynamics(s.device(), 1, 21, FREQ_NRG, 1) == 0
        assert s.L.bm_sync(s.device()) == 0
        solo.append(system.state(s, sp.n, sp.nblocks))
        s.close()
    systems = []
    for sp in specs:
        s = system.build(sp)
        s.run_energy()
        systems.append(s)
    b = system.Batch(systems)
    b.dynamics(1, 21, FREQ_NRG)
    b.sync()
    for k, (sp, s) in enumerate(zip(specs, systems)):
        x, v, th, tv = system.state(s, sp.n, sp.nblocks)
        xs, vs, ths, tvs = solo[k]
        assert np.all(np.isfinite(x))
        assert np.array_equal(x, xs) and np.array_equal(v, vs) and np.array_equal(th, ths), k
    b.close()
    order = np.argsort([sp.n for sp in specs])
    for k in (order[0], order[len(order) // 2], order[-1]):
        sp, s = specs[k], systems[k]
        s.run_energy()
        top, rp = s.topology(), s.runparams()
        x, v, th, tv = system.state(s, sp.n, sp.nblocks)
        lam = orc.theta_to_lambda(top, rp, th[:sp.nblocks])
        fo, do, eo, pairs = orc.forces(top, rp, s.positions(), lam)
        fg, eg, dg = s.forces(), s.energy(), s.dudl()
        frms = rel_rms(fg, fo)
        erel = abs(eg[20] - eo[20]) / abs(eo[20])
        ddl = float(np.max(np.abs(dg - do)))
        print("[batch32 member %d] N=%d force relRMS %.3e energy rel %.3e dudl %.2e"
              % (k, sp.n, frms, erel, ddl), flush=True)
        assert frms <= 1e-5 and erel <= 1e-6 and ddl <= 1e-3
    for s in systems:
        s.close()
