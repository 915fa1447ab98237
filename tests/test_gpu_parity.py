"""GPU parity against the FP64 oracle (run on an MI355X).

Bars (stated per assertion):
  forces          <= 1e-5 relative RMS (all terms together)
  total potential <= 1e-6 relative
  dU/dlambda      <= 1e-3 kcal/mol per block
  lists           bit-exact (slot order, tiles, exclusion masks)
  trajectories    bit-identical run to run, graph replay == eager launches
"""
import numpy as np
import pytest

from blade_amd import system
from oracle import oracle as orc
from tests.helpers import built, lambdas0, rel_rms, relaxed, spec

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6
DUDL_TOL = 1e-3


def _gpu_energy(name):
    s = built(name)
    s.run_energy()
    return s


def _compare(name, report=True):
    s = _gpu_energy(name)
    top, rp = s.topology(), s.runparams()
    pos = s.positions()          # wrapped by the rebuild, as the oracle sees them
    lam = lambdas0(s, name)
    fo, do, eo, pairs = orc.forces(top, rp, pos, lam)
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    if report:
        names = ["bond", "angle", "urey", "dihe", "impr", "cmap", "nb14", "nbdirect", "nbrecip",
                 "self", "excl", "lambda"]
        print("\n[%s] N=%d pairs=%d force relRMS=%.3e  Epot gpu=%.6f oracle=%.6f rel=%.3e"
              % (name, s.n, pairs, frms, eg[20], eo[20], erel), flush=True)
        for i, nm in enumerate(names):
            if eo[i] or eg[i]:
                print("   %-8s gpu %16.6f oracle %16.6f diff %.3e" % (nm, eg[i], eo[i], eg[i] - eo[i]), flush=True)
        if len(do) > 1:
            print("   dudl gpu", dg, "oracle", do, flush=True)
    return s, fo, fg, eo, eg, do, dg, frms, erel


def test_config1_forces_energy():
    s, fo, fg, eo, eg, do, dg, frms, erel = _compare("config1")
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL


def test_msld_small_forces_energy_dudl():
    s, fo, fg, eo, eg, do, dg, frms, erel = _compare("msld_small")
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL
    assert np.max(np.abs(dg - do)) <= DUDL_TOL


def test_config2_small_forces_energy_dudl():
    s, fo, fg, eo, eg, do, dg, frms, erel = _compare("config2_small")
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL
    assert np.max(np.abs(dg - do)) <= DUDL_TOL


def test_config2_forces_energy_dudl():
    """the full 23,558-atom benchmark system"""
    s, fo, fg, eo, eg, do, dg, frms, erel = _compare("config2")
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL
    assert np.max(np.abs(dg - do)) <= DUDL_TOL


def test_two_site_small_forces_energy_dudl():
    """two lambda sites: blocks of different sites scale by the product of
    their lambdas, blocks of one site do not interact"""
    s, fo, fg, eo, eg, do, dg, frms, erel = _compare("two_site_small")
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL
    assert np.max(np.abs(dg - do)) <= DUDL_TOL


def test_config3_precision_form():
    """92,224 atoms, two lambda sites x 4 substituents, PME 96^3 on the
    split-BF16x3 MFMA FFT, from the GPU-relaxed state.  Forces and
    dU/dlambda keep the path's bars; the reciprocal-space energy is held to
    1e-4 relative and the total-potential error is printed."""
    sp = relaxed("config3")
    assert sp.fft_mode == 1
    s = system.build(sp)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    fo, do, eo, pairs = orc.forces(top, rp, s.positions(), lam)
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    frms = rel_rms(fg, fo)
    rrel = abs(eg[8] - eo[8]) / abs(eo[8])
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    ddl = float(np.max(np.abs(dg - do)))
    print("\n[config3] N=%d pairs=%d force relRMS %.3e recip %.6f vs %.6f rel %.3e potential %.4f vs %.4f rel %.3e "
          "max |dU/dl| err %.2e" % (sp.n, pairs, frms, eg[8], eo[8], rrel, eg[20], eo[20], erel, ddl), flush=True)
    assert frms <= FORCE_TOL
    assert rrel <= 1e-4
    assert ddl <= DUDL_TOL


def test_config5_four_site_replica():
    """one replica of the combined workload (60,000 atoms, four lambda sites
    of 6, 6, 6 and 8 substituents: 27 blocks, soft core, atom restraints,
    per-block fixed biases, PME 80^3 on the MFMA FFT) from the GPU-relaxed
    state: forces and dU/dlambda keep the path's bars"""
    sp = relaxed("config5")
    assert sp.nblocks == 27 and sp.fft_mode == 1
    s = system.build(sp)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    fo, do, eo, pairs = orc.forces(top, rp, s.positions(), lam)
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    frms = rel_rms(fg, fo)
    ddl = float(np.max(np.abs(dg - do)))
    print("\n[config5] N=%d blocks=%d force relRMS %.3e potential rel %.3e max |dU/dl| err %.2e"
          % (sp.n, sp.nblocks, frms, abs(eg[20] - eo[20]) / abs(eo[20]), ddl), flush=True)
    assert frms <= FORCE_TOL
    assert ddl <= DUDL_TOL


@pytest.mark.parametrize("name", ["config1", "msld_small", "config2_small", "two_site_small", "config2",
                                  "config5"])
def test_domdec_bit_exact(name):
    """wrapped positions, slot order, block counts, directed tiles and
    exclusion masks bit-exact against the oracle's list semantics, up to the
    23,558-atom box and a 60,000-atom four-site replica"""
    s = built(name)
    top, rp = s.topology(), s.runparams()
    # the device system starts from the rectified coordinates
    x0 = orc.rectify(top, rp, system.soa(spec(name).pos))
    posw, l2g, bc, tiles, masks = orc.domdec(top, rp, x0)
    s.run_energy()
    dev = s.device()
    import ctypes as C
    L = s.L
    nb, nt, ne = C.c_int(), C.c_int(), C.c_int()
    assert L.bm_get_domdec_sizes(dev, C.byref(nb), C.byref(nt), C.byref(ne)) == 0
    gl2g = np.zeros(32 * nb.value, dtype=np.int32)
    gbc = np.zeros(nb.value, dtype=np.int32)
    gti = np.zeros(nt.value, dtype=np.int32)
    gtj = np.zeros(nt.value, dtype=np.int32)
    gtm = np.zeros(32 * nt.value, dtype=np.uint32)
    i32 = C.POINTER(C.c_int)
    assert L.bm_get_domdec(dev, gl2g.ctypes.data_as(i32), gbc.ctypes.data_as(i32), gti.ctypes.data_as(i32),
                           gtj.ctypes.data_as(i32), gtm.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    print("\n[%s] blocks %d/%d tiles %d/%d" % (name, nb.value, len(bc), nt.value, len(tiles)), flush=True)
    assert nb.value == len(bc)
    assert np.array_equal(gl2g, l2g)
    assert np.array_equal(gbc, bc)
    assert nt.value == len(tiles)
    assert np.array_equal(gti, tiles[:, 0]) and np.array_equal(gtj, tiles[:, 1])
    assert np.array_equal(gtm.reshape(-1, 32), masks)
    # the wrapped positions are identical too
    assert np.array_equal(s.positions(), posw)


def _run(sp, nsteps, use_graph):
    s = system.build(sp)
    s.run_dynamics(nsteps, freq_nrg=10, use_graph=use_graph)
    return s.positions(), s.energy(), s.lambdas()


def test_trajectory_bit_reproducible_and_graph_equals_eager():
    sp = relaxed("msld_small")
    xa, ea, la = _run(sp, 40, True)
    xb, eb, lb = _run(sp, 40, True)
    xc, ec, lc = _run(sp, 40, False)
    print("\nlambda after 40 steps", la, flush=True)
    assert np.all(np.isfinite(xa))
    assert np.array_equal(xa, xb)
    assert np.array_equal(xa, xc)
    assert np.array_equal(ea, eb)
    assert np.array_equal(la, lc)


def test_dynamics_constraints_and_stability():
    name = "config1"
    s = built(name)
    s.run_dynamics(50, freq_nrg=10)
    x = s.positions().reshape(3, -1).T
    t = s.topology().contents
    tri = t.array("tri_idx", t.n_tri, 3)
    b0 = t.array("tri_b0", t.n_tri, 3)
    L = np.array(spec(name).box)
    worst = 0.0
    for (a, b, c), (d01, d02, d12) in zip(tri, b0):
        for (i, j, d) in ((a, b, d01), (a, c, d02), (b, c, d12)):
            r = x[i] - x[j]
            r -= L * np.round(r / L)
            worst = max(worst, abs(np.linalg.norm(r) - d) / d)
    print("\nworst relative constraint deviation after 50 steps: %.2e" % worst, flush=True)
    assert worst < 1e-6
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("name", ["msld_small", "two_site_small"])
def test_short_trajectory_tracks_oracle(name):
    """5 steps GPU vs FP64 oracle from the same state and Philox stream"""
    sp = relaxed(name)
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n, nl = sp.n, sp.nblocks
    x = np.concatenate([orc.rectify(top, rp, system.soa(sp.pos)), np.array(sp.theta0, float)])
    v = np.concatenate([system.soa(sp.vel), np.array(sp.theta_v, float)])
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    # the GPU wraps positions at step 0 (rebuild); do the same for the oracle
    posw = orc.domdec(top, rp, x[:3 * n])[0]
    x[:3 * n] = posw
    for k in range(5):
        orc.step(top, rp, k, k % 10 == 0, x, v, lam)
    s.run_dynamics(5, freq_nrg=10)
    xg = s.positions()
    d = np.abs(xg - x[:3 * n])
    print("\nmax |x_gpu - x_oracle| after 5 steps: %.3e A" % d.max(), flush=True)
    assert d.max() < 1e-4


def test_checkpoint_restart_is_bit_identical_and_outputs(tmp_path):
    """40 steps in one run against 20 steps, a binary checkpoint (written on
    a list rebuild step), a fresh system reading it and 20 more steps: the
    states are bit-identical.  The run's .nrg and .lmd files hold the
    reference's line layout at their steps: the potential column is the sum
    of the term columns and the total is potential + kinetic; the continued
    run's XDR .lmd frames hold the same lambdas in single precision."""
    sp = relaxed("msld_small")
    a = system.build(sp)
    a.set_output(nrg=tmp_path / "a.nrg", lmd=tmp_path / "a.lmd", freq_nrg=10, freq_lmd=5)
    a.run_dynamics(40, freq_nrg=10)
    b = system.build(sp)
    b.run_dynamics(20, freq_nrg=10)
    b.write_checkpoint(tmp_path / "ck.bin")
    b.close()
    c = system.build(sp)
    c.read_checkpoint(tmp_path / "ck.bin")
    c.set_output(lmd=tmp_path / "c.lmd", freq_lmd=5, lmd_text=False)
    c.run_dynamics(20, freq_nrg=10)
    c.set_output()
    xa, va, _, _ = system.state(a, sp.n, sp.nblocks)
    xc, vc, _, _ = system.state(c, sp.n, sp.nblocks)
    assert np.array_equal(xa, xc) and np.array_equal(va, vc)
    assert np.array_equal(a.lambdas(), c.lambdas())
    a.set_output()                      # closes (and flushes) the files
    nrg = [l.split() for l in open(tmp_path / "a.nrg")]
    lmd = [l.split() for l in open(tmp_path / "a.lmd")]
    assert [int(l[0]) for l in nrg] == [0, 10, 20, 30]
    assert all(len(l) == 1 + 23 for l in nrg)
    for l in nrg:
        e = np.array([float(v) for v in l[1:]])
        assert np.all(np.isfinite(e)) and e[20] != 0.0
        assert abs(e[:20].sum() - e[20]) < 2e-3
        assert abs(e[20] + e[21] - e[22]) < 2e-4
    assert [int(l[0]) for l in lmd] == list(range(0, 40, 5))
    assert all(len(l) == sp.nblocks for l in lmd)
    lam = [float(v) for v in lmd[-1][1:]]
    assert abs(sum(lam) - 1.0) < 1e-5
    # the continued run's XDR frames (steps 20..35) against the text lines
    xdr = system.read_lmd_xdr(tmp_path / "c.lmd")
    txt = np.array([[float(v) for v in l[1:]] for l in lmd[4:]])
    assert xdr.shape == txt.shape == (4, sp.nblocks - 1)
    assert np.max(np.abs(xdr - txt)) <= 5.1e-7
