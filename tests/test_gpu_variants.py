"""GPU parity of the kernel variants the benchmark system does not use
(run on an MI355X), against the FP64 oracle with the same bars:
forces <= 1e-5 relative RMS, total energy <= 1e-6 relative, dU/dlambda
<= 1e-3 kcal/mol.

  vdW potential switch and shift; electrostatics force switch and force
  shift (no PME); soft-core lambda sites; B-spline orders 4 and 8; grids
  that take the generic FFT path (non-square planes, planes too large for
  LDS, sizes outside the compiled set); the per-term API calls summing to
  the full force evaluation bit for bit.
"""
import copy
import ctypes as C

import numpy as np
import pytest

from blade_amd import system
from oracle import oracle as orc
from tests.helpers import rel_rms, spec

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
ENERGY_TOL = 1e-6
DUDL_TOL = 1e-3


def _variant(base, **kw):
    sp = copy.copy(spec(base))
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


def _check(sp, label):
    s = system.build(sp)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    pos = s.positions()
    t = top.contents
    lam = (orc.theta_to_lambda(top, rp, np.array(sp.theta0, float)) if t.n_blocks > 1 else np.ones(1))
    fo, do, eo, pairs = orc.forces(top, rp, pos, lam)
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    ddl = float(np.max(np.abs(dg - do))) if len(do) > 1 else 0.0
    print("\n[%s] N=%d pairs=%d force relRMS %.3e energy rel %.3e dudl %.2e"
          % (label, sp.n, pairs, frms, erel, ddl), flush=True)
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL
    assert ddl <= DUDL_TOL
    return s


@pytest.mark.parametrize("vdw,elec", [(1, 1), (2, 1), (0, 0), (0, 2), (1, 0), (2, 2)])
def test_switching_methods(vdw, elec):
    _check(_variant("msld_small", vdw_method=vdw, elec_method=elec), "vdw%d elec%d" % (vdw, elec))


def test_atom_restraints_displaced():
    """MSLD atom restraints with the restrained atoms moved apart by up to
    0.5 A per axis, so every group pulls (energy term BM_ECATS)"""
    sp = _variant("two_site_small")
    pos = np.array(sp.pos, float)
    rng = np.random.default_rng(9)
    for g in sp.atom_restraints:
        for a in g:
            pos[a] += rng.uniform(-0.5, 0.5, 3)
    sp.pos = pos
    _check(sp, "atom restraints")


@pytest.mark.parametrize("k,q0,w", [((1.0, 1.0, 3.0), 0.1, 0.05), ((0.0, 0.0, 5.0), -0.3, 0.0)])
def test_charge_restraints(k, q0, w):
    """MSLD charge restraints on the lambda-weighted net charge of the
    blocks (energy term BM_EBIAS, FP64 on the device): the term itself to
    one quantum of the 2^-32 kcal/mol fixed-point energy sum plus 1e-9
    relative, and dU/dlambda to the path's bar"""
    sp = _variant("two_site_small", block_charge=[0.4, 1.0, -1.0, 0.5, -0.5, 0.25],
                  charge_restraints=(k[0], k[1], k[2], q0, w))
    s = _check(sp, "charge restraints %s q0 %g w %g" % (k, q0, w))
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    _, do, eo, _ = orc.forces(top, rp, s.positions(), lam, orc.MSLDBIAS)
    eg = s.energy()
    assert eo[19] != 0.0
    assert abs(eg[19] - eo[19]) <= 2.0 ** -32 + 1e-9 * abs(eo[19]), (eg[19], eo[19])


def test_cmap_terms():
    """CMAP cross terms (24 terms, several lambda-scaled) against the oracle
    at the path's force, energy and dU/dlambda bars"""
    from blade_amd import synthetic
    sp = synthetic.add_cmaps(copy.deepcopy(spec("two_site_small")), count=24)
    _check(sp, "cmap")


def test_soft_core_site():
    _check(_variant("msld_small", soft_core=True), "soft core")


@pytest.mark.parametrize("order,grid", [(4, (24, 20, 18)), (8, (20, 20, 20)), (6, (72, 72, 72)), (6, (30, 30, 30))])
def test_pme_orders_and_fft_paths(order, grid):
    _check(_variant("config1", order=order, grid=grid), "order %d grid %s" % (order, grid))


@pytest.mark.parametrize("base,grid", [("config1", (32, 32, 32)), ("msld_small", (32, 32, 32)),
                                       ("config2_small", (48, 48, 48))])
def test_mfma_split_bf16_fft(base, grid):
    """The PME FFT as DFT contractions on the matrix cores in split-BF16x3
    (FP32 accumulation).  Forces keep the path's own bar (1e-5 relative RMS,
    dU/dlambda 1e-3 kcal/mol).  The split products carry about 2^-17
    relative error per term through five passes, so the reciprocal-space
    energy is held to 1e-4 relative (the FP32 Stockham path is timed beside
    it on the same system and its error printed for comparison)."""
    errs = {}
    for mode in (1, 0):
        sp = _variant(base, grid=grid, fft_mode=mode)
        s = system.build(sp)
        s.run_energy()
        top, rp = s.topology(), s.runparams()
        t = top.contents
        lam = (orc.theta_to_lambda(top, rp, np.array(sp.theta0, float)) if t.n_blocks > 1 else np.ones(1))
        fo, do, eo, pairs = orc.forces(top, rp, s.positions(), lam)
        fg, eg, dg = s.forces(), s.energy(), s.dudl()
        errs[mode] = (rel_rms(fg, fo), abs(eg[8] - eo[8]) / abs(eo[8]), abs(eg[8] - eo[8]),
                      abs(eg[20] - eo[20]) / abs(eo[20]),
                      float(np.max(np.abs(dg - do))) if len(do) > 1 else 0.0)
        print("\n[%s fft %s %s] force relRMS %.3e recip rel %.3e (%.3e kcal/mol) potential rel %.3e dudl %.2e"
              % (("mfma" if mode else "fp32", base, grid) + errs[mode]), flush=True)
    frms, rrel, _, _, ddl = errs[1]
    assert frms <= FORCE_TOL
    assert rrel <= 1e-4
    assert ddl <= DUDL_TOL


@pytest.mark.parametrize("chunk", [1, 4, 8])
def test_tile_chunk_shapes(chunk):
    """directed tiles per nonbonded work unit below the default 16: the FP32
    partial sums group differently, the bars are the same"""
    _check(_variant("msld_small", tile_chunk=chunk), "tile chunk %d" % chunk)


def test_autotune_picks_the_fastest_measured_candidate():
    """the tuner's choice is the fastest of its trials, every trial ran, the
    tuned system keeps the parity bars and its run parameters carry the
    choice"""
    sp = _variant("msld_small", grid=(32, 32, 32))
    s = system.build(sp)
    r = s.autotune(40)
    print("\n[autotune] %s" % r, flush=True)
    trials = r["trials"]
    assert len(trials) >= 7                      # 2 FFT forms, 4 cadences, 3 chunks, shared base
    assert all(t["status"] == 0 and t["ms_per_step"] > 0 for t in trials)
    assert {t["fft_mode"] for t in trials} == {0, 1}
    assert r["ms_per_step"] == min(t["ms_per_step"] for t in trials)
    s.run_energy()
    rp = s.runparams().contents
    assert (rp.fft_mode, rp.freq_domdec, rp.tile_chunk) == (r["fft_mode"], r["freq_domdec"], r["tile_chunk"])
    top = s.topology()
    lam = orc.theta_to_lambda(top, s.runparams(), np.array(sp.theta0, float))
    fo, do, eo, pairs = orc.forces(top, s.runparams(), s.positions(), lam)
    assert rel_rms(s.forces(), fo) <= FORCE_TOL
    assert float(np.max(np.abs(s.dudl() - do))) <= DUDL_TOL


def test_per_term_calls_sum_to_full_force():
    """bm_reset_force, then every per-term entry point once (bonded, PME,
    nonbonded, Ewald self, 1-4, exclusion corrections, MSLD biases,
    restraints): the force array, dU/dlambda and every energy term equal one
    bm_calc_force, bit for bit"""
    from blade_amd import synthetic
    sp = synthetic.add_restraints(_variant("msld_small", var_bias=VB_ALL, theta_inde=[0.0, 0.3]))
    s = system.build(sp)
    s.run_energy()
    L, dev = s.L, s.device()
    n3 = 3 * sp.n
    i64 = C.POINTER(C.c_longlong)

    def snapshot():
        f = np.zeros(n3 + sp.nblocks, dtype=np.int64)
        assert L.bm_get_force_fixed(dev, f.ctypes.data_as(i64)) == 0
        return f[:n3], s.dudl(), s.energy()
    full = snapshot()
    assert L.bm_reset_force(dev, 1) == 0
    for fn in (L.bm_getforce_bonded, L.bm_getforce_ewald, L.bm_getforce_nbdirect, L.bm_getforce_ewaldself,
               L.bm_getforce_nb14, L.bm_getforce_nbex, L.bm_getforce_msldbias, L.bm_getforce_restrain):
        assert fn(dev, 1) == 0
    assert L.bm_sync(dev) == 0
    parts = snapshot()
    assert np.array_equal(parts[0], full[0])
    assert np.array_equal(parts[1], full[1])
    assert np.array_equal(parts[2][:20], full[2][:20])
    assert all(full[2][k] != 0.0 for k in (12, 14, 15, 16, 17))


def test_restraints():
    """NOE, harmonic, bond/angle/dihedral restraint and RESD terms against
    the oracle, each energy term to 1e-8 relative"""
    from blade_amd import synthetic
    sp = synthetic.add_restraints(spec("msld_small"))
    s = _check(sp, "restraints")
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    eo = orc.forces(top, rp, s.positions(), lam, orc.RESTRAIN)[2]
    eg = s.energy()
    for k in (14, 15, 16, 17):
        assert eo[k] != 0.0 and abs(eg[k] - eo[k]) <= 1e-8 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])


def test_term_by_term_step_equals_dynamics():
    """the reference's per-step seam driven through the ABI -
    bm_domdec_update (rebuild every 10), bm_calc_force (energies every 7),
    bm_update - for 25 steps ends bit-identical to bm_dynamics over the
    same steps, eager and graph-replayed"""
    from tests.helpers import relaxed
    sp = relaxed("msld_small")
    out = []
    for mode in ("terms", "eager", "graph"):
        s = system.build(sp)
        s.run_energy()
        L, dev = s.L, s.device()
        if mode == "terms":
            for k in range(1, 26):
                assert L.bm_domdec_update(dev, int(k % 10 == 0)) == 0
                assert L.bm_calc_force(dev, k, int(k % 7 == 0)) == 0
                assert L.bm_update(dev, k, int(k % 7 == 0)) == 0
        else:
            assert L.bm_dynamics(dev, 1, 25, 7, int(mode == "graph")) == 0
        assert L.bm_sync(dev) == 0
        out.append(system.state(s, sp.n, sp.nblocks))
        s.close()
    for q in range(4):
        assert np.array_equal(out[0][q], out[1][q]), q
        assert np.array_equal(out[0][q], out[2][q]), q


def test_holonomic_and_chain_rule_entry_points():
    """bm_holonomic_velocity projects velocities onto the constraints
    (v_ij . r_ij = 0); bm_holonomic_position restores perturbed positions to
    the constraint lengths from the backup; bm_msld_chain_rule writes the
    oracle's dU/dtheta"""
    from tests.helpers import relaxed
    sp = relaxed("msld_small")
    s = system.build(sp)
    s.run_dynamics(10, freq_nrg=0)
    L, dev = s.L, s.device()
    n = sp.n
    t = s.topology().contents
    tri = t.array("tri_idx", t.n_tri, 3)
    f64 = C.POINTER(C.c_double)
    x, v, th, tv = system.state(s, n, sp.nblocks)
    v2 = v + np.random.default_rng(1).normal(scale=0.01, size=v.shape)
    vs = system.soa(v2)
    # set the perturbed velocities through a checkpoint-free path: positions
    # stay, velocities are read back after the projection
    assert L.bm_set_velocities(dev, vs.ctypes.data_as(f64), None) == 0
    assert L.bm_holonomic_velocity(dev) == 0
    assert L.bm_sync(dev) == 0
    x, v, th, tv = system.state(s, n, sp.nblocks)
    Lb = np.array(sp.box)
    worst = 0.0
    for a, b, c in tri:
        for i, j in ((a, b), (a, c), (b, c)):
            r = x[i] - x[j]
            r -= Lb * np.round(r / Lb)
            worst = max(worst, abs(np.dot(v[i] - v[j], r)))
    assert worst < 1e-10
    xp = x + np.random.default_rng(2).normal(scale=0.005, size=x.shape)
    assert L.bm_set_positions(dev, system.soa(xp).ctypes.data_as(f64), None) == 0
    assert L.bm_holonomic_position(dev) == 0
    assert L.bm_sync(dev) == 0
    x = system.state(s, n, sp.nblocks)[0]
    b0 = t.array("tri_b0", t.n_tri, 3)
    dev_len = 0.0
    for (a, b, c), (d01, d02, d12) in zip(tri, b0):
        for (i, j, d) in ((a, b, d01), (a, c, d02), (b, c, d12)):
            r = x[i] - x[j]
            r -= Lb * np.round(r / Lb)
            dev_len = max(dev_len, abs(np.linalg.norm(r) - d) / d)
    assert dev_len < 1e-6
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    x, v, th, tv = system.state(s, n, sp.nblocks)
    lam = orc.theta_to_lambda(top, rp, th[:sp.nblocks])
    _, do, _, _ = orc.forces(top, rp, s.positions(), lam)
    want = orc.chain_rule(top, rp, th[:sp.nblocks], lam, do)
    assert L.bm_msld_chain_rule(dev) == 0
    got = np.zeros(sp.nblocks)
    assert L.bm_get_forces(dev, np.zeros(3 * n).ctypes.data_as(f64), got.ctypes.data_as(f64)) == 0
    assert np.max(np.abs(got - want)) < 1e-3


@pytest.mark.parametrize("rcut,rswitch", [(9.0, 8.0), (10.0, 8.5)])
def test_reference_water216_deck(rcut, rswitch):
    """the reference's test/water216 PSF/CRD/PRM (fixtures), read from
    files; at 10 A (the deck's default cutoff) the 18.856 A box edge is
    below twice the cutoff and pairs interact through their nearest image"""
    from tests.test_files import BOX, CRD, PRM, PSF
    s = system.load_charmm(PRM, PSF, CRD, BOX, beta=1 / 3.125, rcut=rcut, rswitch=rswitch)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    fo, do, eo, pairs = orc.forces(top, rp, s.positions(), np.ones(1))
    fg, eg = s.forces(), s.energy()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    print("\n[water216 deck, rcut %.0f] pairs=%d force relRMS %.3e energy %.6f (oracle %.6f) rel %.3e"
          % (rcut, pairs, frms, eg[20], eo[20], erel), flush=True)
    assert frms <= FORCE_TOL
    assert erel <= ENERGY_TOL


def _lengths_worst(top, P, box):
    t = top.contents
    worst = 0.0
    groups = [(t.array("tri_idx", t.n_tri, 3), t.array("tri_b0", t.n_tri, 3), ((0, 1), (0, 2), (1, 2))),
              (t.array("br1_idx", t.n_br1, 2), t.array("br1_b0", t.n_br1, 1), ((0, 1),)),
              (t.array("br2_idx", t.n_br2, 3), t.array("br2_b0", t.n_br2, 2), ((0, 1), (0, 2))),
              (t.array("br3_idx", t.n_br3, 4), t.array("br3_b0", t.n_br3, 3), ((0, 1), (0, 2), (0, 3)))]
    for idx, b0, pairs in groups:
        b0 = np.asarray(b0).reshape(len(idx), len(pairs))
        for a, b in zip(idx, b0):
            for (i, j), d in zip(pairs, b):
                r = P[a[j]] - P[a[i]]
                r -= box * np.round(r / box)
                worst = max(worst, abs(np.linalg.norm(r) - d) / d)
    return worst


def test_initial_rectify_matches_oracle_then_dynamics():
    """Coordinates whose constraint groups are off their lengths are
    rectified when the device system is created (run/run.cu:606), bit for
    bit as the oracle rectifies them: the reference's water216 deck with every
    hydrogen displaced by 0.05 A, then 100 steps of dynamics that keep the
    lengths; and the small MSLD system, whose SHAKE branches of one to
    three hydrogens take the branch rule."""
    from tests.test_files import BOX, CRD, PRM, PSF
    rng = np.random.default_rng(21)
    s = system.load_charmm(PRM, PSF, CRD, BOX, beta=1 / 3.125, rcut=9.0, rswitch=8.0)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n = s.n
    P = s.positions().reshape(3, n).T.copy()
    for o, h1, h2 in top.contents.array("tri_idx", top.contents.n_tri, 3):
        P[h1] += rng.normal(scale=0.05, size=3)
        P[h2] += rng.normal(scale=0.05, size=3)
    for i in range(n):
        s.add_coordinates_position(i + 1, *P[i])
    box = np.array(BOX)
    assert _lengths_worst(top, P, box) > 1e-2
    s.run_energy()
    xg = s.positions()
    want = orc.domdec(top, rp, orc.rectify(top, rp, np.ascontiguousarray(P.T).reshape(-1)))[0]
    diff = np.abs(xg - want).max()
    print("\n[water216 deck, displaced] rectified: max |gpu - oracle| %.3e A, %d of %d coordinates bit-equal"
          % (diff, int(np.sum(xg == want)), xg.size), flush=True)
    assert np.array_equal(xg, want)
    assert _lengths_worst(top, xg.reshape(3, n).T, box) < 1e-13
    fo, do, eo, pairs = orc.forces(top, rp, xg, np.ones(1))
    assert rel_rms(s.forces(), fo) <= FORCE_TOL
    s.run_dynamics(100, freq_nrg=10)
    x1 = s.positions()
    assert np.all(np.isfinite(x1))
    w = _lengths_worst(top, x1.reshape(3, n).T, box)
    print("   after 100 steps: worst relative constraint error %.3e" % w, flush=True)
    assert w < 1e-9
    s.close()

    sp = copy.deepcopy(spec("msld_small"))
    pos = np.array(sp.pos, float)
    pos += rng.normal(scale=0.02, size=pos.shape)
    sp.pos = pos
    m = system.build(sp)
    m.prepare()
    top, rp = m.topology(), m.runparams()
    t = top.contents
    assert t.n_br1 and t.n_br2 and t.n_br3 and t.n_tri
    m.run_energy()
    xg = m.positions()
    want = orc.domdec(top, rp, orc.rectify(top, rp, system.soa(sp.pos)))[0]
    diff = np.abs(xg - want).max()
    print("[msld_small, displaced] rectified: max |gpu - oracle| %.3e A" % diff, flush=True)
    assert np.array_equal(xg, want)
    assert _lengths_worst(top, xg.reshape(3, sp.n).T, np.array(sp.box)) < 1e-13
    m.close()


def test_barostat_volume_move_restore_and_npt_run():
    """Monte Carlo barostat: a volume move scales free atoms exactly, keeps
    every constraint group's internal offsets, and the forces and energy at
    the new box match the oracle's at that box; restoring gives back the
    positions bit for bit; an NPT run (a move every 10 steps) is
    bit-reproducible and keeps its counters."""
    from tests.helpers import relaxed
    sp = relaxed("msld_small")
    s = system.build(sp)
    s.run_energy()
    L, dev = s.L, s.device()
    n = sp.n
    box0 = np.zeros(3)
    assert L.bm_get_box(dev, box0.ctypes.data_as(C.POINTER(C.c_double))) == 0
    x0 = s.positions().copy()
    scale = np.array([1.01, 0.99, 1.005])
    assert L.bm_volume_move(dev, scale.ctypes.data_as(C.POINTER(C.c_double))) == 0
    x1 = s.positions().copy()
    box1 = np.zeros(3)
    L.bm_get_box(dev, box1.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(box1, box0 * scale)
    top = s.topology()
    t = top.contents
    grouped = set()
    for key, cnt, w in (("tri_idx", t.n_tri, 3), ("br1_idx", t.n_br1, 2), ("br2_idx", t.n_br2, 3),
                        ("br3_idx", t.n_br3, 4)):
        for g in t.array(key, cnt, w):
            grouped |= set(int(a) for a in g)
            for j in g[1:]:
                for k in range(3):
                    d0 = x0[k * n + j] - x0[k * n + g[0]]
                    d0 -= box0[k] * np.floor(d0 / box0[k] + 0.5)
                    d1 = x1[k * n + j] - x1[k * n + g[0]]
                    assert abs(d1 - d0) < 1e-12
    free = np.array([a for a in range(n) if a not in grouped], dtype=np.int64)
    # every first atom of a group and every free atom is scaled exactly
    heads = np.array(sorted({int(g[0]) for key, cnt, w in (("tri_idx", t.n_tri, 3), ("br1_idx", t.n_br1, 2),
                                                          ("br2_idx", t.n_br2, 3), ("br3_idx", t.n_br3, 4))
                             for g in t.array(key, cnt, w)}), dtype=np.int64)
    for idx in (free, heads):
        for k in range(3):
            assert np.array_equal(x1[k * n + idx], x0[k * n + idx] * scale[k])
    # forces and energy at the new box against the oracle
    assert L.bm_domdec_update(dev, 1) == 0
    assert L.bm_calc_force(dev, 0, 1) == 0
    assert L.bm_sync(dev) == 0
    rp = copy.copy(s.runparams().contents)
    for k in range(3):
        rp.box[k] = box1[k]
    lam = orc.theta_to_lambda(top, C.pointer(rp), np.array(sp.theta0, float))
    fo, do, eo, _ = orc.forces(top, C.pointer(rp), s.positions(), lam)
    fg, eg = s.forces(), s.energy()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    print("\n[volume move] force relRMS %.3e energy rel %.3e" % (frms, erel), flush=True)
    assert frms <= FORCE_TOL and erel <= ENERGY_TOL
    assert L.bm_volume_restore(dev, box0.ctypes.data_as(C.POINTER(C.c_double))) == 0
    x2 = s.positions()
    # restored positions equal the backup taken before the move (the list
    # rebuild above does not touch the backup)
    assert np.array_equal(x2, x0)
    s.close()
    # NPT runs: a move every 10 steps, twice from the same state
    runs = []
    for _ in range(2):
        sp2 = copy.copy(sp)
        sp2.freq_npt, sp2.volume_fluct = 10, 30.0
        s2 = system.build(sp2)
        s2.run_dynamics(100, freq_nrg=0)
        moves, acc, box = s2.npt_stats()
        runs.append((s2.positions().copy(), box.copy(), moves, acc))
        s2.close()
    print("[npt] moves %d accepted %d box %s" % (runs[0][2], runs[0][3], runs[0][1]), flush=True)
    assert runs[0][2] == 10 and 0 <= runs[0][3] <= 10
    assert np.all(np.isfinite(runs[0][0]))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][3] == runs[1][3]


def test_term_variety():
    """Urey-Bradley terms, periodic impropers, NBFIX pairs and the
    geometric combination rule"""
    from blade_amd import synthetic
    s = _check(synthetic.term_variety(spec("msld_small")), "UB, periodic impropers, NBFIX, geometric")
    e = s.energy()
    assert e[2] != 0.0 and e[4] != 0.0


def test_soft_core_14_and_soft_core():
    _check(_variant("msld_small", soft_core=True, soft_core14=True), "soft core + soft-core 1-4")


@pytest.mark.parametrize("mode", [1, 3])
def test_msld_ewald_types(mode):
    """msldEwaldType 1 (same-block corrections scaled once) and 3 (every
    correction scaled by both lambdas) on the two-site system"""
    _check(_variant("two_site_small", ewald_type=mode), "msldEwaldType %d" % mode)


VB_ALL = [(1, 2, 1, 0.8, 1.5, 2), (2, 3, 2, 0.1, 1.5, 3), (1, 3, 3, 0.0, 0.7, 2), (2, 1, 4, 0.6, 2.0, 1),
          (3, 1, 5, 0.0, 0.9, 1), (1, 2, 6, 0.0, 1.1, 1), (2, 3, 7, 0.2, 1.3, 1), (3, 2, 8, 0.2, 1.3, 1),
          (1, 3, 9, 0.3, 0.8, 2), (2, 1, 10, -2.0, 0.9, 1), (3, 1, 11, 1.5, 0.9, 2), (1, 2, 12, 0.25, 1.2, 2)]


def test_variable_biases_all_types():
    """all twelve variable-bias types at once (energy term BM_ELAMBDA and
    dU/dlambda)"""
    s = _check(_variant("msld_small", var_bias=VB_ALL), "variable biases 1-12")
    top, rp = s.topology(), s.runparams()
    sp = spec("msld_small")
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    _, do, eo, _ = orc.forces(top, rp, s.positions(), lam, orc.MSLDBIAS)
    eg = s.energy()
    assert abs(eg[11] - eo[11]) <= 1e-8 * max(1.0, abs(eo[11])), (eg[11], eo[11])


@pytest.mark.parametrize("name", ["config2_small", "config2"])
def test_stale_list_forces_match_oracle(name):
    """The list is built at step 0 with a rebuild cadence of 20 (pad
    2 sqrt(30 kT) 20 dt); after 19 steps, on that list, one force and
    energy evaluation (pack only, no rebuild) meets the path's bars: no
    pair has outrun the pad"""
    from tests.helpers import relaxed
    sp = copy.copy(relaxed(name))
    sp.freq_domdec = 20
    s = system.build(sp)
    s.run_dynamics(19, freq_nrg=0)
    L, dev = s.L, s.device()
    assert L.bm_domdec_update(dev, 0) == 0
    assert L.bm_calc_force(dev, 19, 1) == 0
    assert L.bm_sync(dev) == 0
    top, rp = s.topology(), s.runparams()
    assert rp.contents.freq_domdec == 20
    x, v, th, tv = system.state(s, sp.n, sp.nblocks)
    lam = orc.theta_to_lambda(top, rp, th[:sp.nblocks])
    fo, do, eo, pairs = orc.forces(top, rp, system.soa(x), lam)
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    frms = rel_rms(fg, fo)
    erel = abs(eg[20] - eo[20]) / abs(eo[20])
    ddl = float(np.max(np.abs(dg - do)))
    print("\n[stale list %s, 19 steps] pairs %d force relRMS %.3e energy rel %.3e dudl %.2e"
          % (name, pairs, frms, erel, ddl), flush=True)
    assert frms <= FORCE_TOL and erel <= ENERGY_TOL and ddl <= DUDL_TOL


@pytest.mark.parametrize("removescaling", [False, True])
def test_soft_bonds(removescaling):
    """MSLD soft bonds on the three substituent attachment bonds of the small
    MSLD system (radius 0.8 A, bond power 2, angle/dihedral/improper power
    1.5; with removescaling bond angle the soft terms keep their scaling):
    forces, energy and dU/dlambda against the oracle, each bonded term's
    energy to 1e-9 relative"""
    from tests.test_oracle import _soft_spec
    sp = _soft_spec(removescaling=removescaling)
    s = _check(sp, "soft bonds" + (", removescaling" if removescaling else ""))
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    eo = orc.forces(top, rp, s.positions(), lam)[2]
    eg = s.energy()
    for k in range(5):   # bond, angle, Urey-Bradley, dihedral, improper
        assert abs(eg[k] - eo[k]) <= 1e-9 * abs(eo[k]) + 1e-6, (k, eg[k], eo[k])
    assert top.contents.array("bond_soft", top.contents.n_bonds).sum() == 3


def test_pme_phase_timing_leaves_the_state_intact():
    """bm_time_pme (measurement hook): positive phase times, and a later
    energy evaluation and 10 steps give the same results as without it"""
    import ctypes as C
    sp = spec("msld_small")
    runs = []
    for timed in (False, True):
        s = system.build(sp)
        s.run_energy()
        if timed:
            out = np.zeros(3)
            assert s.L.bm_time_pme(s.device(), 3, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
            print("\n[bm_time_pme] spread %.4f ms, FFT %.4f ms, gather %.4f ms" % tuple(out), flush=True)
            assert np.all(out > 0)
        s.run_energy()
        f = s.forces()
        s.run_dynamics(10, freq_nrg=10)
        runs.append((f, s.positions()))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
