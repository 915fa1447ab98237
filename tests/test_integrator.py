"""The oracle's integrator (orc_step, the VRORV Langevin splitting of
update/update.cu:233-305 with the MSLD theta dynamics of msld/msld.cu) pinned
by what the mathematics fixes rather than by a second copy of its formulas:

  * lambda averages of biased MSLD sites against the Boltzmann distribution
    in theta, integrated by quadrature over [-pi, pi)^nsubs (chain rule,
    fnex, theta friction and noise, fixed/variable/theta biases);
  * harmonic pairs (atom restraints, omega^2 = k/m): the BAOAB/VRORV
    splitting samples the exact configurational variance <r^2> = 2 kT / k
    at any stable step (Leimkuhler & Matthews, AMRX 2013);
  * free particles: Ornstein-Uhlenbeck velocities with <v^2> = kT / m;
  * NVE (gamma = 0) on the reference's own water216 deck with SETTLE: no
    energy drift;
  * equipartition <KE> = N_dof kT / 2 on the same deck with the thermostat.
CPU only.
"""
import functools
import os

import numpy as np

from blade_amd import synthetic, system
from oracle import oracle as orc

KB = 0.001987191
HERE = os.path.dirname(os.path.abspath(__file__))
W216 = os.path.join(HERE, "golden", "water216")

NSITES, NSTEPS, BURN = 24, 40000, 500


def probe_spec():
    vb = [(1 + 3 * s, 2 + 3 * s, 6, 0.0, 0.5, 1) for s in range(NSITES)]
    return synthetic.ensemble_probe(nsites=NSITES, var_bias=vb, theta_inde=[0.0] + [0.3] * NSITES)


def lambda_quadrature(sp, grid=96):
    """<lambda> of one site of the probe: U = sum_b F_b lambda_b + k_vb
    lambda_1 lambda_2 + sum_b k_i (1 - l_b^4), l = (1 - sin theta)/2, with
    the softmax lambda = exp(fnex sin theta)/sum, over a periodic grid"""
    th = np.arange(grid) * 2 * np.pi / grid - np.pi
    T = np.stack(np.meshgrid(th, th, th, indexing="ij"), -1).reshape(-1, 3)
    w = np.exp(sp.fnex * np.sin(T))
    lq = w / w.sum(1, keepdims=True)
    F = np.array(sp.fixed_bias[1:4])
    U = lq @ F + 0.5 * lq[:, 0] * lq[:, 1] + 0.3 * (1 - (0.5 - 0.5 * np.sin(T)) ** 4).sum(1)
    kT = KB * sp.temperature
    p = np.exp(-(U - U.min()) / kT)
    return (p / p.sum()) @ lq


@functools.lru_cache(maxsize=None)
def probe_run():
    sp = probe_spec()
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n, nl = sp.n, sp.nblocks
    x = np.concatenate([system.soa(sp.pos), np.array(sp.theta0, float)])
    v = np.concatenate([system.soa(sp.vel), np.zeros(nl)])
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    pairs = np.array(sp.atom_restraints)
    free = np.arange(2 * len(pairs), 2 * len(pairs) + 12)
    L, R, VC, VF, VT = [], [], [], [], []
    for k in range(NSTEPS):
        orc.step(top, rp, k, 0, x, v, lam)
        if k < BURN:
            continue
        X, V = x[:3 * n].reshape(3, n), v[:3 * n].reshape(3, n)
        L.append(lam[1:].reshape(NSITES, 3).copy())
        R.append(((X[:, pairs[:, 0]] - X[:, pairs[:, 1]]) ** 2).mean())
        VC.append(((0.5 * (V[:, pairs[:, 0]] + V[:, pairs[:, 1]])) ** 2).mean())
        VF.append((V[:, free] ** 2).mean())
        VT.append((v[3 * n + 1:] ** 2).mean())
    s.close()
    return sp, np.array(L), np.array(R), np.array(VC), np.array(VF), np.array(VT)


def batch_se(series, nb=20):
    """standard error of the mean of a correlated series by batch means"""
    series = np.asarray(series)
    m = len(series) // nb * nb
    b = series[:m].reshape(nb, -1, *series.shape[1:]).mean(1)
    return b.std(0, ddof=1) / np.sqrt(nb)


def test_theta_dynamics_sample_the_boltzmann_lambda_distribution():
    """24 identical sites of 3 substituents with fixed biases (0, 0.6,
    -0.4 kcal/mol), a product bias 0.5 lambda_1 lambda_2 and independent
    theta biases: <lambda> from 40,000 oracle steps against quadrature.
    A missing fnex or cos factor in the chain rule, a wrong sign or a wrong
    theta noise amplitude moves <lambda> by many standard errors."""
    sp, L, _, _, _, _ = probe_run()
    exact = lambda_quadrature(sp)
    sim = L.mean((0, 1))
    nb = 20
    m = len(L) // nb * nb
    bm = L[:m].reshape(nb, -1, NSITES, 3).mean(1).reshape(-1, 3)   # batches x sites
    se = bm.std(0, ddof=1) / np.sqrt(len(bm))
    z = (sim - exact) / se
    print("\n<lambda> oracle %s exact %s se %s z %s" % (sim, exact, se, z), flush=True)
    assert np.all(se < 0.015)
    assert np.all(np.abs(z) < 4.5)
    # the biases matter: the exact averages are far from uniform
    assert np.max(np.abs(exact - 1 / 3)) > 0.15


def test_harmonic_pairs_have_the_exact_configurational_variance():
    """atom-restrained pairs with omega dt = 0.7: <r_x^2> = 2 kT / k (exact
    for the VRORV splitting; a wrong noise amplitude, friction factor or
    kick/drift split gives an O(1) or O((omega dt)^2) = 12% error)"""
    sp, _, R, _, _, _ = probe_run()
    kT = KB * sp.temperature
    exact = 2 * kT / sp.k_restraint
    ratio = R.mean() / exact
    se = batch_se(R) / exact
    omega_dt = np.sqrt(sp.k_restraint / 12.0) * sp.dt
    print("\nomega dt %.3f  <r^2>/exact %.4f +- %.4f" % (omega_dt, ratio, se), flush=True)
    assert 0.6 < omega_dt < 0.8
    assert abs(ratio - 1) < max(4 * se, 0.01) and se < 0.01


def test_free_particle_and_pair_centre_velocities_are_maxwellian():
    """free atoms and the pairs' centres of mass (no force on either):
    Ornstein-Uhlenbeck velocities with <v^2> = kT/m per component"""
    sp, _, _, VC, VF, VT = probe_run()
    kT = KB * sp.temperature
    for name, series, mass in (("pair centre", VC, 24.0), ("free atom", VF, 12.0)):
        ratio = series.mean() / (kT / mass)
        se = batch_se(series) / (kT / mass)
        print("\n%s <v^2> m/kT %.4f +- %.4f" % (name, ratio, se), flush=True)
        assert abs(ratio - 1) < max(4 * se, 0.01)
    # theta velocities: the theta mass sets their scale, <v^2> m/kT ~ 1
    # within the splitting's small velocity bias
    rt = VT.mean() * 12.0 / kT
    print("theta <v^2> m/kT %.4f" % rt, flush=True)
    assert abs(rt - 1) < 0.06


@functools.lru_cache(maxsize=None)
def water216():
    s = system.load_charmm(os.path.join(W216, "par_water_ions.prm"), os.path.join(W216, "minimized.psf"),
                           os.path.join(W216, "minimized.crd"), (18.856,) * 3, beta=0.34, rcut=9.0,
                           rswitch=8.0, dt_ps=0.002, temperature=298.0, gamma_ps=0.0)
    s.prepare()
    return s


def _water_run(gamma_ps, nsteps, seed=5, v=None, x=None):
    s = water216()
    top, rp = s.topology(), s.runparams()
    rp.contents.gamma = gamma_ps / synthetic.PICOSECOND
    t = top.contents
    n = t.n_atoms
    kT = KB * 298.0
    if x is None:
        x = np.concatenate([s.positions(), [0.0]])
        m = np.array([t.mass[i] for i in range(n)])
        vv = (np.random.default_rng(seed).normal(size=(3, n)) * np.sqrt(kT / m)).ravel()
        v = np.concatenate([orc.holonomic_velocity(top, rp, x[:3 * n], vv), [0.0]])
    lam = np.ones(1)
    E = []
    for k in range(nsteps):
        e = orc.step(top, rp, k, k % 10 == 0, x, v, lam)
        if k % 10 == 0:
            E.append((k, e[20], e[21]))
    return np.array(E), x, v


def test_nve_energy_is_conserved_water216():
    """gamma = 0, 1,000 steps of 2 fs (2 ps) from the reference's water216
    deck (648 atoms, SETTLE, PME 20^3, force-switched vdW 8-9 A): the total
    energy drifts by less than 0.5 kcal/mol/ps (measured 0.02) and
    fluctuates by less than 0.3 kcal/mol rms (measured 0.08), against
    ~360 kcal/mol of kinetic energy"""
    E, _, _ = _water_run(0.0, 1000)
    tot = E[:, 1] + E[:, 2]
    slope = np.polyfit(E[:, 0] * 0.002, tot, 1)[0]
    print("\nNVE drift %.3f kcal/mol/ps, rms %.3f kcal/mol, <KE> %.1f" % (slope, tot.std(), E[:, 2].mean()),
          flush=True)
    assert abs(slope) < 0.5
    assert tot.std() < 0.3


def test_equipartition_water216():
    """500 steps at gamma = 10/ps to reach the bath, then 1,000 steps at
    gamma = 1/ps: <KE> = (3N - 3 N_water) kT / 2 within 5% (the statistical
    error of 2 ps of 648 atoms is about 2%)"""
    E0, x, v = _water_run(10.0, 500)
    E, _, _ = _water_run(1.0, 1000, x=x, v=v)
    s = water216()
    t = s.topology().contents
    ndof = 3 * t.n_atoms - 3 * t.n_tri
    want = 0.5 * ndof * KB * 298.0
    got = E[:, 2].mean()
    print("\n<KE> %.2f expected %.2f (ratio %.4f)" % (got, want, got / want), flush=True)
    assert abs(got / want - 1) < 0.05
