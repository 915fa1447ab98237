"""MSLD theta dynamics on the GPU (run on an MI355X), through the C ABI:

  * the GPU trajectory of atoms, theta, theta velocities and lambda against
    the FP64 oracle over 10 steps from the same state and Philox stream, for
    the softmax and piecewise constraints, theta biases, per-block theta
    friction, a fixed block and the fix flag;
  * dU/dtheta (chain rule plus theta biases) of the first step against the
    oracle;
  * the GPU integrator's sampling against closed forms: lambda averages of
    biased sites against quadrature over theta, the exact configurational
    variance of harmonic pairs and Maxwellian free-particle velocities.
"""
import copy

import numpy as np
import pytest

from blade_amd import synthetic, system
from oracle import oracle as orc
from tests.helpers import relaxed, spec
from tests import test_integrator as ti

pytestmark = pytest.mark.gpu
KB = 0.001987191


def _probe(**kw):
    sp = synthetic.ensemble_probe(nsites=kw.pop("nsites", 4), **{k: kw.pop(k) for k in list(kw)
                                                                  if k in ("var_bias", "theta_inde", "msld_gamma_ps")})
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


def _variant(name):
    if name == "msld_small":
        return relaxed("msld_small")
    if name == "two_site_small":
        return relaxed("two_site_small")
    if name == "friction_fixed":
        sp = copy.copy(relaxed("msld_small"))
        sp.block_friction = [None, 2.0 / synthetic.PICOSECOND, 20.0 / synthetic.PICOSECOND, None]
        sp.block_fixed = [0, 0, 0, 1]
        sp.theta0 = list(sp.theta0[:3]) + [0.2]
        return sp
    if name == "theta_biases":
        return _probe(theta_inde=[0.0] + [0.4] * 4, theta_coll=[(0.0, 2.0)] + [(1.2, 2.0)] * 4,
                      var_bias=[(1, 2, 6, 0.0, 0.5, 1), (4, 5, 3, 0.0, 0.7, 2)])
    if name == "piecewise":
        sp = _probe(piecewise=(1.0, 100.0), theta_edge=(2.0, 1.2, 10.0, 0.0))
        sp.theta0 = [0.0] + [0.5, 0.45, -0.3, 3.4, -0.2, 0.1, 0.1, 0.6, 0.3, -3.6, 0.2, 0.25]
        sp.theta_v = [0.0] + [0.5, -0.4, 0.3] * 4
        return sp
    if name == "fix":
        sp = _probe(fix_lambda=True)
        sp.theta0 = [1.0] + [0.2, 0.3, 0.5] * 4
        return sp
    raise KeyError(name)


VARIANTS = ["msld_small", "two_site_small", "friction_fixed", "theta_biases", "piecewise", "fix"]


def _oracle_state(sp, top, rp):
    n = sp.n
    x = np.concatenate([orc.rectify(top, rp, system.soa(sp.pos)), np.array(sp.theta0, float)])
    v = np.concatenate([system.soa(sp.vel), np.array(sp.theta_v if sp.theta_v is not None
                                                      else [0.0] * sp.nblocks, float)])
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    x[:3 * n] = orc.domdec(top, rp, x[:3 * n])[0]      # the GPU wraps at step 0's rebuild
    return x, v, lam


@pytest.mark.parametrize("name", VARIANTS)
def test_theta_trajectory_tracks_oracle(name):
    """10 steps (energy steps 0 and 5): positions to 1e-4 A, theta to 1e-6,
    theta velocities to 1e-5 relative to their scale, lambda to 1e-6"""
    sp = _variant(name)
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n, nl = sp.n, sp.nblocks
    x, v, lam = _oracle_state(sp, top, rp)
    for k in range(10):
        orc.step(top, rp, k, k % 5 == 0, x, v, lam)
    s.run_dynamics(10, freq_nrg=5)
    xg, vg, thg, tvg = system.state(s, n, nl)
    dx = np.max(np.abs(system.soa(xg) - x[:3 * n]))
    dth = np.max(np.abs(thg[:nl] - x[3 * n:]))
    vscale = max(1e-3, float(np.max(np.abs(v[3 * n:]))))
    dtv = np.max(np.abs(tvg[:nl] - v[3 * n:])) / vscale
    dl = np.max(np.abs(s.lambdas() - lam))
    print("\n[%s] max|dx| %.2e  max|dtheta| %.2e  theta-velocity rel %.2e  max|dlambda| %.2e"
          % (name, dx, dth, dtv, dl), flush=True)
    assert dx < 1e-4
    assert dth < 1e-6 and dl < 1e-6
    assert dtv < 1e-5
    if name == "fix":
        assert np.array_equal(s.lambdas(), np.array(sp.theta0))
    if name == "friction_fixed":
        assert thg[3] == 0.2 and abs(s.lambdas()[3] - 0.2) < 1e-15


@pytest.mark.parametrize("name", VARIANTS)
def test_theta_force_matches_oracle(name):
    """dU/dtheta of step 0 (written by the first step's kick) against the
    oracle's chain rule plus theta biases at the initial state"""
    sp = _variant(name)
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n, nl = sp.n, sp.nblocks
    x, v, lam = _oracle_state(sp, top, rp)
    _, dudl, _, _ = orc.forces(top, rp, x[:3 * n], lam)
    th = x[3 * n:]
    want = orc.chain_rule(top, rp, th, lam, dudl) + orc.theta_bias(top, rp, th)[0]
    s.run_dynamics(1, freq_nrg=1)
    got = np.zeros(nl)
    f = np.zeros(3 * n)
    assert s.L.bm_get_forces(s.device(), f.ctypes.data_as(system._ffi._f64p),
                             got.ctypes.data_as(system._ffi._f64p)) == 0
    err = np.max(np.abs(got - want))
    print("\n[%s] dU/dtheta gpu %s oracle %s max err %.2e" % (name, got, want, err), flush=True)
    assert err < 1e-3
    e = s.energy()
    eo = orc.theta_bias(top, rp, th)[1][12]
    assert abs(e[12] - eo) < 1e-6 * max(1.0, abs(eo))


def test_gpu_theta_sampling_and_harmonic_pairs():
    """the oracle's ensemble checks, run by the GPU integrator: 24 biased
    sites (<lambda> against quadrature), harmonic pairs (<r^2> = 2 kT/k)
    and free particles (<v^2> = kT/m), 40,000 steps sampled every 10"""
    sp = ti.probe_spec()
    s = system.build(sp)
    n, ns = sp.n, ti.NSITES
    pairs = np.array(sp.atom_restraints)
    free = np.arange(2 * len(pairs), 2 * len(pairs) + 12)
    L, R, VF = [], [], []
    s.run_dynamics(500, freq_nrg=0)
    for _ in range(3950):
        s.run_dynamics(10, freq_nrg=0)
        x, v, th, tv = system.state(s, n, sp.nblocks)
        L.append(s.lambdas()[1:].reshape(ns, 3).copy())
        d = x[pairs[:, 0]] - x[pairs[:, 1]]
        d -= np.array(sp.box) * np.round(d / np.array(sp.box))
        R.append((d ** 2).mean())
        VF.append((v[free] ** 2).mean())
    L = np.array(L)
    exact = ti.lambda_quadrature(sp)
    sim = L.mean((0, 1))
    nb = 20
    bm = L[:len(L) // nb * nb].reshape(nb, -1, ns, 3).mean(1).reshape(-1, 3)
    se = bm.std(0, ddof=1) / np.sqrt(len(bm))
    kT = KB * sp.temperature
    rr = np.mean(R) / (2 * kT / sp.k_restraint)
    vv = np.mean(VF) / (kT / 12.0)
    print("\n<lambda> gpu %s exact %s se %s; <r^2> ratio %.4f; <v^2> ratio %.4f" % (sim, exact, se, rr, vv),
          flush=True)
    assert np.all(se < 0.015) and np.all(np.abs(sim - exact) < 4.5 * se)
    assert abs(rr - 1) < 4 * ti.batch_se(np.array(R)) / (2 * kT / sp.k_restraint) + 0.01
    assert abs(vv - 1) < 0.05
