"""MSLD pieces of the oracle pinned independently (CPU):

  * variable biases, all 12 types (msld/msld.cu:1042-1129): dU/dlambda
    equals central differences of the bias energy in lambda, at points
    where each type's active branch applies;
  * theta biases (msld.cu:1155-1264): dU/dtheta equals the theta
    derivative of the energy for the independent, flat-bottom and edge
    biases, and twice it for the collective bias (the reference's force);
  * the piecewise lambda constraint (msld.cu:785-861): lambdas are
    non-negative and sum to one per site, limiting cases, dU/dtheta is the
    theta derivative of U(lambda(theta)), and theta dynamics sample the
    Boltzmann distribution in theta (quadrature over a box);
  * the fix flag (msld.cu:913-935): lambdas stay at the initial theta.
"""
import numpy as np
import pytest

from blade_amd import synthetic, system
from oracle import oracle as orc

KB = 0.001987191
E_LAMBDA, E_THETA = 11, 12


PROBE_ARGS = ("var_bias", "theta_inde", "msld_gamma_ps", "gamma_ps", "bias")


def _probe(**kw):
    args = {k: kw.pop(k) for k in list(kw) if k in PROBE_ARGS}
    sp = synthetic.ensemble_probe(nsites=kw.pop("nsites", 2), **args)
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


def _built(sp):
    s = system.build(sp)
    s.prepare()
    return s, s.topology(), s.runparams()


# (type, l0, k, n, lambda_i, lambda_j): lambda_i on the branch that applies
VB_CASES = [(1, 0.8, 1.5, 2, 0.3, 0.4), (2, 0.1, 1.5, 3, 0.3, 0.4), (3, 0.0, 0.7, 2, 0.3, 0.6),
            (4, 0.6, 2.0, 1, 0.3, 0.4), (5, 0.0, 0.9, 1, 0.3, 0.4), (6, 0.0, 1.1, 1, 0.3, 0.6),
            (7, 0.2, 1.3, 1, 0.3, 0.4), (8, 0.2, 1.3, 1, 0.3, 0.6), (9, 0.3, 0.8, 2, 0.3, 0.6),
            (10, -2.0, 0.9, 1, 0.3, 0.6), (11, 1.5, 0.9, 2, 0.3, 0.6), (12, 0.25, 1.2, 2, 0.3, 0.6)]


@pytest.mark.parametrize("ty,l0,k,n,li,lj", VB_CASES)
def test_variable_bias_dudl_is_the_lambda_derivative(ty, l0, k, n, li, lj):
    """dU/dlambda of each type against central differences of its energy.
    Type 9 keeps the reference's lambda_i force -k lambda_j ((lambda_i +
    l0)/l0)^(n-1) / l0 (msld.cu:1087-1090), which lacks the factor n of the
    energy's derivative: it is 1/n of the difference quotient."""
    sp = _probe(var_bias=[(1, 2, ty, l0, k, n)])
    s, top, rp = _built(sp)
    x = system.soa(sp.pos)
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    lam[1], lam[2] = li, lj
    fb = np.array(sp.fixed_bias)
    _, d, e, _ = orc.forces(top, rp, x, lam, orc.MSLDBIAS)
    h = 1e-6
    for b in (1, 2):
        lp, lm = lam.copy(), lam.copy()
        lp[b] += h
        lm[b] -= h
        fd = (orc.forces(top, rp, x, lp, orc.MSLDBIAS)[2][E_LAMBDA]
              - orc.forces(top, rp, x, lm, orc.MSLDBIAS)[2][E_LAMBDA]) / (2 * h)
        want = fd / n if (ty == 9 and b == 1) else fd
        assert abs(want - d[b]) < 1e-6 * max(1.0, abs(d[b])), (ty, b, fd, d[b])
    # the term is active: energy beyond the fixed biases
    assert abs(e[E_LAMBDA] - fb @ lam) > 1e-3


def _theta_fd(top, rp, th, h=1e-6):
    g, e = orc.theta_bias(top, rp, th)
    fd = np.zeros_like(th)
    for b in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[b] += h
        tm[b] -= h
        fd[b] = (orc.theta_bias(top, rp, tp)[1][E_THETA] - orc.theta_bias(top, rp, tm)[1][E_THETA]) / (2 * h)
    return g, e[E_THETA], fd


def test_theta_independent_and_collective_biases():
    """independent bias: force = dE/dtheta; collective bias: the reference's
    force k n (A h^(n-1) - B l^(n-1)) cos theta is twice dE/dtheta of its
    reported energy k/2 (A^2 + B^2)"""
    rng = np.random.default_rng(3)
    sp = _probe(theta_inde=[0.0, 0.4, 0.7])
    s, top, rp = _built(sp)
    th = rng.uniform(-np.pi, np.pi, sp.nblocks)
    g, e, fd = _theta_fd(top, rp, th)
    assert e != 0.0
    assert np.max(np.abs(g - fd)) < 1e-6, (g, fd)
    sp = _probe(theta_coll=[(0.0, 2.0), (1.5, 2.0), (0.8, 3.0)])
    s, top, rp = _built(sp)
    g, e, fd = _theta_fd(top, rp, th)
    assert e != 0.0
    assert np.max(np.abs(g - 2 * fd)) < 1e-6, (g, fd)
    assert abs(g[0]) == 0.0


def _piecewise(edge=None, nsites=2):
    sp = _probe(nsites=nsites, piecewise=(1.0, 100.0), theta_edge=edge)
    return sp, *_built(sp)


def test_piecewise_lambdas_sum_to_one_and_limits():
    sp, s, top, rp = _piecewise()
    rng = np.random.default_rng(4)
    for _ in range(20):
        th = np.concatenate([[0.0], rng.uniform(-3, 3, sp.nblocks - 1)])
        lam = orc.theta_to_lambda(top, rp, th)
        assert np.all(lam >= 0)
        for st in range(1, 3):
            assert abs(lam[1 + 3 * (st - 1):1 + 3 * st].sum() - 1) < 1e-12
    # equal thetas share equally; a block more than 1 above the rest takes all
    th = np.array([0.0, 0.2, 0.2, 0.2, 2.5, 0.3, -0.1])
    lam = orc.theta_to_lambda(top, rp, th)
    assert np.allclose(lam[1:4], 1 / 3, atol=1e-12)
    assert np.allclose(lam[4:7], [1, 0, 0], atol=1e-12)


def test_piecewise_chain_rule_and_biases_are_theta_derivatives():
    sp, s, top, rp = _piecewise(edge=(2.0, 1.2, 10.0, 0.0))
    rng = np.random.default_rng(5)
    F = rng.normal(size=sp.nblocks)
    F[0] = 0.0
    U = lambda th: F @ orc.theta_to_lambda(top, rp, th)
    for th in (np.array([0.0, 0.5, 0.45, -0.3, 3.4, -0.2, 0.1]),
               np.array([0.0, 0.1, 0.6, 0.3, -3.6, 0.2, 0.25])):
        lam = orc.theta_to_lambda(top, rp, th)
        g = orc.chain_rule(top, rp, th, lam, F)
        h = 1e-6
        for b in range(1, sp.nblocks):
            tp, tm = th.copy(), th.copy()
            tp[b] += h
            tm[b] -= h
            assert abs((U(tp) - U(tm)) / (2 * h) - g[b]) < 1e-6
        gb, e, fd = _theta_fd(top, rp, th)
        assert e > 0.0
        assert np.max(np.abs(gb - fd)) < 1e-6, (gb, fd)


def test_piecewise_theta_dynamics_sample_the_boltzmann_distribution():
    """24 sites of 3 blocks, fixed biases (0, 0.6, -0.4), flat-bottom walls
    at +-3 (k 100): <lambda> from 40,000 oracle steps against quadrature
    over theta in [-6, 6]^3 with the root solved at every grid point"""
    ns = 24
    sp = synthetic.ensemble_probe(nsites=ns, msld_gamma_ps=5.0)
    sp.piecewise = (1.0, 100.0)
    sp.theta0 = [0.0] + list(np.random.default_rng(3).uniform(-1, 1, sp.nblocks - 1))
    s, top, rp = _built(sp)
    n, nl = sp.n, sp.nblocks
    x = np.concatenate([system.soa(sp.pos), np.array(sp.theta0)])
    v = np.concatenate([system.soa(sp.vel), np.zeros(nl)])
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0))
    L = []
    for k in range(40000):
        orc.step(top, rp, k, 0, x, v, lam)
        if k >= 500:
            L.append(lam[1:].reshape(ns, 3).copy())
    L = np.array(L)
    G, lo, hi = 120, -6.0, 6.0
    th = lo + (np.arange(G) + 0.5) * (hi - lo) / G
    T = np.stack(np.meshgrid(th, th, th, indexing="ij"), -1).reshape(-1, 3)
    t0 = T.max(1) - 1.0
    for _ in range(60):
        d = np.clip(T - t0[:, None], 0, None)
        t0 = t0 + ((d ** 3).sum(1) - 1) / (3 * d ** 2).sum(1)
    lq = np.clip(T - t0[:, None], 0, None) ** 3
    wall = np.where(T > 3, T - 3, np.where(T < -3, T + 3, 0.0))
    U = lq @ np.array(sp.fixed_bias[1:4]) + 50.0 * (wall ** 2).sum(1)
    p = np.exp(-(U - U.min()) / (KB * sp.temperature))
    exact = (p / p.sum()) @ lq
    sim = L.mean((0, 1))
    nb = 20
    bm = L[:len(L) // nb * nb].reshape(nb, -1, ns, 3).mean(1).reshape(-1, 3)
    se = bm.std(0, ddof=1) / np.sqrt(len(bm))
    print("\npiecewise <lambda> oracle %s exact %s se %s" % (sim, exact, se), flush=True)
    assert np.all(se < 0.02)
    assert np.all(np.abs(sim - exact) < 4.5 * se)


def test_fix_flag_holds_lambda_at_initial_theta():
    sp = _probe(fix_lambda=True)
    sp.theta0 = [1.0, 0.2, 0.3, 0.5, 0.6, 0.3, 0.1]
    s, top, rp = _built(sp)
    n, nl = sp.n, sp.nblocks
    x = np.concatenate([system.soa(sp.pos), np.array(sp.theta0)])
    v = np.concatenate([system.soa(sp.vel), np.full(nl, 0.3)])
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0))
    assert np.array_equal(lam, sp.theta0)
    for k in range(20):
        orc.step(top, rp, k, k % 10 == 0, x, v, lam)
    assert np.array_equal(lam, sp.theta0)
    assert np.array_equal(x[3 * n:], sp.theta0)
