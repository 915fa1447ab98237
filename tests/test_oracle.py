"""Pins the FP64 oracle (CPU only).

The reference ships no golden vectors for the force/integration path, so
each oracle function is checked against something the mathematics fixes
independently of its own formulas:

  philox           Random123 known-answer vectors for Philox4x32-10
  normal           moments of the Box-Muller stream
  bspline          partition of unity, symmetry, closed-form values
  fft3d            numpy.fft
  PME              Madelung constant of rock salt; direct Ewald k-sum;
                   direct real-space image sum
  every force term central finite differences of the energy
  dU/dlambda       finite differences in lambda
  chain rule       finite differences through theta -> lambda
  switch functions continuity of energy and force at rswitch, zero at rcut
  SETTLE/SHAKE     constraint lengths after the position step; zero bond
                   rate after the velocity projection
  rectify          lengths, kept directions and planes, kept images; a
                   water against its geometry built by hand
  domdec           every pair within rcut lies in a listed tile; masks mark
                   exactly the excluded pairs
"""
import math
import os

import numpy as np
import pytest

from blade_amd import synthetic, system
from oracle import oracle as orc
from tests.helpers import built, spec

KE = 332.0716
E_BOND, E_ANGLE, E_UREY, E_DIHE, E_IMPR, E_CMAP, E_NB14, E_NBDIRECT, E_NBRECIP, E_SELF, E_EXCL, E_LAMBDA = range(12)
E_POT = 20
W216 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "water216")


# ---------------------------------------------------------------------------
# random numbers
# ---------------------------------------------------------------------------
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert orc.philox(ctr, key) == want


def test_normal_moments():
    x = np.array([orc.normal(0x1234, step, dof) for step in range(40) for dof in range(500)])
    assert abs(x.mean()) < 0.02
    assert abs(x.var() - 1) < 0.03
    assert abs(np.mean(x ** 4) - 3) < 0.15
    # the two normals of one Box-Muller pair are independent
    a = np.array([orc.normal(7, s, 2 * k) for s in range(20) for k in range(200)])
    b = np.array([orc.normal(7, s, 2 * k + 1) for s in range(20) for k in range(200)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.05


# ---------------------------------------------------------------------------
# B-splines and FFT
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 6, 8])
def test_bspline_identities(n):
    for u in np.linspace(0.0, 1.0, 7, endpoint=False):
        s = sum(orc.bspline(n, u + t) for t in range(n))
        assert abs(s - 1) < 1e-14
    for u in np.linspace(0.1, n - 0.1, 13):
        assert abs(orc.bspline(n, u) - orc.bspline(n, n - u)) < 1e-14
    assert orc.bspline(n, 0.0) == 0.0 and orc.bspline(n, float(n)) == 0.0


def test_bspline_closed_form_cubic():
    # M_4 at the integers: 1/6, 2/3, 1/6
    assert abs(orc.bspline(4, 1.0) - 1 / 6) < 1e-15
    assert abs(orc.bspline(4, 2.0) - 2 / 3) < 1e-15
    assert abs(orc.bspline(4, 3.0) - 1 / 6) < 1e-15
    # M_2 is the hat function
    for u in (0.25, 0.5, 1.5):
        assert abs(orc.bspline(2, u) - (1 - abs(u - 1))) < 1e-15


def test_bspline_moduli_match_definition():
    # 1/|sum_{k=0}^{n-2} M_n(k+1) exp(2 pi i m k / K)|^2
    for n, K in ((4, 20), (6, 64), (8, 27)):
        got = orc.bspline_moduli(n, K)
        m = np.arange(K)
        s = sum(orc.bspline(n, k + 1.0) * np.exp(2j * np.pi * m * k / K) for k in range(n - 1))
        assert np.allclose(got, 1 / np.abs(s) ** 2, rtol=1e-12)


@pytest.mark.parametrize("shape", [(8, 6, 10), (20, 20, 20), (16, 27, 18)])
def test_fft3d_matches_numpy(shape):
    rng = np.random.default_rng(3)
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    assert np.allclose(orc.fft3d(g, -1), np.fft.fftn(g), atol=1e-9)
    assert np.allclose(orc.fft3d(g, +1), np.fft.ifftn(g) * g.size, atol=1e-9)


# ---------------------------------------------------------------------------
# Ewald electrostatics
# ---------------------------------------------------------------------------
def _ions(positions, charges, box, beta=0.34, rcut=10.0, rswitch=9.0, grid=32, order=6):
    s = synthetic.Spec(name="ions", box=tuple(box))
    s.types.update({"NA": 22.99, "CL": 35.45})
    s.nb_par.update({"NA": (0.0, 1.5, 0.0, 1.5, 1.0, 0), "CL": (0.0, 2.0, 0.0, 2.0, 1.0, 0)})
    for i, q in enumerate(charges):
        t = "NA" if q > 0 else "CL"
        s.atoms.append(("ION", str(i + 1), "ION", t, t, float(q), s.types[t]))
    s.pos = np.asarray(positions, float)
    s.atom_block = np.zeros(len(charges), dtype=np.int32)
    s.shake = False
    s.beta, s.rcut, s.rswitch, s.grid, s.order = beta, rcut, rswitch, (grid, grid, grid), order
    return s


def _rock_salt(ncell, a):
    base_na = [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0)]
    pos, q = [], []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                for b in base_na:
                    p = (np.array(b) + (i, j, k)) * a
                    pos.append(p)
                    q.append(1.0)
                    pos.append(p + (a / 2, 0, 0))
                    q.append(-1.0)
    return np.array(pos), np.array(q)


def test_pme_madelung_rock_salt():
    a, ncell = 5.64, 4
    pos, q = _rock_salt(ncell, a)
    box = (ncell * a,) * 3
    s = system.build(_ions(pos, q, box))
    s.prepare()
    f, d, e, _ = orc.forces(s.topology(), s.runparams(), system.soa(pos), np.ones(1))
    madelung = 1.7475645946331822
    want = -(len(q) // 2) * madelung * KE / (a / 2)
    got = e[E_NBDIRECT] + e[E_NBRECIP] + e[E_SELF]
    assert abs(got - want) / abs(want) < 2e-6
    # every ion of a perfect lattice is at a force minimum
    assert np.abs(f).max() < 1e-4


def _random_ions(n, box, seed):
    rng = np.random.default_rng(seed)
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)   # neutral for even n
    rng.shuffle(q)
    # keep ions at least 2 A apart
    pos = []
    while len(pos) < n:
        p = rng.uniform(0, box, size=3)
        if all(np.linalg.norm((p - o) - box * np.round((p - o) / box)) > 2.0 for o in pos):
            pos.append(p)
    return np.array(pos), q


def test_pme_reciprocal_matches_direct_kspace_sum():
    box = 20.0
    pos, q = _random_ions(40, box, 11)
    s = system.build(_ions(pos, q, (box,) * 3, beta=0.34, rcut=9.0, rswitch=8.0, grid=40, order=8))
    s.prepare()
    top, rp = s.topology(), s.runparams()
    x = system.soa(pos)
    fp, dp, ep, _ = orc.forces(top, rp, x, np.ones(1), orc.PME)
    e, fd, dd = orc.ewald_recip_direct(top, rp, x, np.ones(1), 14)
    assert abs(ep[E_NBRECIP] - e) / abs(e) < 1e-6
    assert np.sqrt(np.mean((fp - fd) ** 2)) / np.sqrt(np.mean(fd ** 2)) < 1e-5


def test_real_space_matches_image_sum():
    box = 20.0
    pos, q = _random_ions(30, box, 5)
    s = system.build(_ions(pos, q, (box,) * 3, beta=0.34, rcut=9.0, rswitch=8.0, grid=20))
    s.prepare()
    f, d, e, _ = orc.forces(s.topology(), s.runparams(), system.soa(pos), np.ones(1), orc.NBDIRECT)
    want = orc.ewald_real_images(q, system.soa(pos), np.array([box] * 3), 0.34, 9.0)
    assert abs(e[E_NBDIRECT] - want) < 1e-9 * abs(want) + 1e-9


# ---------------------------------------------------------------------------
# finite differences of every term
# ---------------------------------------------------------------------------
TERM_GROUPS = [
    ("bonded", orc.BONDED, [E_BOND, E_ANGLE, E_UREY, E_DIHE, E_IMPR]),
    ("nb14", orc.NB14, [E_NB14]),
    ("nbex", orc.NBEX, [E_EXCL]),
    ("nbdirect", orc.NBDIRECT, [E_NBDIRECT]),
    ("pme", orc.PME, [E_NBRECIP]),
]


def _fd_check(sp_name, terms, eidx, atoms, lam=None, h=1e-5, tol=2e-6):
    s = built(sp_name)
    top, rp = s.topology(), s.runparams()
    n = spec(sp_name).n
    x0 = system.soa(spec(sp_name).pos)
    if lam is None:
        lam = lambdas_of(sp_name, s)
    f, _, _, _ = orc.forces(top, rp, x0, lam, terms)
    scale = max(1.0, float(np.abs(f).max()))
    worst = 0.0
    for a in atoms:
        for k in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[k * n + a] += h
            xm[k * n + a] -= h
            ep = sum(orc.forces(top, rp, xp, lam, terms)[2][i] for i in eidx)
            em = sum(orc.forces(top, rp, xm, lam, terms)[2][i] for i in eidx)
            fd = (ep - em) / (2 * h)
            worst = max(worst, abs(fd - f[k * n + a]) / scale)
    assert worst < tol, worst


def lambdas_of(name, s):
    t = s.topology().contents
    if t.n_blocks <= 1:
        return np.ones(1)
    return orc.theta_to_lambda(s.topology(), s.runparams(), np.array(spec(name).theta0, float))


def _polymer_atoms(name, k=6):
    sp = spec(name)
    idx = [i for i, a in enumerate(sp.atoms) if a[0] in ("POLY", "LIG")]
    rng = np.random.default_rng(0)
    return list(rng.choice(idx, size=min(k, len(idx)), replace=False)) + [idx[0]]


@pytest.mark.parametrize("group", TERM_GROUPS, ids=[g[0] for g in TERM_GROUPS])
def test_forces_are_energy_gradients(group):
    name, terms, eidx = group
    atoms = _polymer_atoms("msld_small")
    if name in ("nbdirect", "pme"):
        atoms = atoms[:3] + [0, 1, 2]
    _fd_check("msld_small", terms, eidx, atoms)


def test_atom_restraints_pairwise_identity_and_gradient():
    """MSLD atom restraints: with the restrained atoms moved apart (one
    group across a periodic face), the energy equals the pairwise form
    k/(2n) sum_{i<j} |x_i - x_j|^2 over minimum-image differences, and
    the forces are its gradient."""
    name = "two_site_small"
    sp = spec(name)
    s = built(name)
    top, rp = s.topology(), s.runparams()
    t = top.contents
    assert t.n_cats == len(sp.atom_restraints) >= 2
    n = sp.n
    L = np.array(sp.box)
    x = system.soa(sp.pos).copy()
    rng = np.random.default_rng(5)
    groups = [list(g) for g in sp.atom_restraints]
    for g in groups:
        for a in g:
            for k in range(3):
                x[k * n + a] += rng.uniform(-0.6, 0.6)
    a0 = groups[0][0]
    x[a0] = 0.05                      # first group: one atom just inside x = 0 ...
    x[groups[0][1]] = L[0] - 0.2      # ... and one just below x = L
    lam = lambdas_of(name, s)
    f, _, e, _ = orc.forces(top, rp, x, lam, orc.MSLDBIAS)
    want = 0.0
    for g in groups:
        P = np.array([[x[k * n + a] for k in range(3)] for a in g])
        for i in range(len(g)):
            for j in range(i + 1, len(g)):
                d = P[i] - P[j]
                d -= L * np.round(d / L)
                want += sp.k_restraint / (2 * len(g)) * float(d @ d)
    assert abs(e[13] - want) <= 1e-9 * max(1.0, want), (e[13], want)
    h, worst = 1e-5, 0.0
    for g in groups:
        for a in g:
            for k in range(3):
                xp, xm = x.copy(), x.copy()
                xp[k * n + a] += h
                xm[k * n + a] -= h
                fd = (orc.forces(top, rp, xp, lam, orc.MSLDBIAS)[2][13]
                      - orc.forces(top, rp, xm, lam, orc.MSLDBIAS)[2][13]) / (2 * h)
                worst = max(worst, abs(fd - f[k * n + a]))
    assert worst < 1e-5, worst


E_BIAS = 19
WIGNER_SC = 2.837297479480620   # Wigner lattice-sum constant of a simple cubic box


def set_block_charges(sp, q_blocks):
    """shift the charges of each block's atoms uniformly so that the block's
    atom charges sum to q_blocks[b] (the charge the restraints use)"""
    ab = np.asarray(sp.atom_block)
    atoms = list(sp.atoms)
    for b, qb in enumerate(q_blocks):
        idx = np.nonzero(ab == b)[0]
        shift = (qb - sum(atoms[i][5] for i in idx)) / len(idx)
        for i in idx:
            a = atoms[i]
            atoms[i] = a[:5] + (a[5] + shift,) + a[6:]
    sp.atoms = atoms


def _charge_restrained(k, q_blocks, beta=None):
    sp = synthetic.two_site_small()
    set_block_charges(sp, q_blocks)
    sp.block_charge = [0.0] * len(q_blocks)   # ignored: the atom sums count
    sp.charge_restraints = tuple(k)
    if beta is not None:
        sp.beta = beta
    s = system.build(sp)
    s.prepare()
    return sp, s


def test_charge_restraint_background_term_keeps_ewald_energy_beta_independent():
    """The second charge restraint is the energy of the neutralising
    background of a net charge: added to the reciprocal and self energy of a
    lone ion it makes the Ewald energy independent of beta, equal to the
    Wigner self-energy -kE q^2 xi / (2 L) of a charge in a cubic box."""
    q = 1.0
    vals = []
    for beta in (0.30, 0.40):
        sp, s = _charge_restrained((0.0, 1.0, 0.0, 0.0, 0.0), [q] + [0.0] * 5, beta)
        L = sp.box[0]
        lam = orc.theta_to_lambda(s.topology(), s.runparams(), np.array(sp.theta0, float))
        e_bg = orc.forces(s.topology(), s.runparams(), system.soa(sp.pos), lam, orc.MSLDBIAS)[2][E_BIAS]
        ion = system.build(_ions(np.array([[3.0, 4.0, 5.0]]), [q], (L,) * 3, beta=beta, rcut=10.0,
                                 rswitch=9.0, grid=48, order=8))
        ion.prepare()
        e = orc.forces(ion.topology(), ion.runparams(), np.array([3.0, 4.0, 5.0]), np.ones(1))[2]
        vals.append(e[E_NBRECIP] + e[E_SELF] + e_bg)
    want = -KE * q * q * WIGNER_SC / (2 * L)
    for v in vals:
        assert abs(v - want) / abs(want) < 1e-5, (vals, want)


def test_charge_restraints_dudl_and_flat_band():
    """dU/dlambda of all three charge restraints matches differences of the
    energy in lambda; the harmonic one is zero inside |Q - q0| <= w and
    grows from the band's edge."""
    qb = [0.4, 1.0, -1.0, 0.5, -0.5, 0.25]
    sp, s = _charge_restrained((1.0, 1.0, 3.0, 0.1, 0.05), qb)
    top, rp = s.topology(), s.runparams()
    nb = top.contents.n_blocks
    assert nb == len(qb)
    x = system.soa(sp.pos)
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    _, d, e, _ = orc.forces(top, rp, x, lam, orc.MSLDBIAS)
    h = 1e-6
    for b in range(1, nb):
        lp, lm = lam.copy(), lam.copy()
        lp[b] += h
        lm[b] -= h
        fd = (orc.forces(top, rp, x, lp, orc.MSLDBIAS)[2][E_POT]
              - orc.forces(top, rp, x, lm, orc.MSLDBIAS)[2][E_POT]) / (2 * h)
        assert abs(fd - d[b]) < 1e-6 * max(1.0, abs(d[b])), (b, fd, d[b])
    # harmonic restraint alone, q0 placed relative to Q
    Q = float(np.dot(lam, qb))
    for q0, want in ((Q, 0.0), (Q - 0.04, 0.0), (Q + 0.05, 0.0), (Q - 0.35, 0.5 * 3.0 * 0.3 ** 2),
                     (Q + 0.45, 0.5 * 3.0 * 0.4 ** 2)):
        _, s3 = _charge_restrained((0.0, 0.0, 3.0, q0, 0.05), qb)
        got = orc.forces(s3.topology(), s3.runparams(), x, lam, orc.MSLDBIAS)[2][E_BIAS]
        assert abs(got - want) < 1e-12, (q0, got, want)
    # the first restraint is linear in Q: 2Q gives twice the energy
    _, s1 = _charge_restrained((1.0, 0.0, 0.0, 0.0, 0.0), qb)
    _, s2 = _charge_restrained((1.0, 0.0, 0.0, 0.0, 0.0), [2 * v for v in qb])
    e1 = orc.forces(s1.topology(), s1.runparams(), x, lam, orc.MSLDBIAS)[2][E_BIAS]
    e2 = orc.forces(s2.topology(), s2.runparams(), x, lam, orc.MSLDBIAS)[2][E_BIAS]
    assert e1 != 0.0 and abs(e2 - 2 * e1) < 1e-12 * abs(e1)


def _dihedral(x, n, L, q):
    P = [np.array([x[k * n + a] for k in range(3)]) for a in q]
    mi = lambda d: d - L * np.round(d / L)
    rij, rjk, rkl = mi(P[0] - P[1]), mi(P[1] - P[2]), mi(P[2] - P[3])
    m, nv = np.cross(rij, rjk), np.cross(rjk, rkl)
    phi = math.atan2(np.linalg.norm(np.cross(m, nv)), m @ nv)
    return -phi if rij @ nv > 0 else phi


def test_cmap_interpolates_smooth_surface_and_gradients():
    """CMAP grids sampled from the smooth surface cos(phi + 0.3) + 0.7
    sin(2 psi) + 0.4 cos(phi - 2 psi): the bicubic spline reproduces the
    surface at the terms' own (phi, psi) to 2e-3 kcal/mol (a wrong grid
    origin, axis order, sign or slope fails by far more), forces are the
    energy's gradient and dU/dlambda its lambda derivative."""
    sp = synthetic.add_cmaps(synthetic.two_site_small(), count=24, noise=0.0)
    f = lambda a, b: np.cos(a + 0.3) + 0.7 * np.sin(2 * b) + 0.4 * np.cos(a - 2 * b)
    ang = np.arange(24) * 2 * np.pi / 24 - np.pi
    P, Q = np.meshgrid(ang, ang, indexing="ij")
    for key in sp.cmap_par:
        sp.cmap_par[key] = f(P, Q)
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    t = top.contents
    assert t.n_cmaps == len(sp.cmaps) >= 20
    assert len({t.cmap_sb[3 * c] & 0xFFFF for c in range(t.n_cmaps)} - {0}) >= 2
    n, L = sp.n, np.array(sp.box)
    x = system.soa(sp.pos)
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    fo, do, e, _ = orc.forces(top, rp, x, lam, orc.BONDED)
    e0 = orc.forces(top, rp, x, lam, orc.BONDED)[2]
    want, worst = 0.0, 0.0
    for c, q in enumerate(sp.cmaps):
        ll = np.prod([lam[t.cmap_sb[3 * c + k] & 0xFFFF] if t.cmap_sb[3 * c + k] & 0xFFFF else 1.0
                      for k in range(3)])
        want += ll * f(_dihedral(x, n, L, q[:4]), _dihedral(x, n, L, q[4:]))
    assert abs(e[E_CMAP] - want) < 2e-3 * len(sp.cmaps), (e[E_CMAP], want)
    atoms = sorted({a for q in sp.cmaps[:6] for a in q})
    h = 1e-5
    for a in atoms:
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k * n + a] += h
            xm[k * n + a] -= h
            fd = (orc.forces(top, rp, xp, lam, orc.BONDED)[2][E_POT]
                  - orc.forces(top, rp, xm, lam, orc.BONDED)[2][E_POT]) / (2 * h)
            worst = max(worst, abs(fd - fo[k * n + a]))
    assert worst < 1e-5, worst
    for b in range(1, t.n_blocks):
        lp, lm = lam.copy(), lam.copy()
        lp[b] += 1e-6
        lm[b] -= 1e-6
        fd = (orc.forces(top, rp, x, lp, orc.BONDED)[2][E_POT]
              - orc.forces(top, rp, x, lm, orc.BONDED)[2][E_POT]) / 2e-6
        assert abs(fd - do[b]) < 1e-5 * max(1.0, abs(do[b])), (b, fd, do[b])
    assert e0[E_CMAP] == e[E_CMAP]


def test_dudl_matches_lambda_differences():
    name = "msld_small"
    s = built(name)
    top, rp = s.topology(), s.runparams()
    x0 = system.soa(spec(name).pos)
    lam = lambdas_of(name, s)
    _, dudl, _, _ = orc.forces(top, rp, x0, lam)
    h = 1e-6
    for b in range(1, len(lam)):
        lp, lm = lam.copy(), lam.copy()
        lp[b] += h
        lm[b] -= h
        ep = orc.forces(top, rp, x0, lp)[2][E_POT]
        em = orc.forces(top, rp, x0, lm)[2][E_POT]
        assert abs((ep - em) / (2 * h) - dudl[b]) < 1e-4 * max(1.0, abs(dudl[b]))


def test_chain_rule_matches_theta_differences():
    name = "msld_small"
    s = built(name)
    top, rp = s.topology(), s.runparams()
    th = np.array(spec(name).theta0, float) + np.array([0.0, 0.3, -0.7, 1.1])[: len(spec(name).theta0)]
    lam = orc.theta_to_lambda(top, rp, th)
    dudl = np.array([0.0, 3.0, -2.0, 5.0])[: len(lam)]
    got = orc.chain_rule(top, rp, th, lam, dudl)
    h = 1e-6
    for i in range(1, len(th)):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        dl = (orc.theta_to_lambda(top, rp, tp) - orc.theta_to_lambda(top, rp, tm)) / (2 * h)
        assert abs(got[i] - np.dot(dudl, dl)) < 1e-7


def test_theta_to_lambda_simplex():
    name = "msld_small"
    s = built(name)
    top, rp = s.topology(), s.runparams()
    rng = np.random.default_rng(2)
    for _ in range(10):
        th = rng.uniform(-3, 3, size=len(spec(name).theta0))
        lam = orc.theta_to_lambda(top, rp, th)
        assert abs(lam[1:].sum() - 1) < 1e-14
        assert np.all(lam[1:] > 0)


# ---------------------------------------------------------------------------
# switching functions
# ---------------------------------------------------------------------------
def _lj_pair(r, rcut=10.0, rswitch=8.5):
    s = synthetic.Spec(name="lj", box=(30.0, 30.0, 30.0))
    s.types.update({"AR": 39.9})
    s.nb_par.update({"AR": (-0.238, 1.88, -0.238, 1.88, 1.0, 0)})
    s.atoms = [("A", "1", "AR", "AR", "AR", 0.0, 39.9), ("A", "2", "AR", "AR", "AR", 0.0, 39.9)]
    s.pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
    s.atom_block = np.zeros(2, dtype=np.int32)
    s.shake = False
    s.rcut, s.rswitch, s.grid = rcut, rswitch, (30, 30, 30)
    b = system.build(s)
    b.prepare()
    f, _, e, _ = orc.forces(b.topology(), b.runparams(), system.soa(s.pos), np.ones(1), orc.NBDIRECT)
    return e[E_NBDIRECT], f[1]   # energy, dE/dx of atom 1


def test_vdw_force_switch_continuity():
    rc, rs = 10.0, 8.5
    eps = 1e-7
    for r0 in (rs, rc):
        e1, f1 = _lj_pair(r0 - eps, rc, rs)
        e2, f2 = _lj_pair(r0 + eps, rc, rs)
        assert abs(e1 - e2) < 1e-8
        assert abs(f1 - f2) < 1e-7
    e, f = _lj_pair(rc - 1e-9, rc, rs)
    assert abs(e) < 1e-9 and abs(f) < 1e-8
    # inside the switch the pair force is the plain Lennard-Jones force
    r = 4.0
    e, f = _lj_pair(r, rc, rs)
    eps_lj, rmin = 0.238, 2 * 1.88
    flj = 12 * eps_lj * (rmin ** 6 / r ** 7 - rmin ** 12 / r ** 13)
    assert abs(f - flj) < 1e-10 * abs(flj)


def test_erfc_fit_error_bound():
    import re
    src = open(__file__.replace("tests/test_oracle.py", "blade_amd/csrc/erfc_coef.h")).read()
    c = [float(v) for v in re.findall(r"([-+]?\d\.\d+(?:[eE][-+]?\d+)?)f", src.split("bm_erfc_c")[1])]
    p = float(re.search(r"BM_ERFC_P\s+([\d.eE+-]+)f", src).group(1))
    xmax = float(re.search(r"BM_ERFC_XMAX\s+([\d.eE+-]+)f", src).group(1))
    worst = 0.0
    for x in np.linspace(0, xmax, 2001):
        t = 1 / (1 + p * x)
        poly = sum(ck * t ** k for k, ck in enumerate(c))
        worst = max(worst, abs(math.exp(-x * x) * poly - math.erfc(x)) / math.erfc(x))
    assert worst < 2e-7


# ---------------------------------------------------------------------------
# constraints
# ---------------------------------------------------------------------------
def _constraint_pairs(top, triangles=True, branches=True):
    t = top.contents
    out = []
    if triangles:
        for (a, b, c), (d01, d02, d12) in zip(t.array("tri_idx", t.n_tri, 3), t.array("tri_b0", t.n_tri, 3)):
            out += [(a, b, d01), (a, c, d02), (b, c, d12)]
    if not branches:
        return out
    for (a, b), d in zip(t.array("br1_idx", t.n_br1, 2), t.array("br1_b0", t.n_br1)):
        out.append((a, b, d))
    for (a, b, c), (d1, d2) in zip(t.array("br2_idx", t.n_br2, 3), t.array("br2_b0", t.n_br2, 2)):
        out += [(a, b, d1), (a, c, d2)]
    for (a, b, c, e), (d1, d2, d3) in zip(t.array("br3_idx", t.n_br3, 4), t.array("br3_b0", t.n_br3, 3)):
        out += [(a, b, d1), (a, c, d2), (a, e, d3)]
    return out


def test_holonomic_position_and_velocity():
    name = "msld_small"
    s = built(name)
    top, rp = s.topology(), s.runparams()
    sp = spec(name)
    n = sp.n
    box = np.array(sp.box)
    xprev = system.soa(sp.pos)
    rng = np.random.default_rng(4)
    x = xprev + rng.normal(scale=0.02, size=xprev.shape)
    v = np.zeros_like(x)
    halfdt = 0.5 * sp.dt
    xn, vn = orc.holonomic_position(top, rp, x, xprev, v, halfdt)
    P = xn.reshape(3, n).T

    def worst_of(pairs):
        w = 0.0
        for a, b, d in pairs:
            r = P[a] - P[b]
            r -= box * np.round(r / box)
            w = max(w, abs(np.linalg.norm(r) - d) / d)
        return w
    # water triangles: analytic SETTLE, exact to rounding
    assert worst_of(_constraint_pairs(top, branches=False)) < 1e-10
    # SHAKE groups stop at |b0^2 - r^2| < 2 b0^2 tol
    tol = rp.contents.shake_tol
    assert worst_of(_constraint_pairs(top, triangles=False)) < 1.01 * tol
    pairs = _constraint_pairs(top)
    assert len(pairs) > 0
    # the velocity projection removes every bond-length rate
    vr = rng.normal(size=xn.shape)
    vp = orc.holonomic_velocity(top, rp, xn, vr).reshape(3, n).T
    worst = 0.0
    for a, b, d in pairs:
        r = P[a] - P[b]
        r -= box * np.round(r / box)
        worst = max(worst, abs(np.dot(r, vp[a] - vp[b])) / d)
    assert worst < 1e-9


def _groups(t):
    """(atoms, kind) of every constraint group, kind 0 for a triangle"""
    g = [(tuple(a), 0) for a in t.array("tri_idx", t.n_tri, 3)]
    g += [(tuple(a), 1) for a in t.array("br1_idx", t.n_br1, 2)]
    g += [(tuple(a), 2) for a in t.array("br2_idx", t.n_br2, 3)]
    g += [(tuple(a), 3) for a in t.array("br3_idx", t.n_br3, 4)]
    return g


def test_rectify_puts_groups_at_their_lengths():
    """The initial rectification fixes each group's geometry completely:
    the first atom stays, every length equals its b0, a branch atom and a
    triangle's second atom keep their directions from the first, and the
    third atom of a triangle stays in the plane of the three on its own side
    of the 0-1 line.  Atoms keep their periodic image; constraint-exact
    input comes back unchanged to rounding."""
    name = "msld_small"
    s = built(name)
    top, rp = s.topology(), s.runparams()
    t = top.contents
    sp = spec(name)
    n = sp.n
    box = np.array(sp.box)
    rng = np.random.default_rng(11)
    x0 = system.soa(sp.pos) + rng.normal(scale=0.03, size=3 * n)
    # some groups straddle the box faces
    P0 = x0.reshape(3, n).T.copy()
    groups = _groups(t)
    assert {k for _, k in groups} == {0, 1, 2, 3}
    for (a, _), k in zip(groups[::7], range(10**6)):
        P0[list(a)] += box * (0.5 - P0[a[0]] / box) + (0.02 if k % 2 else -0.02)
    x0 = np.ascontiguousarray(P0.T).reshape(-1)
    x1 = orc.rectify(top, rp, x0)
    P1 = x1.reshape(3, n).T

    def mi(r):
        return r - box * np.round(r / box)
    pairs = _constraint_pairs(top)
    worst = max(abs(np.linalg.norm(mi(P1[a] - P1[b])) - d) / d for a, b, d in pairs)
    assert worst < 1e-13, worst
    moved = np.abs(P1 - P0).max()
    assert 0.01 < moved < 0.5, moved           # same image, small moves
    grouped = set()
    for a, kind in groups:
        grouped.update(a)
        assert np.array_equal(P1[a[0]], P0[a[0]])
        d0 = [mi(P0[b] - P0[a[0]]) for b in a[1:]]
        d1 = [mi(P1[b] - P1[a[0]]) for b in a[1:]]
        keep = d0[:1] if kind == 0 else d0
        for u, w in zip(keep, d1):
            assert np.linalg.norm(np.cross(u, w)) < 1e-12 * np.linalg.norm(u) * np.linalg.norm(w) + 1e-14
            assert np.dot(u, w) > 0
        if kind == 0:
            nrm = np.cross(d0[0], d0[1])
            assert abs(np.dot(nrm, d1[1])) < 1e-12 * np.linalg.norm(nrm) * np.linalg.norm(d1[1])
            assert np.dot(np.cross(d0[0], d1[1]), nrm) > 0
    free = np.array([i for i in range(n) if i not in grouped], dtype=np.int64)
    assert np.array_equal(P1[free], P0[free])
    # exact input stays put
    x2 = orc.rectify(top, rp, x1)
    assert np.abs(x2 - x1).max() < 1e-12


def test_rectify_water_triangle_closed_form():
    """A TIP3P water rectified from a distorted start equals the exact
    geometry built by hand in the frame the rule fixes"""
    s = system.load_charmm(os.path.join(W216, "par_water_ions.prm"), os.path.join(W216, "minimized.psf"),
                           os.path.join(W216, "minimized.crd"), (18.856,) * 3)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    t = top.contents
    n = t.n_atoms
    x = s.positions().copy()
    P = x.reshape(3, n).T
    o, h1, h2 = t.array("tri_idx", t.n_tri, 3)[5]
    b01, b02, b12 = t.array("tri_b0", t.n_tri, 3)[5]
    P[h1] = P[o] + np.array([1.1, 0.05, -0.02])
    P[h2] = P[o] + np.array([-0.3, 0.8, 0.1])
    x = np.ascontiguousarray(P.T).reshape(-1)
    Q = orc.rectify(top, rp, x).reshape(3, n).T
    e0 = np.array([1.1, 0.05, -0.02]) / np.linalg.norm([1.1, 0.05, -0.02])
    e2 = np.cross([1.1, 0.05, -0.02], [-0.3, 0.8, 0.1])
    e2 /= np.linalg.norm(e2)
    e1 = np.cross(e2, e0)
    # TIP3P: O-H 0.9572, H-H 1.5139 (the HOH angle, 104.52 degrees, to 5e-4)
    assert b01 == b02 == 0.9572 and b12 == 1.5139
    ang = 2 * math.asin(b12 / (2 * b01))
    assert abs(math.degrees(ang) - 104.52) < 5e-4
    np.testing.assert_allclose(Q[h1] - Q[o], 0.9572 * e0, atol=1e-12)
    np.testing.assert_allclose(Q[h2] - Q[o], 0.9572 * (math.cos(ang) * e0 + math.sin(ang) * e1), atol=1e-12)


# ---------------------------------------------------------------------------
# neighbour list
# ---------------------------------------------------------------------------
def _water216_default_cutoff():
    # the reference's test/water216/input deck: an 18.856 A box with the
    # default 10 A cutoff, so the box edge is below twice the cutoff and
    # pairs interact through their nearest image only
    s = system.load_charmm(os.path.join(W216, "par_water_ions.prm"), os.path.join(W216, "minimized.psf"),
                           os.path.join(W216, "minimized.crd"), (18.856,) * 3, rcut=10.0, rswitch=8.5)
    s.prepare()
    return s, s.positions(), np.array((18.856,) * 3)


@pytest.mark.parametrize("name", ["msld_small", "water216_rcut10"])
def test_domdec_covers_every_pair_and_exclusion(name):
    if name == "water216_rcut10":
        s, x0, box = _water216_default_cutoff()
        n = s.n
    else:
        s = built(name)
        sp = spec(name)
        n, x0, box = sp.n, system.soa(sp.pos), np.array(sp.box)
    top, rp = s.topology(), s.runparams()
    posw, l2g, bc, tiles, masks = orc.domdec(top, rp, x0)
    slot = np.empty(n, dtype=int)
    slot[l2g[l2g >= 0]] = np.nonzero(l2g >= 0)[0]
    tile_of = {(int(i), int(j)): k for k, (i, j) in enumerate(tiles)}
    assert len(tile_of) == len(tiles)          # each directed block pair once
    P = posw.reshape(3, n).T
    rc = rp.contents.rcut
    d = P[:, None, :] - P[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1))
    t = top.contents
    excl = {tuple(sorted(p)) for p in t.array("excl_idx", t.n_excl, 2).tolist()}
    for i, j in zip(*np.nonzero(r < rc)):
        if i == j:
            continue
        si, sj = slot[i], slot[j]
        k = tile_of[(si // 32, sj // 32)]
        bit = (masks[k][si % 32] >> (sj % 32)) & 1
        assert bit == ((min(i, j), max(i, j)) in excl)


def _fd_spec(sp, terms, eidx, atoms, h=1e-5, tol=2e-6, dl_terms=None):
    """forces of `terms` against central differences of energies `eidx`, and
    (dl_terms) dU/dlambda against differences in lambda"""
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n = sp.n
    x0 = system.soa(sp.pos)
    lam = (orc.theta_to_lambda(top, rp, np.array(sp.theta0, float)) if top.contents.n_blocks > 1 else np.ones(1))
    f, d, e, _ = orc.forces(top, rp, x0, lam, terms)
    scale = max(1.0, float(np.abs(f).max()))
    worst = 0.0
    for a in atoms:
        for k in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[k * n + a] += h
            xm[k * n + a] -= h
            ep = sum(orc.forces(top, rp, xp, lam, terms)[2][i] for i in eidx)
            em = sum(orc.forces(top, rp, xm, lam, terms)[2][i] for i in eidx)
            worst = max(worst, abs((ep - em) / (2 * h) - f[k * n + a]) / scale)
    assert worst < tol, worst
    if dl_terms:
        for b in range(1, len(lam)):
            lp, lm = lam.copy(), lam.copy()
            lp[b] += 1e-6
            lm[b] -= 1e-6
            fd = (sum(orc.forces(top, rp, x0, lp, terms)[2][i] for i in eidx)
                  - sum(orc.forces(top, rp, x0, lm, terms)[2][i] for i in eidx)) / 2e-6
            assert abs(fd - d[b]) < 1e-5 * max(1.0, abs(d[b])), (b, fd, d[b])
    return e


def test_soft_core_14_gradients_and_dudl():
    """useSoftCore14 (bonded/pair.cu:124-150): 1-4 pairs of lambda-scaled
    atoms inside r_soft = 4 (1 - lambda_i lambda_j) A use the soft effective
    distance; forces and dU/dlambda are derivatives of the 1-4 energy"""
    import copy
    sp = copy.copy(spec("msld_small"))
    sp.soft_core14 = True
    s = system.build(sp)
    s.prepare()
    t = s.topology().contents
    idx = t.array("nb14_idx", t.n_nb14, 2)
    sb = t.array("nb14_sb", t.n_nb14, 2)
    lam = orc.theta_to_lambda(s.topology(), s.runparams(), np.array(sp.theta0, float))
    x = np.asarray(sp.pos)
    L = np.array(sp.box)
    inside = []
    for (i, j), (a, b) in zip(idx, sb):
        if a & 0xFFFF:
            ll = lam[a & 0xFFFF] * (lam[b & 0xFFFF] if b & 0xFFFF else 1.0)
            d = x[i] - x[j]
            d -= L * np.round(d / L)
            if np.linalg.norm(d) < 4 * (1 - ll):
                inside += [int(i), int(j)]
    assert inside, "some scaled 1-4 pairs lie inside the soft radius"
    e_soft = _fd_spec(sp, orc.NB14, [E_NB14], sorted(set(inside))[:6], dl_terms=True)
    sp.soft_core14 = False
    s2 = system.build(sp)
    s2.prepare()
    e_hard = orc.forces(s2.topology(), s2.runparams(), system.soa(sp.pos), lam, orc.NB14)[2]
    assert e_soft[E_NB14] != e_hard[E_NB14]


def test_term_variety_gradients():
    """Urey-Bradley, periodic impropers, NBFIX pairs and geometric
    combination: forces are energy gradients for the bonded, 1-4 and
    nonbonded terms"""
    from blade_amd import synthetic
    sp = synthetic.term_variety(spec("msld_small"))
    atoms = _polymer_atoms("msld_small")
    e = _fd_spec(sp, orc.BONDED, [E_BOND, E_ANGLE, E_UREY, E_DIHE, E_IMPR], atoms)
    assert e[E_UREY] != 0.0 and e[E_IMPR] != 0.0
    _fd_spec(sp, orc.NB14, [E_NB14], atoms)
    _fd_spec(sp, orc.NBDIRECT, [E_NBDIRECT], atoms[:3] + [0, 1])


def test_chain_rule_with_a_fixed_block():
    """a fixed block holds lambda = theta; the others share 1 - theta_fixed
    by the softmax (msld.cu:880-898) and the chain rule follows
    theta_to_lambda's derivative (msld.cu:955-971)"""
    import copy
    sp = copy.copy(spec("msld_small"))
    sp.block_fixed = [0, 0, 1, 0]
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    th = np.array([0.0, 0.4, 0.25, -0.6])
    lam = orc.theta_to_lambda(top, rp, th)
    assert lam[2] == 0.25 and abs(lam[1:].sum() - 1) < 1e-14
    dudl = np.array([0.0, 3.0, -2.0, 5.0])
    got = orc.chain_rule(top, rp, th, lam, dudl)
    assert got[2] == 0.0
    for i in (1, 3):
        tp, tm = th.copy(), th.copy()
        tp[i] += 1e-6
        tm[i] -= 1e-6
        dl = (orc.theta_to_lambda(top, rp, tp) - orc.theta_to_lambda(top, rp, tm)) / 2e-6
        assert abs(got[i] - np.dot(dudl, dl)) < 1e-7


E_NOE, E_HARM, E_MMFP, E_RESD = 14, 15, 16, 17


def test_restraint_forces_are_energy_gradients_and_closed_forms():
    """NOE, harmonic, bond/angle/dihedral restraint and RESD terms
    (restrain/restrain.cu): forces are gradients of their energies and
    dU/dlambda of the scaled ones their lambda derivatives; the NOE energy is
    zero inside its band and continuous where the switched tail starts; the
    harmonic energy is k |x - x0|^n"""
    from blade_amd import synthetic
    sp = synthetic.add_restraints(spec("msld_small"))
    atoms = sorted({a for rows in sp.restraints.values() for row in rows
                    for a in (row[:4] if len(row) >= 8 and not isinstance(row[-1], bool) else row[:2])
                    if isinstance(a, (int, np.integer))})
    e = _fd_spec(sp, orc.RESTRAIN, [E_NOE, E_HARM, E_MMFP, E_RESD], atoms, dl_terms=True)
    assert all(e[k] != 0.0 for k in (E_NOE, E_HARM, E_MMFP, E_RESD))
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    x = system.soa(sp.pos)
    n = sp.n
    # harmonic: k |x - x0|^n per term
    want = 0.0
    for (i, k, x0, y0, z0, pw) in sp.restraints["harm"]:
        d = np.array([x[i], x[n + i], x[2 * n + i]]) - [x0, y0, z0]
        want += k * np.linalg.norm(d) ** pw
    only = copy_restraints(sp, "harm")
    s2 = system.build(only)
    s2.prepare()
    eh = orc.forces(s2.topology(), s2.runparams(), x, lam, orc.RESTRAIN)[2][E_HARM]
    assert abs(eh - want) < 1e-10 * max(1.0, want)
    # NOE: zero inside the band, continuous where the power-law tail starts
    i, j = sp.restraints["noe"][0][:2]
    L = np.array(sp.box)
    d0 = np.array([x[i] - x[j], x[n + i] - x[n + j], x[2 * n + i] - x[2 * n + j]])
    d0 -= L * np.round(d0 / L)
    r0 = np.linalg.norm(d0)

    def noe_energy(r, rmin, rmax, rsw):
        one = copy_restraints(sp, None)
        one.restraints = {"noe": [(i, j, rmin, 5.0, rmax, 3.0, 0.8, rsw, 1.0, 0.0, 0.0, 0.0, False)]}
        b = system.build(one)
        b.prepare()
        return orc.forces(b.topology(), b.runparams(), x, lam, orc.RESTRAIN)[2][E_NOE]
    assert noe_energy(r0, r0 - 0.5, r0 + 0.5, 0.0) == 0.0
    rmax = r0 - 1.0
    below = noe_energy(r0, 0.0, rmax, 1.0 - 1e-7)
    above = noe_energy(r0, 0.0, rmax, 1.0 + 1e-7)
    assert abs(below - above) < 1e-5 and abs(below - 0.5 * 3.0 * 1.0) < 1e-5


def copy_restraints(sp, keep):
    import copy
    c = copy.copy(sp)
    c.restraints = {keep: sp.restraints[keep]} if keep else {}
    return c


def test_exclusion_correction_at_and_near_coincidence():
    """The PME exclusion correction e = -kqq erf(beta r)/r takes its Taylor
    series about r = 0 below beta r = 0.01 (msldEwaldType 3 corrects pairs
    between alternative substituents, which may coincide).  Pinned by the
    limit at r = 0 (-kqq 2 beta/sqrt(pi); the force factor kqq beta^3
    4/(3 sqrt(pi))), agreement with math.erf on both sides of the switch,
    and the force factor as the central difference of the energy."""
    kqq, beta = 332.0716 * (-0.3) * 0.45, 0.34
    fr0, e0 = orc.nbex_pair(kqq, beta, 0.0)
    assert e0 == pytest.approx(-kqq * 2 * beta / math.sqrt(math.pi), rel=1e-15)
    assert fr0 == pytest.approx(kqq * beta ** 3 * 4 / (3 * math.sqrt(math.pi)), rel=1e-15)
    rs = 0.01 / beta
    for r in (0.2 * rs, 0.9 * rs, 0.999 * rs, 1.001 * rs, 1.5 * rs, 10 * rs, 3.0):
        fr, e = orc.nbex_pair(kqq, beta, r)
        # closed forms in Python's libm (erf, exp); near r = 0 they lose
        # about eps/(beta r)^2 to cancellation, so the bar follows that
        ec = -kqq * math.erf(beta * r) / r
        frc = kqq * (math.erf(beta * r) / r - 2 * beta / math.sqrt(math.pi) * math.exp(-(beta * r) ** 2)) / r ** 2
        x = beta * r
        assert e == pytest.approx(ec, rel=1e-14)
        assert fr == pytest.approx(frc, rel=max(1e-12, 4e-16 / x ** 2))
        h = 1e-4 * max(r, rs)
        dedr = (orc.nbex_pair(kqq, beta, r + h)[1] - orc.nbex_pair(kqq, beta, r - h)[1]) / (2 * h)
        assert fr * r == pytest.approx(dedr, rel=1e-6, abs=1e-9 * abs(kqq))


# ---------------------------------------------------------------------------
# MSLD soft bonds (msld/msld.cu:478-495, bonded/bonded.cu:81-107, 239-256,
# 408-425)
# ---------------------------------------------------------------------------
def _soft_spec(params=(0.8, 2.0, 1.5), removescaling=False):
    import copy
    sp = copy.deepcopy(spec("msld_small"))
    ab = np.array(sp.atom_block)
    sp.soft_bonds = [(i, j) for (i, j) in sp.bonds if (ab[i] > 0) != (ab[j] > 0)]
    sp.soft_bond_params = params
    # the generator builds bonds at their rest lengths: displace the atoms
    # so the soft factor differs from 1
    sp.pos = np.asarray(sp.pos) + np.random.default_rng(3).normal(scale=0.08, size=np.shape(sp.pos))
    if removescaling:
        sp.scale_terms = (0, 1, 0, 1, 1, 1)   # bond and angle scaling removed
    return sp


def test_soft_bond_terms_gradients_and_dudl():
    """soft bond, Urey-Bradley, angle, dihedral and improper terms: forces
    and dU/dlambda are derivatives of their energies (radius 0.8 A, power 2
    for bonds, 1.5 for the other terms, removescaling bond angle)"""
    sp = _soft_spec(removescaling=True)
    assert len(sp.soft_bonds) == 3
    atoms = sorted({a for b in sp.soft_bonds for a in b})
    _fd_spec(sp, orc.BONDED, [E_BOND, E_ANGLE, E_UREY, E_DIHE, E_IMPR], atoms, dl_terms=True)


def test_soft_terms_topology_and_removescaling():
    """a term holding both atoms of a soft bond is soft and keeps its
    blocks' scaling where removescaling names its type; other terms of the
    same block lose it"""
    sp = _soft_spec(removescaling=True)
    s = system.build(sp)
    s.prepare()
    t = s.topology().contents
    pairs = {tuple(sorted(b)) for b in sp.soft_bonds}
    bi = t.array("bond_idx", t.n_bonds, 2)
    soft = t.array("bond_soft", t.n_bonds)
    sb = t.array("bond_sb", t.n_bonds, 2)
    ai = t.array("angle_idx", t.n_angles, 3)
    asoft = t.array("angle_soft", t.n_angles)
    asb = t.array("angle_sb", t.n_angles, 2)
    for k in range(t.n_bonds12):
        holds = tuple(sorted(bi[k])) in pairs
        assert soft[k] == holds
        if holds:
            assert sb[k][0] & 0xFFFF > 0          # still scaled by its substituent block
    ab = np.array(sp.atom_block)
    plain_sub = [k for k in range(t.n_bonds12) if not soft[k] and ab[bi[k][0]] > 0 and ab[bi[k][1]] > 0]
    assert plain_sub and all(sb[k][0] == 0 for k in plain_sub)   # removescaling applies
    for k in range(t.n_angles):
        holds = any(set(p) <= set(ai[k]) for p in pairs)
        assert asoft[k] == holds
        if holds:
            assert asb[k][0] & 0xFFFF > 0
    assert asoft.sum() > 0 and t.array("dihe_soft", t.n_dihes).sum() > 0


def test_soft_terms_limits():
    """at lambda = 1 every soft term equals the plain term; with the
    non-bond power 1 a soft angle, dihedral or improper equals the plain
    lambda-scaled term at any lambda, while the soft bonds differ"""
    import copy
    sp = _soft_spec(params=(0.8, 2.0, 1.5))
    plain = copy.deepcopy(sp)
    plain.soft_bonds = []
    ones = np.ones(sp.nblocks)
    e = {}
    for name, q in (("soft", sp), ("plain", plain)):
        s = system.build(q)
        s.prepare()
        e[name] = orc.forces(s.topology(), s.runparams(), system.soa(q.pos), ones, orc.BONDED)[2]
    for k in (E_BOND, E_ANGLE, E_UREY, E_DIHE, E_IMPR):
        assert abs(e["soft"][k] - e["plain"][k]) < 1e-9 * max(1.0, abs(e["plain"][k])), k
    # non-bond power 1: angles, dihedrals and impropers equal the plain ones
    sp1 = _soft_spec(params=(0.8, 2.0, 1.0))
    sp1.pos = sp.pos
    plain1 = copy.deepcopy(sp1)
    plain1.soft_bonds = []
    lam = None
    for name, q in (("soft", sp1), ("plain", plain1)):
        s = system.build(q)
        s.prepare()
        if lam is None:
            lam = orc.theta_to_lambda(s.topology(), s.runparams(), np.array(q.theta0, float))
            assert 0.05 < lam[1:].min() and lam[1:].max() < 0.95
        e[name] = orc.forces(s.topology(), s.runparams(), system.soa(q.pos), lam, orc.BONDED)[2]
    for k in (E_ANGLE, E_DIHE, E_IMPR):
        assert abs(e["soft"][k] - e["plain"][k]) < 1e-9 * max(1.0, abs(e["plain"][k])), k
    assert abs(e["soft"][E_BOND] - e["plain"][E_BOND]) > 1e-6   # the bonds do differ
