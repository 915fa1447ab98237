"""Virtual sites (holonomic/virtual.cu, system/structure.cxx:734-800 and
1059-1085, system/potential.cxx:947-995).

The oracle's placement is pinned by closed forms (a TIP4P M site on the HOH
bisector 0.15 A from the oxygen; a two-host site on its hosts' line at
dist + scale |e| from its first host; a three-host site at its distance,
angle and dihedral).  Its force spreading is pinned by central differences
of the energy with the sites re-placed at every displaced host position: the
spread forces on the hosts are the negative gradient of E(place(x)).  The
GPU tests compare the device against that oracle on a 216-water TIP4P box
with extra two- and three-host sites.
"""
import copy
import math
import os

import numpy as np
import pytest

from blade_amd import synthetic, system
from oracle import oracle as orc
from tests.helpers import rel_rms

E_POT = 20
DEGREES = 0.017453292519943   # radians per degree as the reference defines it (main/defines.h:32)


def _box_spec(nwat, box, rcut, rswitch, grid, beta, colinear, offaxis, seed=0x7193):
    sp = synthetic.tip4p_box(nwat, (box,) * 3, seed, colinear=colinear, offaxis=offaxis)
    sp.rcut, sp.rswitch, sp.beta = rcut, rswitch, beta
    sp.grid, sp.order = (grid,) * 3, 6
    return sp


def _small():
    return _box_spec(64, 12.42, 6.0, 5.0, 12, 0.5, colinear=3, offaxis=3)


def _prepared(sp):
    s = system.build(sp)
    s.prepare()
    return s


def _xyz(x, n):
    return x.reshape(3, n).T


def _mi(d, box):
    return d - box * np.round(d / box)


def test_topology_and_exclusions():
    """every site is in the topology with its hosts and parameters (angles
    in radians); a site is excluded from its first host and from that
    host's bond partners, and from nothing else"""
    sp = _small()
    s = _prepared(sp)
    t = s.topology().contents
    assert t.n_virt2 == 3 and t.n_virt3 == 64 + 3
    v2 = t.array("virt2_idx", t.n_virt2, 3)
    p2 = t.array("virt2_par", t.n_virt2, 2)
    v3 = t.array("virt3_idx", t.n_virt3, 4)
    p3 = t.array("virt3_par", t.n_virt3, 3)
    for row, want in zip(v2, sp.virt2):
        assert list(row) == list(want[:3])
    assert np.allclose(p2, [[0.2, 0.1]] * 3)
    for row, par, want in zip(v3, p3, sp.virt3):
        assert list(row) == list(want[:4])
        assert np.array_equal(par, [want[4], DEGREES * want[5], DEGREES * want[6]])
    ex = {tuple(p) for p in t.array("excl_idx", t.n_excl, 2)}
    bonded = {}
    for a, b in sp.bonds:
        bonded.setdefault(a, set()).add(b)
        bonded.setdefault(b, set()).add(a)
    for v, h0 in [(r[0], r[1]) for r in sp.virt2] + [(r[0], r[1]) for r in sp.virt3]:
        want = {h0} | bonded.get(h0, set())
        got = {b if a == v else a for (a, b) in ex if v in (a, b)}
        assert got == want, (v, want, got)


def test_placement_closed_forms():
    """M on the bisector 0.15 A from O; the two-host site beyond H1 on the
    O-H1 line at 0.2 + 0.1 |H1 - O|; the off-plane site 0.5 A from O at 110
    degrees from O-H1 and dihedral 75 degrees (site, O, H1, H2); every site
    in the periodic image nearest its previous position"""
    sp = _small()
    s = _prepared(sp)
    top, rp = s.topology(), s.runparams()
    n, box = sp.n, np.array(sp.box)
    x0 = system.soa(sp.pos)
    x = _xyz(orc.virtual_place(top, rp, x0), n)
    for (v, o, h1, h2, dist, th, ph) in sp.virt3:
        r = _mi(x[v] - x[o], box)
        a1, a2 = _mi(x[h1] - x[o], box), _mi(x[h2] - x[o], box)
        assert abs(np.linalg.norm(r) - abs(dist)) < 1e-12
        if dist < 0:
            bis = (a1 + a2) / np.linalg.norm(a1 + a2)
            assert abs(np.dot(r, bis) / np.linalg.norm(r) - 1) < 1e-12
        else:
            c = np.dot(r, a1) / (np.linalg.norm(r) * np.linalg.norm(a1))
            assert abs(math.degrees(math.acos(c)) - th) < 1e-9
            # dihedral site-O-H1-H2
            b0, b1, b2 = -r, a1, a2 - a1
            nn1, nn2 = np.cross(b0, b1), np.cross(b1, b2)
            m1 = np.cross(nn1, b1 / np.linalg.norm(b1))
            dih = math.degrees(math.atan2(np.dot(m1, nn2), np.dot(nn1, nn2)))
            assert abs(abs(dih) - ph) < 1e-9, dih
    for (v, h0, h1, dist, scale) in sp.virt2:
        e = _mi(x[h0] - x[h1], box)
        r = _mi(x[v] - x[h0], box)
        assert np.allclose(r, (dist / np.linalg.norm(e) + scale) * e, rtol=0, atol=1e-13)
        assert abs(np.linalg.norm(r) - (dist + scale * np.linalg.norm(e))) < 1e-12
    # nearest image of the previous position: a site moved by a box edge
    # comes back one box edge away from its host
    v = sp.virt3[0][0]
    x1 = x0.copy()
    x1[v] += box[0]
    y = _xyz(orc.virtual_place(top, rp, x1), n)
    assert abs(y[v, 0] - x[v, 0] - box[0]) < 1e-12
    assert np.allclose(y[v, 1:], x[v, 1:], rtol=0, atol=0)


def test_spread_forces_are_the_gradient():
    """the force array (the energy gradient dU/dx, as the reference keeps
    it) that the oracle reports on host atoms, with the sites' entries moved
    onto them as J^T F, is the central-difference gradient of the energy
    with every site re-placed from the displaced hosts; the sites' own
    entries are zero"""
    sp = _small()
    s = _prepared(sp)
    top, rp = s.topology(), s.runparams()
    n = sp.n
    lam = np.ones(1)
    x = orc.virtual_place(top, rp, system.soa(sp.pos))
    f, _, e, _ = orc.forces(top, rp, x, lam)
    sites = [r[0] for r in sp.virt2] + [r[0] for r in sp.virt3]
    F = _xyz(f, n)
    assert np.all(F[sites] == 0)
    # hosts of a colinear site, an off-plane site and a plain TIP4P water
    atoms = list(sp.virt2[0][1:3]) + list(sp.virt3[64 + 1][1:4]) + list(sp.virt3[10][1:4])
    h = 1e-5
    worst = 0.0
    for a in atoms:
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[k * n + a] += h
            xm[k * n + a] -= h
            ep = orc.forces(top, rp, orc.virtual_place(top, rp, xp), lam)[2][E_POT]
            em = orc.forces(top, rp, orc.virtual_place(top, rp, xm), lam)[2][E_POT]
            fd = (ep - em) / (2 * h)
            err = abs(fd - f[k * n + a])
            worst = max(worst, err)
            assert err <= 2e-5 + 1e-6 * abs(f[k * n + a]), (a, k, fd, f[k * n + a])
    print("\nworst |FD - spread force| %.2e kcal/mol/A" % worst)


def _fd(top, rp, x, n, a, k, place=(), h=1e-5):
    """central difference of the energy in coordinate k of atom a, with the
    sites in `place` re-placed from the displaced hosts and every other
    site held where it is"""
    out = []
    for sgn in (1, -1):
        y = x.copy()
        y[k * n + a] += sgn * h
        if place:
            z = orc.virtual_place(top, rp, y)
            for v in place:
                y[v::n] = z[v::n]
        out.append(orc.forces(top, rp, y, np.ones(1))[2][E_POT])
    return (out[0] - out[1]) / (2 * h)


def test_gradient_check_rejects_plausible_spreadings():
    """the gradient check above separates the exact J^T F from two
    plausible mistakes (every share measured by central differences, one
    site re-placed at a time): a spreading that gives a site's entry F to
    its first host only, and, for a two-host site, the reference's own
    two-host rule (virtual.cu:155-158: (s - dist (e.F)/|e|^3) F on the
    second host, along F instead of along e), which the exact J^T F
    replaces"""
    sp = _small()
    s = _prepared(sp)
    top, rp = s.topology(), s.runparams()
    n, box = sp.n, np.array(sp.box)
    x = orc.virtual_place(top, rp, system.soa(sp.pos))
    X = _xyz(x, n)
    # an off-plane site: first-host-only spreading misses the rest of J^T F
    v, o = next((r[0], r[1]) for r in sp.virt3 if r[4] > 0)
    G = np.array([_fd(top, rp, x, n, v, k) for k in range(3)])   # the site's own entry
    gap = 0.0
    for k in range(3):
        moved = _fd(top, rp, x, n, o, k, place=(v,))
        held = _fd(top, rp, x, n, o, k)
        gap = max(gap, abs(held + G[k] - moved))
    assert gap > 1e-2, gap
    # a two-host site (hosts h0 = H1, h1 = O): exact J^T F against the
    # reference's rule, on the second host
    v, h0, h1, dist, scale = sp.virt2[0]
    e = _mi(X[h0] - X[h1], box)
    G = np.array([_fd(top, rp, x, n, v, k) for k in range(3)])
    r3 = np.linalg.norm(e) ** 3
    sc = dist / np.linalg.norm(e) + scale
    ref_share = -(sc - dist / r3 * np.dot(e, G)) * G
    exact_share = -(sc * G - dist / r3 * np.dot(e, G) * e)
    worst_exact = worst_ref = 0.0
    for k in range(3):
        moved = _fd(top, rp, x, n, h1, k, place=(v,))
        held = _fd(top, rp, x, n, h1, k)
        worst_exact = max(worst_exact, abs(held + exact_share[k] - moved))
        worst_ref = max(worst_ref, abs(held + ref_share[k] - moved))
    print("\nfirst-host-only spreading off the gradient by %.3e, the reference two-host rule by %.3e, "
          "the exact rule by %.1e kcal/mol/A" % (gap, worst_ref, worst_exact))
    assert worst_exact < 1e-5
    assert worst_ref > 1e-2


def _reference_three_host_rule(X, box, v, h, dist, theta, G):
    """the reference's spreading of a three-host site's entry G onto its
    hosts (holonomic/virtual.cu:160-225), restated in numpy: the angular
    part (s x (r x G))/|s|^2 and, off the axis, the torsional part along
    (s x t)"""
    hx = [X[j].copy() for j in h]
    r = _mi(X[v] - hx[0], box)
    t = _mi(hx[1] - hx[2], box)
    if dist < 0:
        t = 0.5 * t
        hx[1] = hx[1] - t
    s = _mi(hx[0] - hx[1], box)
    sinv2 = 1 / np.dot(s, s)
    hf = [G.copy(), np.zeros(3), np.zeros(3)]
    tf = sinv2 * np.cross(s, np.cross(r, G))
    hf[1] += tf
    hf[0] -= tf
    if math.sin(theta) > 1e-6:
        p, q = np.cross(r, s), np.cross(s, t)
        rr, tt = sinv2 * np.cross(s, p), sinv2 * np.cross(q, s)
        pf = np.dot(p, G) / np.linalg.norm(p)
        pfr = (np.linalg.norm(rr) * pf) / (np.linalg.norm(tt) * np.linalg.norm(q)) * q
        st = np.dot(s, t) * sinv2
        hf[2] += pfr
        hf[0] += st * pfr
        hf[1] += (-1 - st) * pfr
    if dist < 0:
        hf[1] *= 0.5
        hf[2] += hf[1]
    return hf


def test_three_host_spreading_matches_the_reference_rule():
    """for three-host sites the reference's geometric rule (an independent
    derivation) is the gradient as well: with one site re-placed, each
    host's central difference equals its held-site difference plus the
    reference rule's share, for a bisector M and an off-plane site"""
    sp = _small()
    s = _prepared(sp)
    top, rp = s.topology(), s.runparams()
    n, box = sp.n, np.array(sp.box)
    x = orc.virtual_place(top, rp, system.soa(sp.pos))
    X = _xyz(x, n)
    worst = 0.0
    for row in (sp.virt3[10], next(r for r in sp.virt3 if r[4] > 0)):
        v, h0, h1, h2, dist, th, ph = row
        G = np.array([_fd(top, rp, x, n, v, k) for k in range(3)])
        hf = _reference_three_host_rule(X, box, v, (h0, h1, h2), dist, DEGREES * th, G)
        for j, a in enumerate((h0, h1, h2)):
            for k in range(3):
                moved = _fd(top, rp, x, n, a, k, place=(v,))
                held = _fd(top, rp, x, n, a, k)
                worst = max(worst, abs(held + hf[j][k] - moved))
    print("\nreference three-host rule against the gradient: %.1e kcal/mol/A" % worst)
    assert worst < 1e-5


_TIP4P_PSF = """PSF EXT XPLOR

         1 !NTITLE
* one TIP4P water

         4 !NATOM
         1 W        1        TIP4     OH2      OT4      0.000000       15.9994           0
         2 W        1        TIP4     H1       HT       0.520000        1.0080           0
         3 W        1        TIP4     H2       HT       0.520000        1.0080           0
         4 W        1        TIP4     OM       LP      -1.040000        0.0000           0

         3 !NBOND: bonds
         1         2         1         3         2         3

         1 !NTHETA: angles
         2         1         3

         0 !NPHI: dihedrals


         0 !NIMPHI: impropers


         0 !NDON: donors


         0 !NACC: acceptors


         0 !NNB

         0         0         0         0

         1         0 !NGRP NST2
         0         0         0

         1         4 !NUMLP NUMLPH
         3         1   F     -0.150000      0.000000      0.000000
         4         1         2         3

"""


def test_psf_lone_pair_section(tmp_path):
    """blade_read_psf takes the NUMLP/NUMLPH section (structure.cxx:734-800):
    a bisector lone pair with three hosts becomes a three-host site"""
    p = tmp_path / "tip4p.psf"
    p.write_text(_TIP4P_PSF)
    s = system.BladeSystem((0,))
    s.init_parameters()
    for t, m in synthetic.TIP4P_TYPES.items():
        s.add_parameter_atoms(t, m)
    s.add_parameter_bonds("OT4", "HT", 450.0, 0.9572)
    s.add_parameter_bonds("HT", "HT", 0.0, 1.5139)
    s.add_parameter_angles("HT", "OT4", "HT", 55.0, 104.52, 0.0, 0.0)
    for t, par in synthetic.TIP4P_NB.items():
        s.add_parameter_nbonds(t, *par)
    s.add_parameter_nbonds("HT", *synthetic.WATER_NB["HT"])
    s.init_structure()
    s.read_psf(str(p))
    s.add_shake(1)
    s.init_coordinates(4)
    f0, f1, f2 = synthetic._water_frame()
    for k, q in enumerate((f0, f1, f2, np.zeros(3))):
        s.add_coordinates_position(k + 1, *(q + 5.0))
    s.add_coordinates_box(0, 20.0, 20.0, 20.0)
    s.init_msld(0)
    s.init_run()
    s.add_run_flags(1.0 / synthetic.PICOSECOND, 0.34, 9.0, 8.0, 0, 1, 1.0, -1, -1, -1, 6, 2e-7)
    s.add_run_dynopts(0, 0, 0, 0.002 * synthetic.PICOSECOND, 300.0)
    s.n, s.nblocks = 4, 1
    s.prepare()
    t = s.topology().contents
    assert (t.n_virt2, t.n_virt3) == (0, 1)
    assert list(t.array("virt3_idx", 1, 4)[0]) == [3, 0, 1, 2]
    assert np.allclose(t.array("virt3_par", 1, 3)[0], [-0.15, 0.0, 0.0])
    assert t.n_tri == 1
    x0 = np.concatenate([q + 5.0 for q in (f0, f1, f2, np.zeros(3))]).reshape(4, 3).T.reshape(-1)
    x = _xyz(orc.virtual_place(s.topology(), s.runparams(), x0), 4)
    assert abs(np.linalg.norm(x[3] - x[0]) - 0.15) < 1e-12


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _gpu_spec():
    sp = _box_spec(216, 18.856, 9.0, 8.0, 20, 0.34, colinear=6, offaxis=6)
    sp.temperature = 298.0
    return sp


@pytest.mark.gpu
def test_gpu_forces_energy_and_sites_match_oracle():
    """216 TIP4P waters plus 6 two-host and 6 off-plane sites: the device
    places every site as the oracle does, its forces (with the sites'
    forces on their hosts and zero on the sites) are within 1e-5 relative
    RMS of the oracle's, its energy within 1e-6; a second evaluation is bit
    identical"""
    sp = _gpu_spec()
    s = system.build(sp)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    n, box = sp.n, np.array(sp.box)
    xg = s.positions()
    xo = orc.virtual_place(top, rp, xg)
    dpos = np.abs(_mi(_xyz(xg, n) - _xyz(xo, n), box)).max()
    fo, _, eo, _ = orc.forces(top, rp, xo, np.ones(1))
    fg, eg = s.forces(), s.energy()
    sites = [r[0] for r in sp.virt2] + [r[0] for r in sp.virt3]
    frms = rel_rms(fg, fo)
    erel = abs(eg[E_POT] - eo[E_POT]) / abs(eo[E_POT])
    print("\n[tip4p + sites] N=%d sites=%d |x_gpu - x_oracle| %.2e A force relRMS %.3e energy rel %.3e"
          % (n, len(sites), dpos, frms, erel), flush=True)
    assert dpos < 1e-12
    assert np.all(_xyz(fg, n)[sites] == 0)
    assert frms <= 1e-5
    assert erel <= 1e-6
    s.run_energy()
    assert np.array_equal(s.forces(), fg)
    s.close()


@pytest.mark.gpu
def test_gpu_dynamics_tracks_oracle_and_paths_agree():
    """5 steps against the FP64 oracle from the same state and Philox
    stream (positions within 1e-4 A modulo the box); 25 steps through the
    per-term seam (bm_domdec_update, bm_calc_force, bm_update), eager
    bm_dynamics and graph replay end bit-identical; after 200 steps the
    SETTLE lengths hold and the sites sit at their construction; a batch
    refuses a member with sites"""
    import ctypes as C
    sp = _gpu_spec()
    s = system.build(sp)
    s.prepare()
    top, rp = s.topology(), s.runparams()
    n, box = sp.n, np.array(sp.box)
    x = np.concatenate([orc.rectify(top, rp, system.soa(sp.pos)), np.zeros(1)])
    v = np.concatenate([system.soa(sp.vel), np.zeros(1)])
    lam = np.ones(1)
    x[:3 * n] = orc.domdec(top, rp, x[:3 * n])[0]
    for k in range(5):
        orc.step(top, rp, k, k % 10 == 0, x, v, lam)
    s.run_dynamics(5, freq_nrg=10)
    d = np.abs(_mi(_xyz(s.positions(), n) - _xyz(x[:3 * n], n), box)).max()
    print("\n[tip4p + sites] max |x_gpu - x_oracle| after 5 steps: %.3e A" % d, flush=True)
    assert d < 1e-4
    s.close()

    out = []
    for mode in ("terms", "eager", "graph"):
        g = system.build(sp)
        g.run_energy()
        L, dev = g.L, g.device()
        if mode == "terms":
            for k in range(1, 26):
                assert L.bm_domdec_update(dev, int(k % 10 == 0)) == 0
                assert L.bm_calc_force(dev, k, int(k % 7 == 0)) == 0
                assert L.bm_update(dev, k, int(k % 7 == 0)) == 0
        else:
            assert L.bm_dynamics(dev, 1, 25, 7, int(mode == "graph")) == 0
        assert L.bm_sync(dev) == 0
        out.append(system.state(g, sp.n, sp.nblocks))
        if mode == "graph":
            g.run_dynamics(200, freq_nrg=50)
            # a step places the sites before its forces, so after the
            # update they trail their hosts by one step's motion; an energy
            # evaluation places them again
            g.run_energy()
            xg = _xyz(g.positions(), n)
            assert np.all(np.isfinite(xg))
            worst = 0.0
            for (vs, o, h1, h2, dist, th, ph) in sp.virt3:
                worst = max(worst, abs(np.linalg.norm(_mi(xg[vs] - xg[o], box)) - abs(dist)))
            for (o, h1, h2) in [(r[1], r[2], r[3]) for r in sp.virt3[:216]]:
                worst = max(worst, abs(np.linalg.norm(_mi(xg[h1] - xg[o], box)) - 0.9572))
            print("   after 200 more steps: worst site/bond length error %.2e A" % worst, flush=True)
            assert worst < 1e-8
            b = C.c_void_p()
            arr = (C.c_void_p * 1)(dev)
            assert L.bm_batch_create(1, arr, C.byref(b)) != 0
        g.close()
    for q in range(4):
        assert np.array_equal(out[0][q], out[1][q]), q
        assert np.array_equal(out[0][q], out[2][q]), q
