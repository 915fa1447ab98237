"""Topology derivation (CPU): the exclusion, 1-4 and PME exclusion-correction
lists and the 1-4 parameters derived by the library (blade_prepare) against
an independent restatement of the reference's rules:

  src/system/potential.cxx:970-1072  exclusions grown along bonds: bonded
      pairs, then partners of bonded partners, then one more bond; a pair is
      added only if its atoms interact (src/msld/msld.cu:675-678: same block,
      or blocks of different sites) and it is not already excluded; every
      pair of atoms in different blocks of one site is excluded as well
  src/system/potential.cxx:1093-1142 1-4 pairs = the third-bond exclusions,
      eps14 = sqrt(eps14_i eps14_j), Rmin14 = Rmin14/2_i + Rmin14/2_j,
      c12 = eps Rmin^12, c6 = 2 eps Rmin^6, e14fac = sqrt(e_i) sqrt(e_j)
  src/system/potential.cxx:1144-1163 + src/msld/msld.cu:623-668 (mode 2)
      PME corrections: every exclusion that is not a 1-4 pair, except pairs
      of different blocks of one site
"""
import numpy as np
import pytest

from tests.helpers import built, spec


def _derive(sp):
    n = sp.n
    block = np.asarray(sp.atom_block)
    site = np.asarray(sp.block_site)

    def interacting(i, j):
        return block[i] == block[j] or site[block[i]] != site[block[j]]

    excluded = set()
    levels = []
    # level 1: bonded pairs
    lvl = set()
    for a, b in sp.bonds:
        p = (min(a, b), max(a, b))
        if p not in excluded and interacting(a, b):
            lvl.add(p)
            excluded.add(p)
    levels.append(lvl)
    # levels 2 and 3: extend a pair of the previous level by one bond
    for _ in range(2):
        prev = levels[-1]
        part = [set() for _ in range(n)]
        for i, j in prev:
            part[i].add(j)
            part[j].add(i)
        lvl = set()
        for a, b in sp.bonds:
            for x, y in ((a, b), (b, a)):
                for k in part[y]:
                    if k == x:
                        continue
                    p = (min(x, k), max(x, k))
                    if p not in excluded and interacting(x, k):
                        lvl.add(p)
                        excluded.add(p)
        levels.append(lvl)
    msld = set()
    for i in range(n):
        for j in range(i + 1, n):
            if not interacting(i, j):
                msld.add((i, j))
    allx = excluded | msld
    nb14 = levels[2]
    nbex = {p for p in allx - nb14
            if not (block[p[0]] != block[p[1]] and site[block[p[0]]] == site[block[p[1]]])}
    return allx, nb14, nbex


@pytest.mark.parametrize("name", ["config1", "msld_small", "config2_small", "two_site_small"])
def test_exclusion_lists_follow_the_reference_rules(name):
    sp = spec(name)
    s = built(name)                     # owns the topology tables
    t = s.topology().contents
    allx, nb14, nbex = _derive(sp)
    got_x = {tuple(sorted(p)) for p in t.array("excl_idx", t.n_excl, 2).tolist()}
    got_14 = {tuple(sorted(p)) for p in t.array("nb14_idx", t.n_nb14, 2).tolist()}
    got_ex = {tuple(sorted(p)) for p in t.array("nbex_idx", t.n_nbex, 2).tolist()}
    assert got_x == allx
    assert got_14 == nb14
    assert got_ex == nbex
    assert len(got_14) == t.n_nb14 and len(got_ex) == t.n_nbex


@pytest.mark.parametrize("name", ["msld_small", "two_site_small"])
def test_bonded_scaling_follows_the_reference_rule(name):
    """src/msld/msld.cu:507-553 without soft bonds: a term carries the
    distinct blocks of its atoms in descending order as (site << 16) | block
    codes, padded with 0.  For a type named by `removescaling` an atom's
    block counts only when every atom of the term in that block is
    restrained.  msld_small scales every type; two_site_small only
    dihedrals (and CMAP), with each substituent's first atom restrained."""
    sp = spec(name)
    s = built(name)
    t = s.topology().contents
    block = np.asarray(sp.atom_block)
    site = np.asarray(sp.block_site)
    restrained = {a for g in sp.atom_restraints for a in g}
    seen_scaled = seen_unscaled = seen_kept = 0
    for key, count, width, typ in (("bond", t.n_bonds12, 2, 0), ("angle", t.n_angles, 3, 2),
                                   ("dihe", t.n_dihes, 4, 3), ("impr", t.n_imprs, 4, 4)):
        idx = t.array(key + "_idx", count, width)
        sb = t.array(key + "_sb", count, 2)
        for atoms, got in zip(idx, sb):
            if sp.scale_terms[typ]:
                kept = {int(block[a]) for a in atoms if block[a]}
            else:
                kept = {int(block[a]) for a in atoms if block[a] and
                        all(b in restrained for b in atoms if block[b] == block[a])}
                seen_kept += bool(kept)
            bl = sorted(kept, reverse=True)
            want = [(int(site[b]) << 16) | b for b in bl] + [0] * (2 - len(bl))
            seen_scaled += bool(bl)
            seen_unscaled += bool(not bl and any(block[a] for a in atoms))
            assert list(got) == want, (key, atoms, got, want)
    assert seen_scaled > 0
    if name == "two_site_small":
        assert seen_unscaled > 0 and seen_kept > 0


def test_nb14_parameters():
    name = "config2_small"
    sp = spec(name)
    s = built(name)
    t = s.topology().contents
    idx = t.array("nb14_idx", t.n_nb14, 2)
    qq, c12, c6, e14 = (t.array(k, t.n_nb14) for k in ("nb14_qq", "nb14_c12", "nb14_c6", "nb14_e14fac"))
    q = t.array("charge", t.n_atoms)
    for k, (i, j) in enumerate(idx):
        ti, tj = sp.atoms[i][4], sp.atoms[j][4]
        eps_i, _, e14_i, r14_i, f_i, _ = sp.nb_par[ti]
        eps_j, _, e14_j, r14_j, f_j, _ = sp.nb_par[tj]
        eps = np.sqrt(e14_i * e14_j)     # the API negates the input eps (both negative)
        rmin = r14_i + r14_j
        s6 = rmin ** 6
        assert abs(c12[k] - eps * s6 * s6) <= 1e-12 * abs(eps * s6 * s6)
        assert abs(c6[k] - 2 * eps * s6) <= 1e-12 * abs(2 * eps * s6)
        assert abs(e14[k] - np.sqrt(f_i) * np.sqrt(f_j)) < 1e-15
        assert abs(qq[k] - q[i] * q[j]) < 1e-15


def _variety(name):
    from blade_amd import synthetic, system
    sp = synthetic.term_variety(spec(name))
    s = system.build(sp)
    s.prepare()
    return sp, s


def test_vdw_table_nbfix_and_geometric_combination():
    """src/system/potential.cxx:1359-1407: a type pair with an NBFIX entry
    (either order) takes its eps and Rmin as given; otherwise eps =
    sqrt(eps_i eps_j) and Rmin = Rmin/2_i + Rmin/2_j, or 2 sqrt(Rmin/2_i
    Rmin/2_j) when both types select the geometric rule; c12 = eps Rmin^12,
    c6 = 2 eps Rmin^6.  1-4 pairs use the eps14/Rmin14 columns the same way."""
    sp, s = _variety("msld_small")
    t = s.topology().contents
    T = t.n_types
    tidx = t.array("type_idx", t.n_atoms)
    c12 = t.array("vdw_c12", T * T)
    c6 = t.array("vdw_c6", T * T)
    first = {}
    for i, a in enumerate(sp.atoms):
        first.setdefault(a[4], i)

    def want(ta, tb, is14):
        fx = sp.nbfix_par.get((ta, tb), sp.nbfix_par.get((tb, ta)))
        if fx is not None:
            eps, rmin = (-fx[2], fx[3]) if is14 else (-fx[0], fx[1])
        else:
            A, B = sp.nb_par[ta], sp.nb_par[tb]
            ea, ra = (-A[2], A[3]) if is14 else (-A[0], A[1])
            eb, rb = (-B[2], B[3]) if is14 else (-B[0], B[1])
            eps = np.sqrt(ea * eb)
            rmin = 2 * np.sqrt(ra * rb) if (A[5] == 1 and B[5] == 1) else ra + rb
        return eps * rmin ** 12, 2 * eps * rmin ** 6

    seen_fix = seen_geo = 0
    for ta, ia in first.items():
        for tb, ib in first.items():
            k = tidx[ia] * T + tidx[ib]
            w12, w6 = want(ta, tb, False)
            assert abs(c12[k] - w12) <= 1e-12 * abs(w12) and abs(c6[k] - w6) <= 1e-12 * abs(w6), (ta, tb)
            seen_fix += ((ta, tb) in sp.nbfix_par)
            seen_geo += (sp.nb_par[ta][5] == 1 and sp.nb_par[tb][5] == 1)
    assert seen_fix > 0 and seen_geo > 0
    idx = t.array("nb14_idx", t.n_nb14, 2)
    p12, p6 = t.array("nb14_c12", t.n_nb14), t.array("nb14_c6", t.n_nb14)
    for k, (i, j) in enumerate(idx):
        w12, w6 = want(sp.atoms[i][4], sp.atoms[j][4], True)
        assert abs(p12[k] - w12) <= 1e-12 * abs(w12) and abs(p6[k] - w6) <= 1e-12 * abs(w6)


def test_urey_bradley_terms_follow_the_angles():
    """src/system/potential.cxx:597-645: every angle whose parameter has a
    Urey-Bradley constant adds a 1-3 bond term after the 1-2 bonds, with
    kb = 2 kub (the setter doubles it) and b0 = s0"""
    sp, s = _variety("msld_small")
    t = s.topology().contents
    nub = t.n_bonds - t.n_bonds12
    idx = t.array("bond_idx", t.n_bonds, 2)[t.n_bonds12:]
    kb = t.array("bond_kb", t.n_bonds)[t.n_bonds12:]
    b0 = t.array("bond_b0", t.n_bonds)[t.n_bonds12:]
    want = []
    for a, b, c in sp.angles:
        key = tuple(sp.atoms[x][4] for x in (a, b, c))
        par = sp.angle_par.get(key, sp.angle_par.get(key[::-1]))
        if par[2] != 0:
            want.append((a, c, 2 * par[2], par[3]))
    assert nub == len(want) > 0
    for (i, j), k, r, w in zip(idx, kb, b0, want):
        assert (i, j) == w[:2] and abs(k - w[2]) < 1e-12 and abs(r - w[3]) < 1e-12


@pytest.mark.parametrize("mode", [1, 3])
def test_pme_exclusion_corrections_ewald_types(mode):
    """src/msld/msld.cu:623-668: mode 1 scales a same-block correction by
    that block's lambda once and drops pairs of different blocks of one
    site; mode 3 keeps every non-1-4 exclusion, scaled by both atoms'
    lambdas (siteBlock codes in descending block order)"""
    import copy
    from blade_amd import system
    sp = copy.copy(spec("two_site_small"))
    sp.ewald_type = mode
    s = system.build(sp)
    s.prepare()
    t = s.topology().contents
    allx, nb14, _ = _derive(sp)
    block = np.asarray(sp.atom_block)
    site = np.asarray(sp.block_site)
    want = {}
    for i, j in allx - nb14:
        b = sorted((int(block[i]), int(block[j])), reverse=True)
        if mode == 1:
            if b[0] == b[1]:
                b[1] = 0
            if b[0] != b[1] and site[b[0]] == site[b[1]]:
                continue
        want[(i, j)] = [(int(site[x]) << 16) | x for x in b]
    got = {}
    for (i, j), sb in zip(t.array("nbex_idx", t.n_nbex, 2), t.array("nbex_sb", t.n_nbex, 2)):
        got[(min(i, j), max(i, j))] = list(sb)
    assert got == want
    same = [p for p in want if block[p[0]] == block[p[1]] and block[p[0]]]
    assert same, "the system has same-block corrections"
