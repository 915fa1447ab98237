"""Batch workload generator (CPU): members of uneven size with one lambda
site each, as the batched benchmark builds them.

  * no two interacting ligand atoms closer than 1.7 A (interacting: more
    than two bonds apart and not alternative substituents of the site);
  * every substituent neutral with positive hydrogens, the system neutral;
  * the member set is reproducible from its seed.
"""
import numpy as np

from blade_amd import synthetic


def _graph(sp):
    nbr = [set() for _ in range(sp.n)]
    for a, b in sp.bonds:
        nbr[a].add(b)
        nbr[b].add(a)
    return nbr


def test_member_set_reproducible():
    assert synthetic.batch_set(32) == synthetic.batch_set(32)
    m = synthetic.batch_set(32)
    assert len(m) == 32 and all(3000 <= n <= 25000 and 2 <= s <= 6 for n, s, _ in m)


def test_small_members_clash_free_and_neutral():
    for natoms, nsubs, seed in ((3300, 6, 11), (4100, 4, 12), (3600, 2, 13)):
        sp = synthetic.batch_member(seed, natoms, nsubs)
        lig = [i for i, a in enumerate(sp.atoms) if a[0] in ("POLY", "LIG")]
        nbr = _graph(sp)
        block = np.asarray(sp.atom_block)
        site = np.asarray(sp.block_site)
        P = np.asarray(sp.pos, float).reshape(-1, 3)
        L = np.array(sp.box)
        worst = np.inf
        for a in lig:
            near = set(nbr[a])
            for b in list(nbr[a]):
                near |= nbr[b]
            for b in lig:
                if b <= a or b in near:
                    continue
                if block[a] != block[b] and site[block[a]] == site[block[b]]:
                    continue
                d = P[a] - P[b]
                d -= L * np.round(d / L)
                worst = min(worst, float(np.linalg.norm(d)))
        assert worst >= 1.7, (seed, worst)
        q = np.array([a[5] for a in sp.atoms])
        assert abs(q.sum()) < 1e-9
        for s in range(1, nsubs + 1):
            sub = [a for a in sp.atoms if a[2] == "SUB%d" % s]
            assert len(sub) == 8
            assert abs(sum(a[5] for a in sub)) < 1e-9
            assert all(a[5] > 0 for a in sub if a[4].startswith("H"))
