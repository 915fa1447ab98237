"""bm_batch_exchange_noalloc, the replica exchange that allocates nothing
(its swap staging is a static device array), called through the C-ABI
against bm_batch_exchange on identical batches: the exchanged states,
replica labels, acceptances and dW are bit-identical.  Two calls in a row,
the second on a smaller system (the array keeps the first call's
contents), each for both exchange parities."""
import copy
import ctypes as C

import numpy as np
import pytest

from blade_amd import synthetic, system
from blade_amd._ffi import check, ptr
from tests.helpers import relaxed

pytestmark = pytest.mark.gpu


def _replicas(name):
    base = relaxed(name)
    specs = []
    for k in range(3):
        sp = copy.copy(base)
        sp.fixed_bias = [0.0] + [0.7 * (k + 1) * (-1) ** j for j in range(1, base.nblocks)]
        sp.vel = synthetic._maxwell(np.random.default_rng(90 + k), [a[6] for a in sp.atoms], sp.temperature)
        sp.theta0 = [t if j == 0 else t + 0.4 * k * (j % 2 * 2 - 1) for j, t in enumerate(sp.theta0)]
        sp.theta_v = None
        specs.append(sp)
    return specs


def _batch(specs):
    systems = [system.build(sp) for sp in specs]
    for s in systems:
        s.run_energy()
    b = system.Batch(systems)
    b.dynamics(1, 5, 0)
    b.sync()
    return b, systems


def _exchange(b, fn, step, parity, seed, replica):
    n = len(b.systems)
    rep = np.ascontiguousarray(replica, dtype=np.int32).copy()
    acc = np.zeros(n, dtype=np.int32)
    dw = np.zeros(n)
    i32 = C.POINTER(C.c_int32)
    check(fn(b.h, int(step), int(parity), int(seed), rep.ctypes.data_as(i32), acc.ctypes.data_as(i32), ptr(dw)),
          "exchange")
    return rep, acc, dw


def test_noalloc_exchange_equals_exchange():
    for name in ("two_site_small", "msld_small"):          # the second call on a smaller system
        specs = _replicas(name)
        n, nb = specs[0].n, specs[0].nblocks
        ba, sa = _batch(specs)
        bb, sb = _batch(specs)
        L = ba.L
        rep_a, rep_b = [0, 1, 2], [0, 1, 2]
        for parity in (0, 1):
            ra, aa, wa = _exchange(ba, L.bm_batch_exchange, 5, parity, 4321, rep_a)
            rb, ab, wb = _exchange(bb, L.bm_batch_exchange_noalloc, 5, parity, 4321, rep_b)
            print("\n[%s parity %d] accepted %s dW %s" % (name, parity, aa, wa), flush=True)
            assert np.array_equal(ra, rb) and np.array_equal(aa, ab) and np.array_equal(wa, wb)
            for x, y in zip(sa, sb):
                qa, qb = system.state(x, n, nb), system.state(y, n, nb)
                for q in range(4):
                    assert np.array_equal(qa[q], qb[q])
            rep_a, rep_b = ra, rb
        ba.close()
        bb.close()
        for s in sa + sb:
            s.close()
