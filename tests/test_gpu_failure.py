"""Failure handling on the device (run on an MI355X).  Every condition is
forced from the host, so every kernel still exits normally:

  list overflow: the tile list and the PME brick bins start far below what
  the system needs (bm_set_list_capacity).  The device halts the state at
  the step that overflowed; bm_dynamics grows the buffer that overflowed
  (x1.5 per retry, as domdec/cull.cu:337-375) and runs the halted steps
  again, and the trajectory is bit-identical to one with ample capacity.

  a non-finite position (set from the host, as the NaN check of
  update/update.cu:19-24 would see it): the state halts at once, the call
  returns BM_ENAN, the flag names the atom and no other coordinate moves.
"""
import ctypes as C

import numpy as np
import pytest

from blade_amd import system
from blade_amd._ffi import check
from tests.helpers import relaxed

pytestmark = pytest.mark.gpu

BM_ENAN = -3


def _caps(s):
    mt, pc = C.c_longlong(), C.c_int()
    check(s.L.bm_get_list_capacity(s.device(), C.byref(mt), C.byref(pc)), "bm_get_list_capacity")
    return mt.value, pc.value


def _tiles(s):
    nb, nt, ne = C.c_int(), C.c_int(), C.c_int()
    check(s.L.bm_get_domdec_sizes(s.device(), C.byref(nb), C.byref(nt), C.byref(ne)), "sizes")
    return nt.value


def test_list_overflow_regrows_and_resumes_bit_identical():
    sp = relaxed("msld_small")
    a = system.build(sp)
    a.run_energy()
    nt = _tiles(a)
    mt0, pc0 = _caps(a)
    a.run_dynamics(40, freq_nrg=10)
    ref = system.state(a, sp.n, sp.nblocks)
    a.close()

    b = system.build(sp)
    b.run_energy()
    small_t, small_p = nt // 3, 8
    check(b.L.bm_set_list_capacity(b.device(), small_t, small_p), "bm_set_list_capacity")
    assert _caps(b) == (small_t, small_p)
    b.run_dynamics(40, freq_nrg=10)
    got = system.state(b, sp.n, sp.nblocks)
    mt1, pc1 = _caps(b)
    print("\ntiles %d (capacity %d): grown %d -> %d; brick records %d: grown %d -> %d"
          % (nt, mt0, small_t, mt1, pc0, small_p, pc1), flush=True)
    assert mt1 > small_t and pc1 > small_p
    for q in range(4):
        assert np.array_equal(got[q], ref[q]), q
    b.close()


def test_nonfinite_position_halts_and_names_the_atom():
    sp = relaxed("msld_small")
    s = system.build(sp)
    s.run_energy()
    x, v, th, tv = system.state(s, sp.n, sp.nblocks)
    bad = 37
    x2 = x.copy()
    x2[bad, 1] = np.nan
    s.set_positions(np.ascontiguousarray(x2.T).reshape(-1))
    rc = s.L.blade_run_dynamics(s.h, 20, 0, 1)
    assert rc == BM_ENAN
    flag = C.c_int(-1)
    check(s.L.bm_get_nan_flag(s.device(), C.byref(flag)), "bm_get_nan_flag")
    assert flag.value == bad
    x3, v3, th3, tv3 = system.state(s, sp.n, sp.nblocks)
    keep = np.ones(sp.n, bool)
    keep[bad] = False
    assert np.array_equal(x3[keep], x[keep])
    assert np.array_equal(v3, v)
    assert np.array_equal(th3, th) and np.array_equal(tv3, tv)
    assert np.isnan(x3[bad, 1])
    s.close()
