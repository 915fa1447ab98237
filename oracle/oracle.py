"""ctypes bindings of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker or the reported CPU
baseline, never as a product code path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "liboracle.so")

NBDIRECT, EWALDSELF, PME, BONDED, NB14, NBEX, MSLDBIAS, RESTRAIN = (1 << k for k in range(8))
ALL = 0xFF

_L = None
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(PATH)
        vp = C.c_void_p
        L.orc_forces.restype = C.c_long
        L.orc_forces.argtypes = [vp, vp, _f64p, _f64p, C.c_uint, _f64p, _f64p, _f64p]
        L.orc_ewald_recip_direct.restype = C.c_double
        L.orc_ewald_recip_direct.argtypes = [vp, vp, _f64p, _f64p, C.c_int, _f64p, _f64p]
        L.orc_ewald_real_images.restype = C.c_double
        L.orc_ewald_real_images.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_double]
        L.orc_nbex_pair.restype = None
        L.orc_nbex_pair.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_bspline.restype = C.c_double
        L.orc_bspline.argtypes = [C.c_int, C.c_double]
        L.orc_bspline_moduli.restype = None
        L.orc_bspline_moduli.argtypes = [C.c_int, C.c_int, _f64p]
        L.orc_fft3d.restype = None
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_int]
        L.orc_domdec_columns.restype = None
        L.orc_domdec_columns.argtypes = [vp, C.c_int, _i32p, _i32p]
        L.orc_domdec.restype = C.c_long
        L.orc_domdec.argtypes = [vp, vp, _f64p, C.c_int, C.c_int, C.c_int, C.c_long, _i32p, _i32p, _i32p,
                                 _i32p, _i32p, C.POINTER(C.c_uint32)]
        L.orc_theta_to_lambda.restype = None
        L.orc_theta_to_lambda.argtypes = [vp, vp, _f64p, _f64p]
        L.orc_chain_rule.restype = None
        L.orc_chain_rule.argtypes = [vp, vp, _f64p, _f64p, _f64p, _f64p]
        L.orc_theta_bias.restype = None
        L.orc_theta_bias.argtypes = [vp, vp, _f64p, _f64p, _f64p]
        L.orc_holonomic_position.restype = None
        L.orc_holonomic_position.argtypes = [vp, vp, _f64p, _f64p, _f64p, C.c_double]
        L.orc_holonomic_velocity.restype = None
        L.orc_holonomic_velocity.argtypes = [vp, vp, _f64p, _f64p]
        L.orc_rectify.argtypes = [vp, vp, _f64p]
        L.orc_virtual_place.restype = None
        L.orc_virtual_place.argtypes = [vp, vp, _f64p]
        L.orc_philox4x32.restype = None
        L.orc_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_normal.restype = C.c_double
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [vp, vp, C.c_long, C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.orc_threads.restype = C.c_int
        _L = L
    return _L


def _p(a):
    return a.ctypes.data_as(_f64p)


def _addr(p):
    return C.cast(p, C.c_void_p)


def forces(top, rp, pos, lam, terms=ALL):
    """(force[3N], dudl[nL], energy[23], pairs) at SoA positions `pos`."""
    t = top.contents
    n, nl = t.n_atoms, max(1, t.n_blocks)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    f = np.zeros(3 * n)
    d = np.zeros(nl)
    e = np.zeros(23)
    pairs = lib().orc_forces(_addr(top), _addr(rp), _p(pos), _p(lam), terms, _p(f), _p(d), _p(e))
    return f, d, e, pairs


def ewald_recip_direct(top, rp, pos, lam, kmax):
    t = top.contents
    f = np.zeros(3 * t.n_atoms)
    d = np.zeros(max(1, t.n_blocks))
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    e = lib().orc_ewald_recip_direct(_addr(top), _addr(rp), _p(pos), _p(lam), kmax, _p(f), _p(d))
    return e, f, d


def ewald_real_images(q, pos, box, beta, rcut):
    q = np.ascontiguousarray(q, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    return lib().orc_ewald_real_images(len(q), _p(q), _p(pos), _p(box), beta, rcut)


def bspline(order, u):
    return lib().orc_bspline(order, u)


def bspline_moduli(order, K):
    out = np.zeros(K)
    lib().orc_bspline_moduli(order, K, _p(out))
    return out


def fft3d(grid, sign=-1):
    """complex 3D FFT of a (K0,K1,K2) complex128 array (unnormalised)."""
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    lib().orc_fft3d(g.shape[0], g.shape[1], g.shape[2], g.ctypes.data_as(_f64p), sign)
    return g


def domdec_columns(rp, n):
    a, b = C.c_int(), C.c_int()
    lib().orc_domdec_columns(_addr(rp), n, C.byref(a), C.byref(b))
    return a.value, b.value


def domdec(top, rp, pos):
    """list semantics: (wrapped pos, l2g[32*nb], block_count[nb], directed tiles (nt,2),
    masks (nt,32))"""
    t = top.contents
    n = t.n_atoms
    ncx, ncy = domdec_columns(rp, n)
    maxb = n // 32 + 1 + ncx * ncy
    maxt = maxb * maxb
    pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
    l2g = np.zeros(32 * maxb, dtype=np.int32)
    bc = np.zeros(maxb, dtype=np.int32)
    nb = C.c_int()
    ti = np.zeros(maxt, dtype=np.int32)
    tj = np.zeros(maxt, dtype=np.int32)
    tm = np.zeros(32 * maxt, dtype=np.uint32)
    nt = lib().orc_domdec(_addr(top), _addr(rp), _p(pos), ncx, ncy, maxb, maxt,
                          l2g.ctypes.data_as(_i32p), bc.ctypes.data_as(_i32p), C.byref(nb),
                          ti.ctypes.data_as(_i32p), tj.ctypes.data_as(_i32p),
                          tm.ctypes.data_as(C.POINTER(C.c_uint32)))
    if nt < 0:
        raise RuntimeError("orc_domdec capacity")
    nbv = nb.value
    return (pos, l2g[:32 * nbv], bc[:nbv], np.stack([ti[:nt], tj[:nt]], 1), tm[:32 * nt].reshape(nt, 32))


def theta_to_lambda(top, rp, theta):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    lam = np.zeros_like(theta)
    lib().orc_theta_to_lambda(_addr(top), _addr(rp), _p(theta), _p(lam))
    return lam


def chain_rule(top, rp, theta, lam, dudl):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    dudl = np.ascontiguousarray(dudl, dtype=np.float64)
    out = np.zeros_like(theta)
    lib().orc_chain_rule(_addr(top), _addr(rp), _p(theta), _p(lam), _p(dudl), _p(out))
    return out


def theta_bias(top, rp, theta):
    """(dU/dtheta[nL], energy[23]) of the theta biases at theta"""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.zeros_like(theta)
    e = np.zeros(23)
    lib().orc_theta_bias(_addr(top), _addr(rp), _p(theta), _p(out), _p(e))
    return out, e


def holonomic_position(top, rp, x, xprev, v, halfdt):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    xprev = np.ascontiguousarray(xprev, dtype=np.float64)
    lib().orc_holonomic_position(_addr(top), _addr(rp), _p(x), _p(xprev), _p(v), halfdt)
    return x, v


def holonomic_velocity(top, rp, x, v):
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    lib().orc_holonomic_velocity(_addr(top), _addr(rp), _p(x), _p(v))
    return v


def rectify(top, rp, x):
    """positions (SoA, 3N) with every constraint group at exactly its lengths"""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().orc_rectify(_addr(top), _addr(rp), _p(x))
    return x


def virtual_place(top, rp, x):
    """positions (SoA, 3N) with every virtual site placed from its hosts"""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    lib().orc_virtual_place(_addr(top), _addr(rp), _p(x))
    return x


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32(c, k, o)
    return list(o)


def normal(seed, step, dof):
    return lib().orc_normal(seed, step, dof)


def step(top, rp, stepno, energy_step, x, v, lam):
    """one FP64 VRORV step in place on x[3N+nL], v[3N+nL], lam[nL]; returns energy[23]"""
    e = np.zeros(23)
    rc = lib().orc_step(_addr(top), _addr(rp), stepno, int(energy_step), _p(x), _p(v), _p(lam), _p(e))
    if rc:
        raise RuntimeError("oracle step produced NaN")
    return e


def threads():
    return lib().orc_threads()


def nbex_pair(kqq, beta, r):
    """(fr, e) of one PME exclusion correction at distance r"""
    fr, e = C.c_double(), C.c_double()
    lib().orc_nbex_pair(kqq, beta, r, C.byref(fr), C.byref(e))
    return fr.value, e.value
