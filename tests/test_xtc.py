"""XTC trajectory frames (CPU).  The writer (blade_amd/csrc/xtc.cpp) is
pinned byte for byte against the reference's vendored xdrfile library
(src/xdr/xdrfile.cxx, xdrfile_xtc.cxx: third-party, compiled where it lies
into oracle/_ref/ by oracle/Makefile), which the reference's io/io.cxx:345
calls to write its frames; the frames are then read back with that
library's reader."""
import ctypes as C
import os

import numpy as np
import pytest

from blade_amd import _ffi, synthetic

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "oracle", "_ref", "libxdrfile.so")
W216 = os.path.join(HERE, "golden", "water216", "minimized.crd")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference xdrfile library not built here")


def _ref():
    L = C.CDLL(REF)
    L.xdrfile_open.restype = C.c_void_p
    L.xdrfile_open.argtypes = [C.c_char_p, C.c_char_p]
    L.xdrfile_close.argtypes = [C.c_void_p]
    f32p = C.POINTER(C.c_float)
    L.write_xtc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, f32p, f32p, C.c_float]
    L.read_xtc_natoms.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
    L.read_xtc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), f32p, f32p, f32p, f32p]
    return L


def _mine():
    L = _ffi.lib()
    f32p = C.POINTER(C.c_float)
    L.blade_write_xtc_frame.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_float, f32p, f32p, C.c_float]
    L.blade_write_xtc_frame.restype = C.c_int
    return L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _write_both(tmp, frames, prec=1000.0):
    ref, mine = _ref(), _mine()
    a, b = str(tmp / "ref.xtc"), str(tmp / "mine.xtc")
    fp = ref.xdrfile_open(a.encode(), b"w")
    for step, time, box, x in frames:
        box, x = _f(box), _f(x)
        assert ref.write_xtc(fp, x.shape[0], step, time, _p(box), _p(x), prec) == 0
        rc = mine.blade_write_xtc_frame(b.encode(), int(step != frames[0][0]), x.shape[0], step, time, _p(box),
                                        _p(x), prec)
        assert rc == 0
    ref.xdrfile_close(fp)
    return open(a, "rb").read(), open(b, "rb").read(), b


def _read_all(path, natoms):
    ref = _ref()
    fp = ref.xdrfile_open(path.encode(), b"r")
    out = []
    while True:
        step, time, prec = C.c_int(), C.c_float(), C.c_float()
        box = np.zeros(9, np.float32)
        x = np.zeros((natoms, 3), np.float32)
        if ref.read_xtc(fp, natoms, C.byref(step), C.byref(time), _p(box), _p(x), C.byref(prec)) != 0:
            break
        out.append((step.value, time.value, box, x))
    ref.xdrfile_close(fp)
    return out


def _water216_nm():
    rows = [l.split() for l in open(W216).read().split("\n") if l.strip() and not l.startswith("*")][1:]
    return np.array([[float(r[4]), float(r[5]), float(r[6])] for r in rows]) / 10.0


@pytest.mark.parametrize("case", ["water216", "msld_small", "gas", "nine_atoms", "fine_precision"])
def test_frames_byte_identical_to_reference_writer(tmp_path, case):
    rng = np.random.default_rng(5)
    prec = 1000.0
    if case == "water216":
        x = _water216_nm()
        box = np.diag([1.8856] * 3)
    elif case == "msld_small":
        sp = synthetic.msld_small()
        x = np.array(sp.pos) / 10.0
        box = np.diag(np.array(sp.box) / 10.0)
    elif case == "gas":          # sparse atoms: large first size class, no water runs
        x = rng.uniform(0, 30.0, size=(400, 3))
        box = np.diag([30.0] * 3)
    elif case == "nine_atoms":   # stored as plain floats
        x = rng.uniform(0, 2.0, size=(9, 3))
        box = np.diag([2.0] * 3)
    else:                        # extents beyond 2^24 integers: per-axis bit fields
        x = _water216_nm()
        box = np.diag([1.8856] * 3)
        prec = 1e7
    frames = [(0, 0.0, box, x), (1000, 2.0, box, x + rng.normal(scale=0.01, size=x.shape)),
              (2000, 4.0, box, x[::-1].copy())]
    a, b, path = _write_both(tmp_path, frames, prec)
    assert len(a) > 0 and a == b
    got = _read_all(path, x.shape[0])
    assert len(got) == 3
    for (step, time, bx, xx), (gs, gt, gb, gx) in zip(frames, got):
        assert gs == step and abs(gt - time) < 1e-6
        np.testing.assert_allclose(gb.reshape(3, 3), bx, rtol=1e-6)
        tol = 0.5 / prec + 1e-6 * np.abs(xx).max() if x.shape[0] > 9 else 1e-6
        assert np.abs(gx - xx).max() <= tol
