"""The binary checkpoint format (CPU): a file written in the documented
layout (magic "BLADEMCK", int64 step, int32 atoms, int32 blocks, f64 box,
then f64 SoA positions, velocities, theta, theta velocities) is read back
into the system's coordinates bit for bit; a wrong magic or a mismatched
atom count is refused."""
import numpy as np
import pytest

from blade_amd import system
from tests.helpers import spec


def _write(path, step, pos, vel, theta, theta_v, box, n=None, magic=b"BLADEMCK"):
    n = len(pos[0]) if n is None else n
    with open(path, "wb") as f:
        f.write(magic)
        f.write(np.int64(step).tobytes())
        f.write(np.array([n, len(theta)], dtype=np.int32).tobytes())
        f.write(np.array(box, dtype=np.float64).tobytes())
        for a in (pos, vel):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.array(theta, dtype=np.float64).tobytes())
        f.write(np.array(theta_v, dtype=np.float64).tobytes())


def test_checkpoint_read_restores_exact_coordinates(tmp_path):
    sp = spec("msld_small")
    n = sp.n
    rng = np.random.default_rng(3)
    pos = np.array(sp.box)[:, None] * rng.random((3, n))
    vel = rng.normal(size=(3, n))
    th = rng.normal(size=sp.nblocks)
    _write(tmp_path / "ck.bin", 120, pos, vel, th, np.zeros(sp.nblocks), sp.box)
    s = system.build(sp)
    s.read_checkpoint(tmp_path / "ck.bin")
    x = np.zeros(3 * n)
    assert s.L.blade_get_positions(s.h, x.ctypes.data_as(system.C.POINTER(system.C.c_double))) == 0
    assert np.array_equal(x, pos.reshape(-1))


def test_checkpoint_rejects_wrong_files(tmp_path):
    sp = spec("msld_small")
    n = sp.n
    pos = np.zeros((3, n))
    _write(tmp_path / "bad.bin", 0, pos, pos, np.zeros(sp.nblocks), np.zeros(sp.nblocks), sp.box, magic=b"NOTACKPT")
    s = system.build(sp)
    with pytest.raises(RuntimeError):
        s.read_checkpoint(tmp_path / "bad.bin")
    pos2 = np.zeros((3, n - 1))
    _write(tmp_path / "short.bin", 0, pos2, pos2, np.zeros(sp.nblocks), np.zeros(sp.nblocks), sp.box)
    s2 = system.build(sp)
    with pytest.raises(RuntimeError):
        s2.read_checkpoint(tmp_path / "short.bin")
