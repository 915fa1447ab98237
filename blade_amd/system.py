"""Python mirror of the reference-compatible blade_* API.

`BladeSystem` wraps one `System*` of libblade_mi355x.so.  Its methods carry
the reference's names without the `blade_` prefix (init_parameters,
add_parameter_bonds, add_atom, add_msld_flags, add_run_flags, run_energy,
...), with the same argument meaning, 1-based indices and AKMA units, so a
test reads like the reference's CHARMM-side calls.  `build(spec)` drives
those calls from a synthetic `Spec`.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import check, ptr

ATMOSPHERE = 0.000014584319218586   # kcal/mol/A^3 (main/defines.h:58)


def _b(s):
    return s.encode() if isinstance(s, str) else s


class BladeSystem:
    def __init__(self, gpus=(0,)):
        self.L = _ffi.lib()
        arr = (C.c_int * len(gpus))(*gpus)
        self.h = self.L.blade_init_system(len(gpus), arr)
        self.n = 0
        self.nblocks = 1

    def close(self):
        if self.h:
            self.L.blade_dest_system(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference-compatible calls (include/blade.h) -----------------------
    def init_parameters(self):
        self.L.blade_init_parameters(self.h)

    def add_parameter_atoms(self, name, mass):
        self.L.blade_add_parameter_atoms(self.h, _b(name), mass)

    def add_parameter_bonds(self, t1, t2, kb, b0):
        self.L.blade_add_parameter_bonds(self.h, _b(t1), _b(t2), kb, b0)

    def add_parameter_angles(self, t1, t2, t3, k, t0, kub, s0):
        self.L.blade_add_parameter_angles(self.h, _b(t1), _b(t2), _b(t3), k, t0, kub, s0)

    def add_parameter_dihes(self, t1, t2, t3, t4, k, n, d0):
        self.L.blade_add_parameter_dihes(self.h, _b(t1), _b(t2), _b(t3), _b(t4), k, n, d0)

    def add_parameter_imprs(self, t1, t2, t3, t4, k, n, d0):
        self.L.blade_add_parameter_imprs(self.h, _b(t1), _b(t2), _b(t3), _b(t4), k, n, d0)

    def add_parameter_nbonds(self, t, eps, sig, eps14, sig14, e14fac, combine):
        self.L.blade_add_parameter_nbonds(self.h, _b(t), eps, sig, eps14, sig14, e14fac, combine)

    def add_parameter_nbfixs(self, t1, t2, eps, sig, eps14, sig14):
        self.L.blade_add_parameter_nbfixs(self.h, _b(t1), _b(t2), eps, sig, eps14, sig14)

    def init_structure(self):
        self.L.blade_init_structure(self.h)

    def add_atom(self, idx, seg, resid, resname, name, typ, q, m):
        self.L.blade_add_atom(self.h, idx, _b(seg), _b(resid), _b(resname), _b(name), _b(typ), q, m)
        self.n = max(self.n, idx)

    def add_bond(self, i, j):
        self.L.blade_add_bond(self.h, i, j)

    def add_angle(self, i, j, k):
        self.L.blade_add_angle(self.h, i, j, k)

    def add_dihe(self, i, j, k, l):
        self.L.blade_add_dihe(self.h, i, j, k, l)

    def add_impr(self, i, j, k, l):
        self.L.blade_add_impr(self.h, i, j, k, l)

    def add_shake(self, on):
        self.L.blade_add_shake(self.h, int(on))

    def init_coordinates(self, n):
        self.L.blade_init_coordinates(self.h, n)

    def add_coordinates_position(self, i, x, y, z):
        self.L.blade_add_coordinates_position(self.h, i, x, y, z)

    def add_coordinates_velocity(self, i, x, y, z):
        self.L.blade_add_coordinates_velocity(self.h, i, x, y, z)

    def add_coordinates_box(self, name, a, b, c, al=90.0, be=90.0, ga=90.0):
        self.L.blade_add_coordinates_box(self.h, name, a, b, c, al, be, ga)

    def init_msld(self, nblocks):
        self.L.blade_init_msld(self.h, nblocks)
        self.nblocks = nblocks

    def add_msld_atomassignment(self, atom, block):
        self.L.blade_add_msld_atomassignment(self.h, atom, block)

    def add_msld_initialconditions(self, block, site, theta0, thetav, mass, bias, charge):
        self.L.blade_add_msld_initialconditions(self.h, block, site, theta0, thetav, mass, bias, charge)

    def add_msld_flags(self, gamma, fnex, soft, soft14, ewaldtype, krest, sbr, sbe, snbe, fix):
        self.L.blade_add_msld_flags(self.h, gamma, fnex, soft, soft14, ewaldtype, krest, sbr, sbe, snbe, fix)

    def add_msld_atomrestraint(self, atoms):
        """one restraint group of 1-based atom indices"""
        self.L.blade_add_msld_atomrestraint(self.h)
        for a in atoms:
            self.L.blade_add_msld_atomrestraint_element(self.h, int(a))

    def add_msld_charges(self, k1=0.0, k2=0.0, k3=0.0, q0=0.0, w=0.0):
        """charge restraints on the lambda-weighted net charge of the blocks'
        initial-condition charges (msld/msld.h:154)"""
        self.L.blade_add_msld_charges(self.h, float(k1), float(k2), float(k3), float(q0), float(w))

    def add_msld_termscaling(self, bond, urey, angle, dihe, impr, cmap):
        self.L.blade_add_msld_termscaling(self.h, bond, urey, angle, dihe, impr, cmap)

    def add_msld_bias(self, i, j, typ, l0, k, n):
        self.L.blade_add_msld_bias(self.h, i, j, typ, l0, k, n)

    def init_run(self):
        self.L.blade_init_run(self.h)

    def add_run_flags(self, gamma, beta, rcut, rswitch, vdw, elec, gridspace, gx, gy, gz, order, shaketol):
        self.L.blade_add_run_flags(self.h, gamma, beta, rcut, rswitch, vdw, elec, gridspace, gx, gy, gz,
                                   order, shaketol)

    def add_run_dynopts(self, step, step0, nsteps, dt, T, freqnpt=0, volfluct=0.0, pressure=0.0):
        self.L.blade_add_run_dynopts(self.h, step, step0, nsteps, dt, T, freqnpt, volfluct, pressure)

    def npt_stats(self):
        """(volume moves tried, kept, current box) of the barostat"""
        m, a = C.c_long(), C.c_long()
        box = np.zeros(3)
        check(self.L.blade_get_npt_stats(self.h, C.byref(m), C.byref(a), ptr(box)), "blade_get_npt_stats")
        return m.value, a.value, box

    def set_seed(self, seed):
        self.L.blade_set_seed(self.h, seed)

    def set_domdec_frequency(self, steps):
        self.L.blade_set_domdec_frequency(self.h, steps)

    def set_fft_mode(self, mode):
        self.L.blade_set_fft_mode(self.h, mode)

    def set_output(self, nrg=None, lmd=None, freq_nrg=10, freq_lmd=10, lmd_text=True):
        """.nrg and .lmd files written by run_dynamics every freq steps; the
        .lmd frames are text lines, or XDR frames with lmd_text=False"""
        self.L.blade_set_lmd_format(self.h, int(bool(lmd_text)))
        check(self.L.blade_set_output_nrg(self.h, str(nrg).encode() if nrg else None, int(freq_nrg) if nrg else 0),
              "blade_set_output_nrg")
        check(self.L.blade_set_output_lmd(self.h, str(lmd).encode() if lmd else None, int(freq_lmd) if lmd else 0),
              "blade_set_output_lmd")

    def write_checkpoint(self, path):
        check(self.L.blade_write_checkpoint(self.h, str(path).encode()), "blade_write_checkpoint")

    def read_checkpoint(self, path):
        check(self.L.blade_read_checkpoint(self.h, str(path).encode()), "blade_read_checkpoint")

    def set_tile_chunk(self, tiles):
        self.L.blade_set_tile_chunk(self.h, tiles)

    def autotune(self, steps=300):
        """time candidate FFT forms, list cadences and tile chunks from the
        current coordinates and keep the fastest (before the first energy
        or dynamics call); returns the chosen values and every trial"""
        from ._ffi import TuneResult
        r = TuneResult()
        check(self.L.blade_autotune(self.h, int(steps), C.byref(r)), "blade_autotune")
        return r.as_dict()

    def _read(self, fn, path):
        if fn(self.h, str(path).encode()) != 0:
            raise RuntimeError(self.L.blade_last_error().decode())

    def read_parameter_file(self, path):
        self._read(self.L.blade_read_parameter_file, path)

    def read_psf(self, path):
        self._read(self.L.blade_read_psf, path)

    def read_crd(self, path):
        self._read(self.L.blade_read_crd, path)

    def read_pdb(self, path):
        self._read(self.L.blade_read_pdb, path)

    def run_energy(self):
        self.L.blade_run_energy(self.h)
        err = self.L.blade_system_error(self.h).decode()
        if err:
            raise RuntimeError(err)

    # -- extensions -----------------------------------------------------------
    def prepare(self):
        check(self.L.blade_prepare(self.h), "blade_prepare")

    def run_dynamics(self, nsteps, freq_nrg=0, use_graph=True):
        check(self.L.blade_run_dynamics(self.h, nsteps, freq_nrg, int(use_graph)), "blade_run_dynamics")

    def topology(self):
        p = self.L.blade_get_topology(self.h)
        if not p:
            self.prepare()
            p = self.L.blade_get_topology(self.h)
        return p

    def runparams(self):
        p = self.L.blade_get_runparams(self.h)
        if not p:
            self.prepare()
            p = self.L.blade_get_runparams(self.h)
        return p

    def device(self):
        d = self.L.blade_get_device_system(self.h)
        if not d:
            raise RuntimeError("device system not initialised")
        return d

    def energy(self):
        e = np.zeros(_ffi.NTERMS)
        check(self.L.blade_get_energy(self.h, ptr(e)), "get_energy")
        return e

    def forces(self):
        f = np.zeros(3 * self.n)
        check(self.L.blade_get_forces(self.h, ptr(f)), "get_forces")
        return f

    def positions(self):
        x = np.zeros(3 * self.n)
        check(self.L.blade_get_positions(self.h, ptr(x)), "get_positions")
        return x

    def set_positions(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(self.L.blade_set_positions(self.h, ptr(x)), "set_positions")

    def lambdas(self):
        v = np.zeros(max(1, self.nblocks))
        check(self.L.blade_get_lambda(self.h, ptr(v)), "get_lambda")
        return v

    def dudl(self):
        v = np.zeros(max(1, self.nblocks))
        check(self.L.blade_get_dudl(self.h, ptr(v)), "get_dudl")
        return v


def soa(pos):
    """(N,3) -> SoA flat [x..., y..., z...]"""
    return np.ascontiguousarray(np.asarray(pos, dtype=np.float64).T).ravel()


def read_lmd_xdr(path):
    """XDR .lmd frames (io/io.cxx:347-370): big-endian int32 block count,
    then that many big-endian IEEE singles; returns (frames, count) float32"""
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw, ">i4", 1, o)[0])
        out.append(np.frombuffer(raw, ">f4", n, o + 4).astype(np.float32))
        o += 4 + 4 * n
    return np.array(out)


def build(spec, gpus=(0,)):
    """Create a BladeSystem from a synthetic Spec through the blade_* API."""
    s = BladeSystem(gpus)
    s.init_parameters()
    for t, m in spec.types.items():
        s.add_parameter_atoms(t, m)
    for (a, b), (kb, b0) in spec.bond_par.items():
        s.add_parameter_bonds(a, b, kb, b0)
    for (a, b, c), (k, t0, kub, s0) in spec.angle_par.items():
        s.add_parameter_angles(a, b, c, k, t0, kub, s0)
    for key, terms in spec.dihe_par.items():
        for (k, n, d0) in terms:
            s.add_parameter_dihes(*key, k, n, d0)
    for key, (k, n, d0) in spec.impr_par.items():
        s.add_parameter_imprs(*key, k, n, d0)
    for key, grid in getattr(spec, "cmap_par", {}).items():
        g = np.asarray(grid, float)
        tb = [_b(x) for x in key]
        s.L.blade_add_parameter_cmaps(s.h, *tb, g.shape[0])
        for i in range(g.shape[0]):
            for j in range(g.shape[1]):
                s.L.blade_add_parameter_cmaps_fill(s.h, *tb, i + 1, j + 1, float(g[i, j]))
    for t, (eps, sig, eps14, sig14, e14, comb) in spec.nb_par.items():
        s.add_parameter_nbonds(t, eps, sig, eps14, sig14, e14, comb)
    for (t1, t2), (eps, sig, eps14, sig14) in getattr(spec, "nbfix_par", {}).items():
        s.add_parameter_nbfixs(t1, t2, eps, sig, eps14, sig14)
    s.init_structure()
    for i, (seg, resid, resn, nm, t, q, m) in enumerate(spec.atoms):
        s.add_atom(i + 1, seg, resid, resn, nm, t, q, m)
    for a, b in spec.bonds:
        s.add_bond(a + 1, b + 1)
    for a, b, c in spec.angles:
        s.add_angle(a + 1, b + 1, c + 1)
    for a, b, c, d in spec.dihes:
        s.add_dihe(a + 1, b + 1, c + 1, d + 1)
    for a, b, c, d in spec.imprs:
        s.add_impr(a + 1, b + 1, c + 1, d + 1)
    for q in getattr(spec, "cmaps", []):
        s.L.blade_add_cmap(s.h, *[int(a) + 1 for a in q])
    for (v, h0, h1, dist, scale) in getattr(spec, "virt2", None) or []:
        s.L.blade_add_virt2(s.h, int(v) + 1, int(h0) + 1, int(h1) + 1, float(dist), float(scale))
    for (v, h0, h1, h2, dist, theta, phi) in getattr(spec, "virt3", None) or []:
        s.L.blade_add_virt3(s.h, int(v) + 1, int(h0) + 1, int(h1) + 1, int(h2) + 1, float(dist), float(theta),
                            float(phi))
    rs = getattr(spec, "restraints", None) or {}
    for (i, j, rmin, kmin, rmax, kmax, rpeak, rsw, nsw, cx, cy, cz, pnoe) in rs.get("noe", []):
        s.L.blade_add_noe(s.h, i + 1, j + 1, rmin, kmin, rmax, kmax, rpeak, rsw, nsw, cx, cy, cz, bool(pnoe))
    for (i, k, x0, y0, z0, n) in rs.get("harm", []):
        s.L.blade_add_harmonic(s.h, i + 1, k, x0, y0, z0, n)
    for (i, j, kr, r0, b) in rs.get("borest", []):
        s.L.blade_add_borest(s.h, i + 1, j + 1, kr, r0, b + 1)
    for (i, j, k, kt, t0, b) in rs.get("anrest", []):
        s.L.blade_add_anrest(s.h, i + 1, j + 1, k + 1, kt, t0, b + 1)
    for (i, j, k, l, kp, n, p0, w, b) in rs.get("direst", []):
        s.L.blade_add_direst(s.h, i + 1, j + 1, k + 1, l + 1, kp, n, p0, w, b + 1)
    for (i1, i2, j1, j2, ci, cj, rd, kd) in rs.get("resd", []):
        s.L.blade_add_resd(s.h, i1 + 1, i2 + 1, j1 + 1, j2 + 1, ci, cj, rd, kd)
    s.add_shake(spec.shake)
    s.init_coordinates(spec.n)
    for i in range(spec.n):
        s.add_coordinates_position(i + 1, *spec.pos[i])
        if spec.vel is not None:
            s.add_coordinates_velocity(i + 1, *spec.vel[i])
    s.add_coordinates_box(0, *spec.box)
    s.init_msld(spec.nblocks - 1)          # substituent blocks; block 0 is the environment
    if spec.nblocks > 1:
        for i in np.nonzero(spec.atom_block)[0]:
            s.add_msld_atomassignment(int(i) + 1, int(spec.atom_block[i]) + 1)
        for b in range(spec.nblocks):
            tv = spec.theta_v[b] if spec.theta_v is not None else 0.0
            s.add_msld_initialconditions(b + 1, spec.block_site[b] + 1, spec.theta0[b], tv,
                                         spec.theta_mass[b], spec.fixed_bias[b],
                                         spec.block_charge[b] if getattr(spec, "block_charge", None) else 0.0)
        s.add_msld_flags(spec.msld_gamma, spec.fnex, int(spec.soft_core), int(getattr(spec, "soft_core14", False)),
                         int(getattr(spec, "ewald_type", 2)), getattr(spec, "k_restraint", 59.2),
                         *getattr(spec, "soft_bond_params", (1.0, 2.0, 1.0)),
                         int(getattr(spec, "fix_lambda", False)))
        for i, j in getattr(spec, "soft_bonds", None) or []:
            s.L.blade_add_msld_softbond(s.h, int(i) + 1, int(j) + 1)
        s.add_msld_termscaling(*[int(x) for x in spec.scale_terms])
        for b, g in enumerate(getattr(spec, "block_friction", None) or []):
            if g is not None:
                s.L.blade_add_msld_block_friction(s.h, b + 1, float(g))
        for b, fx in enumerate(getattr(spec, "block_fixed", None) or []):
            if fx:
                s.L.blade_set_msld_block_fixed(s.h, b + 1, 1)
        nsite = max(spec.block_site) + 1
        for i, kn in enumerate(getattr(spec, "theta_coll", None) or []):
            s.L.blade_add_msld_thetacollbias(s.h, nsite, i + 1, float(kn[0]), float(kn[1]))
        for i, k in enumerate(getattr(spec, "theta_inde", None) or []):
            s.L.blade_add_msld_thetaindebias(s.h, nsite, i + 1, float(k))
        if getattr(spec, "piecewise", None):
            s.L.blade_set_msld_piecewise_constraint(s.h, 1, float(spec.piecewise[0]), float(spec.piecewise[1]))
        if getattr(spec, "theta_edge", None):
            s.L.blade_set_msld_thetaedgebias(s.h, *[float(v) for v in spec.theta_edge])
        if getattr(spec, "charge_restraints", None):
            s.add_msld_charges(*spec.charge_restraints)
        for g in getattr(spec, "atom_restraints", []):
            s.add_msld_atomrestraint([a + 1 for a in g])
        for (i, j, t, l0, k, n) in spec.var_bias:
            s.add_msld_bias(i + 1, j + 1, t, l0, k, n)
    s.init_run()
    gx, gy, gz = spec.grid
    s.add_run_flags(spec.gamma, spec.beta, spec.rcut, spec.rswitch, spec.vdw_method, spec.elec_method,
                    spec.grid_space if gx <= 0 else -1.0, gx, gy, gz, spec.order, spec.shake_tol)
    s.add_run_dynopts(0, 0, spec.nsteps, spec.dt, spec.temperature, getattr(spec, "freq_npt", 0),
                      getattr(spec, "volume_fluct", 100.0), getattr(spec, "pressure_atm", 1.0) * ATMOSPHERE)
    s.set_seed(spec.seed)
    s.set_fft_mode(getattr(spec, "fft_mode", 0))
    s.set_domdec_frequency(getattr(spec, "freq_domdec", 10))
    s.set_tile_chunk(getattr(spec, "tile_chunk", 0))
    s.n = spec.n
    s.nblocks = spec.nblocks
    return s


def load_charmm(prm, psf, crd, box, shake=True, gpus=(0,), dt_ps=0.002, temperature=300.0, gamma_ps=1.0,
                beta=1.0 / 3.2, rcut=10.0, rswitch=8.5, grid_space=1.0, order=6, seed=0xB1ADE000):
    """BladeSystem from CHARMM files, as the script commands `parameters
    file`, `structure file psf`, `structure shake hbond`, `coordinates file
    crd`, `coordinates box`, `msld nblocks 0` and `run setvariable` would
    set it up (no lambda sites)."""
    from .synthetic import PICOSECOND
    s = BladeSystem(gpus)
    s.init_parameters()
    for p in ([prm] if isinstance(prm, (str, bytes)) or not hasattr(prm, "__iter__") else prm):
        s.read_parameter_file(p)
    s.read_psf(psf)
    s.add_shake(1 if shake else 0)
    if crd.lower().endswith(".pdb"):
        s.read_pdb(crd)
    else:
        s.read_crd(crd)
    s.add_coordinates_box(0, *box)
    s.init_msld(0)
    s.init_run()
    s.add_run_flags(gamma_ps / PICOSECOND, beta, rcut, rswitch, 0, 1, grid_space, -1, -1, -1, order, 2e-7)
    s.add_run_dynopts(0, 0, 0, dt_ps * PICOSECOND, temperature)
    s.set_seed(seed)
    s.n = s.L.blade_structure_atom_count(s.h)
    s.nblocks = 1
    return s


# Relaxation schedule for synthetic starts: (dt fs, gamma 1/ps, steps,
# steps between list rebuilds).  Lattice-built systems start far from
# equilibrium (random water orientations release ~10 kcal/mol per water);
# heavily damped short steps drain that energy before production-length
# steps, and lists are rebuilt every step while atoms still move faster than
# the thermal list pad assumes.
RELAX_SCHEDULE = ((0.1, 1000.0, 400, 1), (0.5, 200.0, 400, 1), (1.0, 50.0, 400, 1),
                  (2.0, 5.0, 300, 10))


class Batch:
    """Independent BladeSystems on one GPU advanced together: one graph
    launch per step runs every member's step as a parallel branch."""

    def __init__(self, systems):
        self.systems = list(systems)
        self.L = self.systems[0].L
        devs = (C.c_void_p * len(self.systems))(*[s.device() for s in self.systems])
        h = C.c_void_p()
        check(self.L.bm_batch_create(len(self.systems), devs, C.byref(h)), "bm_batch_create")
        self.h = h

    def dynamics(self, step0, nsteps, freq_nrg=0):
        check(self.L.bm_batch_dynamics(self.h, step0, nsteps, freq_nrg), "bm_batch_dynamics")

    def capture_graphs(self):
        """capture every batch step graph variant now (none is captured
        by a later dynamics call)"""
        check(self.L.bm_batch_capture_graphs(self.h), "bm_batch_capture_graphs")

    def sync(self):
        check(self.L.bm_batch_sync(self.h), "bm_batch_sync")

    def exchange(self, step, parity, seed, replica):
        """one round of neighbour replica exchange (bm_batch_exchange):
        returns (replica labels, accepted per lower member, dW per lower
        member)"""
        n = len(self.systems)
        rep = np.ascontiguousarray(replica, dtype=np.int32).copy()
        acc = np.zeros(n, dtype=np.int32)
        dw = np.zeros(n)
        i32 = C.POINTER(C.c_int32)
        check(self.L.bm_batch_exchange(self.h, int(step), int(parity), int(seed), rep.ctypes.data_as(i32),
                                       acc.ctypes.data_as(i32), ptr(dw)), "bm_batch_exchange")
        return rep, acc, dw

    def close(self):
        if self.h:
            self.L.bm_batch_destroy(self.h)
            self.h = None


def autotune_batch(systems, steps=150):
    """Time shared list cadences and tile chunks for BladeSystems that will
    run together as a Batch (before their first energy or dynamics call)
    and set the fastest on all of them; returns the choice and every trial."""
    from ._ffi import TuneResult
    L = systems[0].L
    hs = (C.c_void_p * len(systems))(*[s.h for s in systems])
    r = TuneResult()
    check(L.blade_batch_autotune(len(systems), hs, int(steps), C.byref(r)), "blade_batch_autotune")
    return r.as_dict()


def state(s, n, nblocks):
    """(pos (N,3), vel (N,3), theta, theta_v) of a running BladeSystem"""
    dev = s.device()
    x = np.zeros(3 * n)
    v = np.zeros(3 * n)
    th = np.zeros(max(1, nblocks))
    tv = np.zeros(max(1, nblocks))
    check(s.L.bm_get_positions(dev, ptr(x), ptr(th)), "bm_get_positions")
    check(s.L.bm_get_velocities(dev, ptr(v), ptr(tv)), "bm_get_velocities")
    return x.reshape(3, n).T.copy(), v.reshape(3, n).T.copy(), th, tv


def relax(spec, schedule=RELAX_SCHEDULE, gpus=(0,), log=None, use_graph=True):
    """Copy of `spec` advanced through the relaxation schedule on the GPU
    (positions, velocities and theta replaced; run options kept)."""
    import copy

    from .synthetic import PICOSECOND
    cur = copy.copy(spec)
    for dt_fs, gamma_ps, steps, freq_dd in schedule:
        stage = copy.copy(cur)
        stage.dt = dt_fs * 0.001 * PICOSECOND
        stage.gamma = gamma_ps / PICOSECOND
        s = build(stage, gpus)
        s.set_domdec_frequency(freq_dd)
        s.run_dynamics(steps - 1, freq_nrg=0, use_graph=use_graph)
        s.run_dynamics(1, freq_nrg=1, use_graph=use_graph)
        if log:
            e = s.energy()
            log("relax dt=%.1f fs gamma=%g/ps steps=%d: Epot %.1f KE %.1f" % (dt_fs, gamma_ps, steps, e[20], e[21]))
        pos, vel, th, tv = state(s, spec.n, spec.nblocks)
        s.close()
        cur = copy.copy(cur)
        cur.pos, cur.vel = pos, vel
        if spec.nblocks > 1:
            cur.theta0, cur.theta_v = list(th[:spec.nblocks]), list(tv[:spec.nblocks])
    return cur
