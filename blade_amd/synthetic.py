"""Seeded synthetic MSLD systems for the benchmark workloads.

Each generator returns a `Spec`: CHARMM-style parameters, structure,
coordinates, lambda sites and run options, which `blade_amd.system.build`
feeds through the reference-compatible blade_* API.  Units are the
reference's library units: kcal/mol, A, amu, e, AKMA time.

Workloads (sizes as named in the benchmark descriptions):
  water_box     N TIP3P waters on a jittered lattice, random orientations
  config1       216 TIP3P waters, 18.856 A box, switch 8-9 A, kappa 0.34,
                PME 20^3 order 6, 298 K, gamma 1/ps, dt 2 fs, LJ on O only
  config2       23,558 atoms in a 62.23 A cube: 7,023 TIP3P waters plus a
                2,489-atom synthetic C/N/O/H polymer carrying one lambda
                site of 3 substituents x 8 atoms; switch 9-10 A, PME 64^3
                order 6, fnex 5
  config3       92,224 atoms in a 96 x 96 x 100 A box: about 29,000 TIP3P
                waters, 78 Na+/Cl- pairs (150 mM), 16 seeded synthetic chains
                of 1.5-3 kDa, two of them carrying a lambda site of 4
                substituents x 8 atoms; switch 9-10 A, PME 96^3 order 6 with
                the split-BF16x3 MFMA FFT
  config5       one of 8 lambda-replicas of a 60,000-atom box: TIP3P water
                plus a 6,000-atom set of seeded chains, four of them
                carrying a lambda site of 4-8 substituents (up to 29
                blocks), per-block fixed biases, switch 9-10 A, PME 80^3
                order 6 on the MFMA FFT; replicas of one relaxed box differ
                in biases and velocities
  msld_small    a few-hundred-atom water box with a 3-substituent site
                (fast parity cases)
  two_site_small  a thousand-atom box with two chains, each carrying a
                lambda site (3 and 2 substituents), and two ion pairs; only
                dihedrals scale with lambda, and the substituents' first
                atoms are held together by atom restraints
"""
import math
from dataclasses import dataclass, field

import numpy as np

PICOSECOND = 20.454828284385908  # AKMA time units per ps
KBOLTZ = 0.001987191

# ---------------------------------------------------------------------------
# parameters (CHARMM conventions: K in kcal/mol/A^2 per the PRM files, the
# library doubles bond/angle constants; eps negative, Rmin/2)
# ---------------------------------------------------------------------------
WATER_TYPES = {"OT": 15.9994, "HT": 1.008}
WATER_BONDS = {("OT", "HT"): (450.0, 0.9572), ("HT", "HT"): (0.0, 1.5139)}
WATER_ANGLES = {("HT", "OT", "HT"): (55.0, 104.52, 0.0, 0.0)}
# LJ on the oxygen only (the benchmark variant of CHARMM TIP3P)
WATER_NB = {"OT": (-0.1521, 1.7682, -0.1521, 1.7682, 1.0, 0),
            "HT": (0.0, 0.2245, 0.0, 0.2245, 1.0, 0)}

POLY_TYPES = {"CT": 12.011, "NT": 14.007, "OP": 15.9994, "HA": 1.008, "HN": 1.008, "HO": 1.008}
POLY_BONDS = {("CT", "CT"): (222.5, 1.530), ("CT", "NT"): (320.0, 1.470), ("CT", "OP"): (360.0, 1.430),
              ("CT", "HA"): (309.0, 1.090), ("NT", "HN"): (440.0, 1.010), ("OP", "HO"): (545.0, 0.960)}
POLY_NB = {"CT": (-0.0780, 2.0500, -0.0100, 1.9000, 1.0, 0),
           "NT": (-0.2000, 1.8500, -0.2000, 1.5500, 1.0, 0),
           "OP": (-0.1521, 1.7700, -0.1521, 1.7700, 1.0, 0),
           "HA": (-0.0220, 1.3200, -0.0220, 1.3200, 1.0, 0),
           "HN": (-0.0460, 0.2245, -0.0460, 0.2245, 1.0, 0),
           "HO": (-0.0460, 0.2245, -0.0460, 0.2245, 1.0, 0)}
# CHARMM sodium and chloride
ION_TYPES = {"SOD": 22.98977, "CLA": 35.45}
ION_NB = {"SOD": (-0.0469, 1.41075, -0.0469, 1.41075, 1.0, 0),
          "CLA": (-0.1500, 2.27000, -0.1500, 2.27000, 1.0, 0)}
HEAVY_BOND = {frozenset(("CT", "CT")): 1.530, frozenset(("CT", "NT")): 1.470, frozenset(("CT", "OP")): 1.430}
H_OF = {"CT": ("HA", 1.090), "NT": ("HN", 1.010), "OP": ("HO", 0.960)}


@dataclass
class Spec:
    name: str
    box: tuple
    types: dict = field(default_factory=dict)        # name -> mass
    bond_par: dict = field(default_factory=dict)     # (t1,t2) -> (kb, b0)
    angle_par: dict = field(default_factory=dict)    # (t1,t2,t3) -> (k, t0deg, kub, s0)
    dihe_par: dict = field(default_factory=dict)     # (t1..t4) -> [(k, n, d0deg)]
    impr_par: dict = field(default_factory=dict)     # (t1..t4) -> (k, n, d0deg)
    nb_par: dict = field(default_factory=dict)       # t -> (eps, rmin2, eps14, rmin2_14, e14fac, combine)
    nbfix_par: dict = field(default_factory=dict)    # (t1,t2) -> (eps, rmin, eps14, rmin14) pair overrides
    atoms: list = field(default_factory=list)        # (seg, resid, resname, name, type, q, mass)
    bonds: list = field(default_factory=list)        # 0-based
    angles: list = field(default_factory=list)
    dihes: list = field(default_factory=list)
    imprs: list = field(default_factory=list)
    cmap_par: dict = field(default_factory=dict)     # (t1..t8) -> (ngrid, ngrid) kcal/mol, phi rows
    cmaps: list = field(default_factory=list)        # 8 atoms, 0-based
    pos: np.ndarray = None                           # (N,3) A
    vel: np.ndarray = None                           # (N,3) A/AKMA
    shake: bool = True
    # msld
    nblocks: int = 1
    atom_block: np.ndarray = None                    # (N,) 0-based blocks
    block_site: list = field(default_factory=lambda: [0])
    theta0: list = field(default_factory=lambda: [0.0])
    theta_v: list = None                             # theta velocities (None = 0)
    theta_mass: list = field(default_factory=lambda: [0.0])
    fixed_bias: list = field(default_factory=lambda: [0.0])
    var_bias: list = field(default_factory=list)     # (i, j, type, l0, k, n) 0-based blocks
    fnex: float = 5.5
    soft_core: bool = False
    scale_terms: tuple = (1, 1, 1, 1, 1, 1)           # lambda-scaled bond, urey, angle, dihe, impr, cmap
    atom_restraints: list = field(default_factory=list)  # groups of 0-based atoms held to their centroid
    k_restraint: float = 59.2                         # kcal/mol/A^2
    block_charge: list = None                         # blockCharge of the initial conditions (the library
                                                      # uses each block's atom-charge sum, as the reference)
    charge_restraints: tuple = None                   # (k1, k2, k3, q0, w) on Q = sum lambda_b q_b
    soft_core14: bool = False                         # soft-core 1-4 pairs (useSoftCore14)
    soft_bonds: list = field(default_factory=list)    # MSLD soft bonds, 0-based atom pairs
    soft_bond_params: tuple = (1.0, 2.0, 1.0)         # softBondRadius A, softBondExponent, softNotBondExponent
    ewald_type: int = 2                               # msldEwaldType: 1 on, 2 ex, 3 nn
    block_friction: list = None                       # per-block theta friction, 1/AKMA time (None = msld_gamma)
    block_fixed: list = None                          # per-block fixed flags (lambda held at theta)
    theta_coll: list = None                           # per site (k, n) collective theta bias
    theta_inde: list = None                           # per site k independent theta bias
    piecewise: tuple = None                           # (width, k): piecewise lambda constraint
    theta_edge: tuple = None                          # (k, N, alpha, phi) edge bias of the piecewise form
    fix_lambda: bool = False                          # msld fix flag: lambda held at the initial theta
    virt2: list = None                                # two-host virtual sites (site, h0, h1, dist, scale), 0-based
    virt3: list = None                                # three-host sites (site, h0, h1, h2, dist, theta deg, phi deg)
    restraints: dict = None                           # {"noe"|"harm"|"borest"|"anrest"|"direst"|"resd": [rows]},
                                                      # 0-based atoms and blocks, angles in degrees
    # run options
    dt: float = 0.002 * PICOSECOND
    temperature: float = 298.0
    gamma: float = 1.0 / PICOSECOND
    msld_gamma: float = 1.0 / PICOSECOND
    beta: float = 1.0 / 3.2
    rcut: float = 10.0
    rswitch: float = 9.0
    grid: tuple = (-1, -1, -1)
    grid_space: float = 1.0
    order: int = 6
    vdw_method: int = 0      # 0 force switch, 1 potential switch, 2 shift
    elec_method: int = 1     # 0 force switch, 1 PME, 2 force shift
    fft_mode: int = 0        # PME FFT: 0 FP32 Stockham, 1 MFMA split-BF16x3
    freq_domdec: int = 10    # neighbour-list rebuild cadence (steps)
    freq_npt: int = 0        # Monte Carlo barostat move every freq_npt steps (0: constant volume)
    volume_fluct: float = 100.0   # rms volume move, A^3
    pressure_atm: float = 1.0
    tile_chunk: int = 0      # directed tiles per nonbonded work unit (0: 16)
    shake_tol: float = 2e-7
    seed: int = 0xB1ADE000
    nsteps: int = 1000

    @property
    def n(self):
        return len(self.atoms)


def _rot(rng):
    """uniform random rotation matrix (unit quaternion)"""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _water_frame():
    b, th = 0.9572, math.radians(104.52)
    h1 = np.array([b * math.sin(th / 2), b * math.cos(th / 2), 0.0])
    h2 = np.array([-b * math.sin(th / 2), b * math.cos(th / 2), 0.0])
    return np.zeros(3), h1, h2


class _Grid:
    """periodic cell list for clash checks while building"""

    def __init__(self, box, cell=3.0):
        self.box = np.asarray(box, float)
        self.nc = np.maximum(1, (self.box // cell).astype(int))
        self.cells = {}
        self.pts = []

    def key(self, p):
        f = np.floor((p % self.box) / self.box * self.nc).astype(int) % self.nc
        return tuple(f)

    def add(self, p):
        self.pts.append(np.asarray(p, float))
        self.cells.setdefault(self.key(p), []).append(len(self.pts) - 1)

    def remove(self, i):
        """drop point i from the cell lists (its index stays reserved)"""
        self.cells[self.key(self.pts[i])].remove(i)

    def near(self, p):
        k = self.key(p)
        out = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    kk = ((k[0] + dx) % self.nc[0], (k[1] + dy) % self.nc[1], (k[2] + dz) % self.nc[2])
                    out += self.cells.get(kk, [])
        return out

    def mindist(self, p, skip=()):
        idx = [i for i in self.near(p) if i not in skip]
        if not idx:
            return 1e9
        d = np.array([self.pts[i] for i in idx]) - p
        d -= self.box * np.round(d / self.box)
        return float(np.sqrt((d * d).sum(1)).min())


def _maxwell(rng, masses, T):
    sig = np.sqrt(KBOLTZ * T / np.asarray(masses))[:, None]
    v = rng.normal(size=(len(masses), 3)) * sig
    m = np.asarray(masses)[:, None]
    v -= (m * v).sum(0) / m.sum()
    return v


def _add_water(spec, o, rot, resid):
    base = len(spec.atoms)
    f0, f1, f2 = _water_frame()
    c = o - f0
    for nm, t, q, f in (("OH2", "OT", -0.834, f0), ("H1", "HT", 0.417, f1), ("H2", "HT", 0.417, f2)):
        spec.atoms.append(("WAT", str(resid), "TIP3", nm, t, q, WATER_TYPES[t]))
        spec._pos.append(o + rot @ f)
    spec.bonds += [(base, base + 1), (base, base + 2), (base + 1, base + 2)]
    spec.angles.append((base + 1, base, base + 2))


def _water_params(spec):
    spec.types.update(WATER_TYPES)
    spec.bond_par.update(WATER_BONDS)
    spec.angle_par.update(WATER_ANGLES)
    spec.nb_par.update(WATER_NB)


def water_box(nwat, box, seed, name="water"):
    rng = np.random.default_rng(seed)
    spec = Spec(name=name, box=tuple(float(b) for b in box))
    spec._pos = []
    _water_params(spec)
    n1 = int(math.ceil(round(nwat ** (1 / 3), 9)))
    sites = [(i, j, k) for i in range(n1) for j in range(n1) for k in range(n1)]
    rng.shuffle(sites)
    sp = np.array(box) / n1
    for r, (i, j, k) in enumerate(sorted(sites[:nwat])):
        o = (np.array([i, j, k]) + 0.5) * sp
        _add_water(spec, o, _rot(rng), r + 1)
    spec.pos = np.array(spec._pos)
    spec.atom_block = np.zeros(spec.n, dtype=np.int32)
    return spec, rng


# TIP4P (Jorgensen et al. 1983): charge on a massless site M 0.15 A from
# the oxygen along the HOH bisector, LJ on the oxygen
TIP4P_TYPES = {"OT4": 15.9994, "HT": 1.008, "LP": 0.0}
TIP4P_NB = {"OT4": (-0.155, 1.7699, -0.155, 1.7699, 1.0, 0),
            "LP": (0.0, 0.2245, 0.0, 0.2245, 1.0, 0)}


def tip4p_box(nwat, box, seed, colinear=0, offaxis=0, name="tip4p"):
    """nwat TIP4P waters (O, H1, H2, M) on a shuffled lattice; M is a
    three-host bisector site (dist -0.15, angle 0, dihedral 0).  The first
    `colinear` waters also carry a two-host site beyond H1 (hosts H1, O;
    dist 0.2 A, scale 0.1), the next `offaxis` a three-host site off the
    molecular plane (hosts O, H1, H2; dist 0.5 A, angle 110, dihedral 75);
    each takes 0.1 e from its H1."""
    rng = np.random.default_rng(seed)
    spec = Spec(name=name, box=tuple(float(b) for b in box))
    spec.types.update(TIP4P_TYPES)
    spec.bond_par.update({("OT4", "HT"): (450.0, 0.9572), ("HT", "HT"): (0.0, 1.5139)})
    spec.angle_par.update({("HT", "OT4", "HT"): (55.0, 104.52, 0.0, 0.0)})
    spec.nb_par.update(TIP4P_NB)
    spec.nb_par["HT"] = WATER_NB["HT"]
    spec.virt2, spec.virt3 = [], []
    pos = []
    n1 = int(math.ceil(round(nwat ** (1 / 3), 9)))
    sites = [(i, j, k) for i in range(n1) for j in range(n1) for k in range(n1)]
    rng.shuffle(sites)
    sp = np.array(box) / n1
    f0, f1, f2 = _water_frame()
    bis = (f1 + f2) / np.linalg.norm(f1 + f2)
    for r, (i, j, k) in enumerate(sorted(sites[:nwat])):
        o = (np.array([i, j, k]) + 0.5) * sp
        rot = _rot(rng)
        base = len(spec.atoms)
        extra = 1 if r < colinear + offaxis else 0
        qh1 = 0.52 - 0.1 * extra
        for nm, t, q, f in (("OH2", "OT4", 0.0, f0), ("H1", "HT", qh1, f1), ("H2", "HT", 0.52, f2),
                            ("OM", "LP", -1.04, 0.15 * bis)):
            spec.atoms.append(("WAT", str(r + 1), "TIP4", nm, t, q, TIP4P_TYPES[t]))
            pos.append(o + rot @ f)
        spec.bonds += [(base, base + 1), (base, base + 2), (base + 1, base + 2)]
        spec.angles.append((base + 1, base, base + 2))
        spec.virt3.append((base + 3, base, base + 1, base + 2, -0.15, 0.0, 0.0))
        if extra:
            spec.atoms.append(("WAT", str(r + 1), "TIP4", "LX", "LP", 0.1, 0.0))
            pos.append(o + rot @ (1.5 * f1))   # any start: the site is placed from its hosts
            if r < colinear:
                spec.virt2.append((base + 4, base + 1, base, 0.2, 0.1))
            else:
                spec.virt3.append((base + 4, base, base + 1, base + 2, 0.5, 110.0, 75.0))
    spec.pos = np.array(pos)
    spec.atom_block = np.zeros(spec.n, dtype=np.int32)
    m = np.array([a[6] for a in spec.atoms])
    v = np.zeros((spec.n, 3))
    heavy = m > 0
    v[heavy] = _maxwell(rng, m[heavy], spec.temperature)
    spec.vel = v
    return spec


def config1(seed=0xB1ADE001):
    """216 TIP3P waters, 648 atoms (the water216 benchmark case)."""
    spec, rng = water_box(216, (18.856,) * 3, seed, name="config1-water216")
    spec.rcut, spec.rswitch, spec.beta = 9.0, 8.0, 0.34
    spec.grid, spec.order = (20, 20, 20), 6
    spec.temperature = 298.0
    spec.nsteps = 1000
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


# ---------------------------------------------------------------------------
# synthetic polymer
# ---------------------------------------------------------------------------
def _poly_params(spec):
    spec.types.update(POLY_TYPES)
    spec.bond_par.update(POLY_BONDS)
    heavy = ["CT", "NT", "OP"]
    hyd = ["HA", "HN", "HO"]
    for a in heavy + hyd:
        for b in heavy:
            for c in heavy + hyd:
                spec.angle_par[(a, b, c)] = (50.0, 111.0 if b == "CT" else 110.0, 0.0, 0.0)
    for b in heavy:
        for c in heavy:
            spec.dihe_par[("X", b, c, "X")] = [(0.2, 3, 0.0), (0.1, 1, 180.0)]
    spec.impr_par[("NT", "X", "X", "HN")] = (20.0, 0, 0.0)
    spec.nb_par.update(POLY_NB)


def _place(prev, prev2, prev3, b, theta, phi):
    """position at bond length b, angle theta (prev2-prev-new), dihedral phi"""
    bc = prev - prev2
    bc /= np.linalg.norm(bc)
    n = np.cross(prev2 - prev3, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-6:
        n = np.cross(bc, np.array([1.0, 0.3, 0.1]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    d2 = np.array([-b * math.cos(theta), b * math.sin(theta) * math.cos(phi), b * math.sin(theta) * math.sin(phi)])
    return prev + d2[0] * bc + d2[1] * m + d2[2] * n


def _polymer(spec, rng, natoms, center, radius, grid, site_at=None, nsubs=3, sub_size=8, typed_subs=False):
    """Self-avoiding chain of C/N/O heavy atoms with hydrogens, confined to a
    sphere, exactly `natoms` atoms including the lambda-site substituents.
    typed_subs: substituent charges by the body's rule (hydrogens positive by
    type, heavy atoms balancing their hydrogens, each substituent neutral)
    instead of plain N(0, 0.35) draws."""
    _poly_params(spec)
    body = natoms - (nsubs * sub_size if site_at is not None else 0)
    # the attachment carbon gives up one hydrogen to the substituents
    budget = body + (1 if site_at is not None else 0)
    nH = {"CT": 2, "NT": 1, "OP": 0}
    pos, types, gidx = [], [], []

    def ok(p, confine):
        d = p - center
        d -= np.array(spec.box) * np.round(d / np.array(spec.box))
        if confine and np.linalg.norm(d) >= radius:
            return False
        # the 1-4 partner may sit gauche; everything farther along the chain
        # keeps room for the hydrogens (no syn-pentane contacts)
        if len(pos) >= 3:
            r14 = pos[-3] - p
            r14 -= np.array(spec.box) * np.round(r14 / np.array(spec.box))
            if np.linalg.norm(r14) < 2.9:
                return False
        return grid.mindist(p, skip=set(gidx[-3:])) > 3.7

    def extend(t, confine=True, tries=60):
        if len(types) == 0:
            p = np.array(center, float)
        elif len(types) == 1:
            p = pos[-1] + np.array([1.53, 0, 0])
        elif len(types) == 2:
            th = math.radians(111.0)
            p = pos[-1] + 1.53 * np.array([-math.cos(th), math.sin(th), 0.0])
        else:
            b = HEAVY_BOND[frozenset((types[-1], t))]
            for _ in range(tries):
                phi = rng.choice([math.pi, math.pi / 3, -math.pi / 3]) + rng.normal(0, 0.25)
                p = _place(pos[-1], pos[-2], pos[-3], b, math.radians(111.0), phi)
                if ok(p, confine):
                    break
            else:
                return False
        pos.append(np.asarray(p, float))
        types.append(t)
        gidx.append(len(grid.pts))
        grid.add(p)
        return True

    def retract(k):
        """drop the last k heavy atoms (dead end)"""
        nonlocal count
        for _ in range(min(k, len(types) - 3)):
            grid.remove(gidx.pop())
            pos.pop()
            count -= 1 + nH[types.pop()]

    # methyl start (3 H), then interior units; stop with room for a
    # terminal carbon carrying 0..3 hydrogens; back off on dead ends
    extend("CT")
    count = 4
    fails = 0
    while True:
        t = "CT" if len(types) < 3 else rng.choice(["CT", "CT", "CT", "NT", "OP"], p=[0.25, 0.25, 0.2, 0.15, 0.15])
        if t != "CT" and types[-1] != "CT":
            t = "CT"
        size = 1 + nH[t]
        if count + size + 4 > budget:
            break
        if extend(t):
            count += size
            continue
        fails += 1
        if fails > 4000:
            raise RuntimeError("polymer growth failed")
        retract(1 + fails % 6)
    while budget - count > 4:
        if extend("CT", confine=False, tries=200):
            count += 3
        else:
            retract(3)
    rem = budget - count
    while not extend("CT", confine=False, tries=200):
        retract(2)
        while budget - count > 4:
            if extend("CT", confine=False, tries=200):
                count += 3
        rem = budget - count
    hcount = [nH[t] for t in types]
    hcount[0] = 3
    hcount[-1] = rem - 1
    assert len(types) + sum(hcount) == budget
    q = rng.normal(0.0, 0.35, size=body)
    q -= q.mean()
    qi = iter(q)
    heavy_idx = []
    site_heavy = None
    if site_at is not None:
        site_heavy = max(1, int(site_at * len(types)))
        while types[site_heavy] != "CT" or hcount[site_heavy] != 2 or site_heavy >= len(types) - 1:
            site_heavy += 1
        hcount[site_heavy] -= 1   # replaced by the substituent attachment
    for k, t in enumerate(types):
        idx = len(spec.atoms)
        spec.atoms.append(("POLY", str(k // 4 + 1), "SYN", "%s%d" % (t[0], k), t, float(next(qi)), POLY_TYPES[t]))
        spec._pos.append(pos[k])
        heavy_idx.append(idx)
    hatoms = {k: [] for k in range(len(types))}
    def best_h(center_p, bonded, skip, blen, tries=400):
        """hydrogen direction away from bonded partners with the largest
        clearance from everything already placed"""
        best, bestd = None, -1.0
        units = [(x - center_p) / np.linalg.norm(x - center_p) for x in bonded]
        # near-tetrahedral window (100-125 degrees to every bonded partner),
        # widened until some direction fits
        for lo, hi, n_try in ((-0.57, -0.17, tries), (-0.75, -0.05, 4 * tries), (-0.95, 0.2, 8 * tries)):
            for _ in range(n_try):
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                if not all(lo < np.dot(d, u) < hi for u in units):
                    continue
                p = center_p + blen * d
                m = grid.mindist(p, skip=skip)
                if m > bestd:
                    best, bestd = p, m
            if best is not None:
                return best
        raise RuntimeError("no hydrogen position found")

    for k, t in enumerate(types):
        ht, hb = H_OF[t]
        nb = [pos[k - 1]] if k > 0 else []
        if k + 1 < len(types):
            nb.append(pos[k + 1])
        skip = set(gidx[max(0, k - 1):k + 2])
        for h in range(hcount[k]):
            if t == "NT" and len(nb) == 2:
                # planar amine hydrogen on the bisector
                u = (pos[k] - nb[0]) / np.linalg.norm(pos[k] - nb[0]) + (pos[k] - nb[1]) / np.linalg.norm(pos[k] - nb[1])
                p = pos[k] + hb * u / np.linalg.norm(u)
            else:
                p = best_h(pos[k], nb, skip, hb)
            idx = len(spec.atoms)
            spec.atoms.append(("POLY", str(k // 4 + 1), "SYN", "H%d_%d" % (k, h), ht, float(next(qi)), POLY_TYPES[ht]))
            spec._pos.append(p)
            skip.add(len(grid.pts))
            grid.add(p)
            hatoms[k].append(idx)
            nb.append(p)
    # charges: hydrogens positive by type with a small seeded spread, each
    # heavy atom balances its own hydrogens plus a larger spread, the body is
    # neutral.  Negative hydrogens would meet other hydrogens with no LJ core
    # between them and collapse.
    hq0 = {"HA": 0.09, "HN": 0.31, "HO": 0.42}
    nheavy = len(types)
    hq = iter(q[nheavy:])
    heavy_q = []
    for k in range(nheavy):
        hs = 0.0
        for h in hatoms[k]:
            a = spec.atoms[h]
            qh = max(0.03, hq0[a[4]] + 0.1 * float(next(hq)))
            spec.atoms[h] = a[:5] + (qh,) + a[6:]
            hs += qh
        heavy_q.append(-hs + 0.5 * float(q[k]))
    shift = (sum(heavy_q) + sum(spec.atoms[h][5] for k in range(nheavy) for h in hatoms[k])) / nheavy
    for k in range(nheavy):
        a = spec.atoms[heavy_idx[k]]
        spec.atoms[heavy_idx[k]] = a[:5] + (heavy_q[k] - shift,) + a[6:]
    # bonds
    for k in range(len(types)):
        if k > 0:
            spec.bonds.append((heavy_idx[k - 1], heavy_idx[k]))
        for h in hatoms[k]:
            spec.bonds.append((heavy_idx[k], h))
    # substituents: CH2-CH2-OH (8 atoms) per block, all bonded to site_heavy
    subs = []
    if site_at is not None:
        k = site_heavy
        anchor = pos[k]
        nbp = [pos[k - 1], pos[k + 1]] + [np.asarray(spec._pos[h]) for h in hatoms[k]]
        u = sum((anchor - x) / np.linalg.norm(anchor - x) for x in nbp)
        u /= np.linalg.norm(u)
        near = set(gidx[max(0, k - 1):k + 2])
        placed_all = []
        for s in range(nsubs):
            atoms_s = []
            qs = rng.normal(0.0, 0.35, size=sub_size)
            qs -= qs.mean()
            c1 = anchor + 1.53 * u
            # C2 and O on the largest-clearance torsions (alternative
            # substituents of one site overlap each other by design)
            best, bestd = None, -1
            for _ in range(24):
                phi = rng.uniform(-math.pi, math.pi)
                c2 = _place(c1, anchor, pos[k - 1], 1.53, math.radians(111.0), phi)
                phi2 = rng.uniform(-math.pi, math.pi)
                o = _place(c2, c1, anchor, 1.43, math.radians(111.0), phi2)
                bonded = {0: [anchor, c2], 1: [c1, o], 2: [c2]}
                hs = []
                for (hk, nh) in [(0, 2), (1, 2), (2, 1)]:
                    hb = H_OF[("CT", "CT", "OP")[hk]][1]
                    for j in range(nh):
                        p = best_h((c1, c2, o)[hk], bonded[hk], near, hb, tries=120)
                        bonded[hk].append(p)
                        hs.append((hk, p))
                m = min(grid.mindist(x, skip=near) for x in [c2, o] + [p for _, p in hs])
                if m > bestd:
                    best, bestd = (c2, o, hs), m
            c2, o, hs = best
            hv = [("C1", "CT", c1), ("C2", "CT", c2), ("O", "OP", o)]
            hidx = []
            for nm, t, p in hv:
                idx = len(spec.atoms)
                spec.atoms.append(("LIG", str(s + 1), "SUB%d" % (s + 1), nm, t, 0.0, POLY_TYPES[t]))
                spec._pos.append(p)
                hidx.append(idx)
                atoms_s.append(idx)
            for j, (hk, p) in enumerate(hs):
                ht = H_OF[hv[hk][1]][0]
                idx = len(spec.atoms)
                spec.atoms.append(("LIG", str(s + 1), "SUB%d" % (s + 1), "H%d%d" % (hk, j), ht, 0.0, POLY_TYPES[ht]))
                spec._pos.append(p)
                atoms_s.append(idx)
                spec.bonds.append((hidx[hk], idx))
            spec.bonds += [(heavy_idx[k], hidx[0]), (hidx[0], hidx[1]), (hidx[1], hidx[2])]
            if typed_subs:
                # hydrogens positive by type, heavy atoms balance their own
                # hydrogens plus a spread, the substituent is neutral
                qh = [max(0.03, hq0[spec.atoms[a][4]] + 0.1 * float(qv)) for a, qv in zip(atoms_s[3:], qs[3:])]
                qheavy = [-sum(q for (hk, _), q in zip(hs, qh) if hk == j) + 0.5 * float(qs[j]) for j in range(3)]
                shift = (sum(qheavy) + sum(qh)) / 3.0
                qs = [q - shift for q in qheavy] + qh
            for a, qv in zip(atoms_s, qs):
                rec = list(spec.atoms[a])
                rec[5] = float(qv)
                spec.atoms[a] = tuple(rec)
            subs.append(atoms_s)
            placed_all += atoms_s
        for a in placed_all:
            grid.add(spec._pos[a])
    return heavy_idx, subs


def _angles_dihes(spec, block_of):
    """all angles and proper dihedrals of the bond graph, skipping terms that
    mix two blocks of one site; impropers on NT with three partners"""
    n = spec.n
    nbr = [[] for _ in range(n)]
    for a, b in spec.bonds:
        nbr[a].append(b)
        nbr[b].append(a)
    have = set(tuple(sorted((a[0], a[2]))) + (a[1],) for a in spec.angles)

    def mixed(idx):
        bl = {block_of[i] for i in idx if block_of[i] != 0}
        return len(bl) > 1

    def water(i):
        return spec.atoms[i][4] in ("OT", "HT")

    for j in range(n):
        if water(j):
            continue
        for x in range(len(nbr[j])):
            for y in range(x + 1, len(nbr[j])):
                i, k = nbr[j][x], nbr[j][y]
                key = tuple(sorted((i, k))) + (j,)
                if key in have or mixed((i, j, k)):
                    continue
                ti, tk = spec.atoms[i][4], spec.atoms[k][4]
                if ti.startswith("H") and tk.startswith("H") and spec.atoms[j][4] == "OP":
                    continue
                spec.angles.append((i, j, k))
                have.add(key)
    for a, b in spec.bonds:
        if water(a) or spec.atoms[a][4].startswith("H") or spec.atoms[b][4].startswith("H"):
            continue
        for i in nbr[a]:
            if i == b:
                continue
            for l in nbr[b]:
                if l == a or l == i:
                    continue
                if mixed((i, a, b, l)):
                    continue
                spec.dihes.append((i, a, b, l))
    for j in range(n):
        if spec.atoms[j][4] == "NT" and len(nbr[j]) == 3:
            hs = [x for x in nbr[j] if spec.atoms[x][4] == "HN"]
            cs = [x for x in nbr[j] if spec.atoms[x][4] != "HN"]
            if len(hs) == 1 and len(cs) == 2:
                spec.imprs.append((j, cs[0], cs[1], hs[0]))


def _fill_water(spec, rng, nwat, grid, clearance=2.5, lattice=None):
    L = np.array(spec.box)
    n1 = lattice or int(math.ceil((nwat * 1.32) ** (1 / 3)))
    sp = L / n1
    cand = []
    for i in range(n1):
        for j in range(n1):
            for k in range(n1):
                o = (np.array([i, j, k]) + 0.5) * sp + rng.uniform(-0.1, 0.1, 3)
                if grid.mindist(o) > clearance:
                    cand.append(o)
    if len(cand) < nwat:
        raise RuntimeError("only %d water sites for %d waters" % (len(cand), nwat))
    pick = rng.choice(len(cand), size=nwat, replace=False)
    pick.sort()
    resid = 1
    for pi in pick:
        o = cand[pi]
        best, bestd = None, -1
        for _ in range(12):
            R = _rot(rng)
            f0, f1, f2 = _water_frame()
            h1, h2 = o + R @ f1, o + R @ f2
            dmin = min(grid.mindist(h1), grid.mindist(h2))
            if dmin > bestd:
                best, bestd = R, dmin
            if dmin > 1.9:
                break
        _add_water(spec, o, best, resid)
        f0, f1, f2 = _water_frame()
        for f in (f0, f1, f2):
            grid.add(o + best @ f)
        resid += 1


def _msld_site(spec, subs, fnex, theta_mass=12.0):
    _msld_sites(spec, [subs], fnex, theta_mass)


def _msld_sites(spec, sites, fnex, theta_mass=12.0):
    """lambda sites, each a list of substituent atom lists: block 0 is the
    environment, then the blocks of site 1, of site 2, ..."""
    blk = np.zeros(spec.n, dtype=np.int32)
    block_site = [0]
    for si, subs in enumerate(sites):
        for atoms in subs:
            block_site.append(si + 1)
            blk[atoms] = len(block_site) - 1
    nb = len(block_site)
    spec.nblocks = nb
    spec.atom_block = blk
    spec.block_site = block_site
    spec.theta0 = [0.0] * nb
    spec.theta_mass = [0.0] + [theta_mass] * (nb - 1)
    spec.fixed_bias = [0.0] * nb
    spec.fnex = fnex


def _add_ions(spec, rng, npairs, grid, clearance=3.0):
    """npairs Na+/Cl- at uniform random sites at least `clearance` from
    everything placed so far (the water fill then keeps its own clearance
    from them)"""
    spec.types.update(ION_TYPES)
    spec.nb_par.update(ION_NB)
    L = np.array(spec.box)
    resid = 1
    for kind, q in [("SOD", 1.0)] * npairs + [("CLA", -1.0)] * npairs:
        for _ in range(1000):
            p = rng.uniform(0.0, 1.0, 3) * L
            if grid.mindist(p) > clearance:
                break
        else:
            raise RuntimeError("no room for an ion")
        spec.atoms.append(("ION", str(resid), kind, kind, kind, q, ION_TYPES[kind]))
        spec._pos.append(p)
        grid.add(p)
        resid += 1


def _chains(spec, rng, sizes, grid, site_nsubs, rscale=28.0, sep=1.0):
    """seeded chains of the given atom counts, each confined to a sphere
    (the polymer density of the 23,558-atom box) at a random centre whose
    sphere overlaps no earlier one (centres at least sep times the sum of
    the radii apart); chain c carries a lambda site of
    site_nsubs[c] substituents when that entry is nonzero.  Returns the
    sites (lists of substituent atom lists)."""
    L = np.array(spec.box)
    sites, placed = [], []
    for c, n in enumerate(sizes):
        radius = max(7.0, rscale * (n / 2489.0) ** (1 / 3))
        for _ in range(10000):
            center = rng.uniform(0.0, 1.0, 3) * L
            ok = True
            for c0, r0 in placed:
                d = center - c0
                d -= L * np.round(d / L)
                ok = ok and float(np.linalg.norm(d)) > sep * (radius + r0)
            if ok:
                break
        else:
            raise RuntimeError("no room for chain %d" % c)
        placed.append((center, radius))
        ns = site_nsubs[c] if c < len(site_nsubs) else 0
        _, subs = _polymer(spec, rng, n, center, radius, grid, site_at=0.5 if ns else None,
                           nsubs=max(ns, 1), typed_subs=True)
        if ns:
            sites.append(subs)
    return sites


def _sites_block_of(spec, sites):
    block_of = np.zeros(spec.n, dtype=np.int32)
    b = 0
    for subs in sites:
        for atoms in subs:
            b += 1
            block_of[atoms] = b
    return block_of


def config3(seed=0xB1ADE003, box=(96.0, 96.0, 100.0), n_chains=16, chain_atoms=5302,
            n_ion_pairs=78, n_water=28922):
    """92,224-atom precision-form box: chains of 240-480 atoms (1.5-3 kDa)
    summing to chain_atoms, the first two carrying a lambda site of 4
    substituents x 8 atoms, Na+/Cl- pairs at 150 mM, TIP3P water filling the
    rest; PME 96^3 order 6 on the MFMA FFT.  The sites use the options of
    the reference's MSLD test decks: soft-core nonbonded terms and
    `removescaling bond urey angle impr` (only dihedrals scale with lambda),
    so a substituent at small lambda keeps its shape."""
    rng = np.random.default_rng(seed)
    spec = Spec(name="config3-precision-92224", box=tuple(float(b) for b in box))
    spec._pos = []
    _water_params(spec)
    grid = _Grid(spec.box, 3.0)
    sizes = rng.integers(240, 481, n_chains).astype(float)
    sizes = np.floor(sizes * chain_atoms / sizes.sum()).astype(int)
    sizes[-1] += chain_atoms - int(sizes.sum())
    sites = _chains(spec, rng, [int(x) for x in sizes], grid, [4, 4])
    assert spec.n == chain_atoms, (spec.n, chain_atoms)
    _angles_dihes(spec, _sites_block_of(spec, sites))
    _add_ions(spec, rng, n_ion_pairs, grid)
    _fill_water(spec, rng, n_water, grid, lattice=35)
    spec.pos = np.array(spec._pos) % np.array(spec.box)   # chain spheres may cross the faces
    _msld_sites(spec, sites, fnex=5.0)
    spec.soft_core = True
    spec.scale_terms = (0, 0, 0, 1, 0, 1)
    spec.rcut, spec.rswitch = 10.0, 9.0
    spec.beta = 1.0 / 3.2
    spec.grid, spec.order = (96, 96, 96), 6
    spec.fft_mode = 1
    spec.temperature = 298.0
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


def config5(replica=0, seed=0xB1ADE005, natoms=60000, chain_atoms=6000, n_chains=10, box=84.8):
    """One lambda-replica of the combined workload.  The structure (chains,
    sites, water) comes from `seed` alone, so every replica is the same box;
    the replica index seeds the initial theta, the per-block fixed biases
    (N(0, 1) kcal/mol) and the velocities.  Sites use the reference decks'
    options: soft core, only dihedrals lambda-scaled, each site's
    substituents' first atoms restrained together."""
    rng = np.random.default_rng(seed)
    spec = Spec(name="config5-combined-60000-r%d" % replica, box=(box,) * 3)
    spec._pos = []
    _water_params(spec)
    grid = _Grid(spec.box, 3.0)
    sizes = rng.integers(500, 701, n_chains).astype(float)
    sizes = np.floor(sizes * chain_atoms / sizes.sum()).astype(int)
    sizes[-1] += chain_atoms - int(sizes.sum())
    nsubs = [int(x) for x in rng.integers(4, 9, 4)]
    sites = _chains(spec, rng, [int(x) for x in sizes], grid, nsubs, rscale=26.0)
    assert spec.n == chain_atoms, (spec.n, chain_atoms)
    _angles_dihes(spec, _sites_block_of(spec, sites))
    _fill_water(spec, rng, (natoms - chain_atoms) // 3, grid)
    spec.pos = np.array(spec._pos) % np.array(spec.box)
    _msld_sites(spec, sites, fnex=5.0)
    spec.soft_core = True
    spec.scale_terms = (0, 0, 0, 1, 0, 1)
    spec.atom_restraints = [[subs[0] for subs in site] for site in sites]
    spec.rcut, spec.rswitch = 10.0, 9.0
    spec.beta = 1.0 / 3.2
    spec.grid, spec.order = (80, 80, 80), 6
    spec.fft_mode = 1
    spec.temperature = 298.0
    return replica_of(spec, replica, seed, new_theta=True)


def replica_of(base, replica, seed=0xB1ADE005, new_theta=False):
    """a lambda-replica of `base` (positions kept): fixed biases N(0, 1)
    kcal/mol and Maxwell velocities from seed + 1000 + replica, and with
    new_theta an initial theta uniform in [-1, 1] (otherwise the base's
    theta, so a relaxed base stays consistent with its lambdas)"""
    import copy
    spec = copy.copy(base)
    spec.name = base.name.rsplit("-r", 1)[0] + "-r%d" % replica
    rr = np.random.default_rng(seed + 1000 + replica)
    nb = spec.nblocks
    theta = [0.0] + [float(x) for x in rr.uniform(-1.0, 1.0, nb - 1)]
    if new_theta:
        spec.theta0, spec.theta_v = theta, None
    spec.fixed_bias = [0.0] + [float(x) for x in rr.normal(0.0, 1.0, nb - 1)]
    spec.vel = _maxwell(rr, [a[6] for a in spec.atoms], spec.temperature)
    return spec


def two_site_small(seed=0xB1ADE0C0, box=24.0, grid=24):
    """two 50-atom chains carrying lambda sites of 3 and 2 substituents,
    two Na+/Cl- pairs and water: the multi-site parity case"""
    rng = np.random.default_rng(seed)
    spec = Spec(name="two-site-small", box=(box,) * 3)
    spec._pos = []
    _water_params(spec)
    g = _Grid(spec.box, 3.0)
    sites = _chains(spec, rng, [50, 50], g, [3, 2])
    _angles_dihes(spec, _sites_block_of(spec, sites))
    _add_ions(spec, rng, 2, g)
    _fill_water(spec, rng, 330, g, lattice=9)
    spec.pos = np.array(spec._pos)
    _msld_sites(spec, sites, fnex=5.0)
    spec.scale_terms = (0, 0, 0, 1, 0, 1)   # `removescaling bond urey angle impr`
    # the first heavy atom of each substituent of a site held together
    spec.atom_restraints = [[subs[0] for subs in site] for site in sites]
    spec.rcut, spec.rswitch = 9.0, 8.0
    spec.beta = 0.34
    spec.grid, spec.order = (grid, grid, grid), 6
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


def config2(seed=0xB1ADE002, n_water=7023, n_poly=2489, box=62.23):
    """23,558-atom MSLD box: TIP3P waters + synthetic polymer with one site
    of 3 substituents x 8 atoms (the polymer count includes them)."""
    rng = np.random.default_rng(seed)
    spec = Spec(name="config2-msld-23558", box=(box,) * 3)
    spec._pos = []
    _water_params(spec)
    grid = _Grid(spec.box, 3.0)
    radius = 26.0 * (n_poly / 2489.0) ** (1 / 3)
    _, subs = _polymer(spec, rng, n_poly, np.array(spec.box) / 2, radius, grid, site_at=0.5)
    npoly = spec.n
    assert npoly == n_poly, (npoly, n_poly)
    block_of = np.zeros(spec.n, dtype=np.int32)
    for s, atoms in enumerate(subs):
        block_of[atoms] = s + 1
    _angles_dihes(spec, block_of)
    _fill_water(spec, rng, n_water, grid)
    spec.pos = np.array(spec._pos)
    _msld_site(spec, subs, fnex=5.0)
    spec.rcut, spec.rswitch = 10.0, 9.0
    spec.beta = 1.0 / 3.2
    spec.grid, spec.order = (64, 64, 64), 6
    spec.temperature = 298.0
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


WATER_DENSITY = 0.0978   # atoms per A^3 in a TIP3P box


def _ligand_clash(spec, block_of, dmin=1.7):
    """True if two atoms of the ligand built so far (every atom of `spec`)
    sit closer than dmin although they interact: not within two bonds of
    each other, and not in different blocks of the one lambda site
    (alternative substituents overlap by design)."""
    n = len(spec._pos)
    nbr = [set() for _ in range(n)]
    for a, b in spec.bonds:
        nbr[a].add(b)
        nbr[b].add(a)
    P = np.asarray(spec._pos, float)
    L = np.array(spec.box)
    for a in range(n):
        near = set(nbr[a])
        for b in nbr[a]:
            near |= nbr[b]
        d = P[a + 1:] - P[a]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=1)
        for b in np.nonzero(r < dmin)[0] + a + 1:
            b = int(b)
            if b in near:
                continue
            if block_of[a] and block_of[b] and block_of[a] != block_of[b]:
                continue
            return True
    return False


def batch_member(seed, natoms, nsubs, core=None):
    """One member of the batched workload: TIP3P water at box density plus
    a ligand (a 30-60-atom core, seeded) carrying one lambda site of
    `nsubs` substituents x 8 atoms; cutoff 10 A, switch 9-10 A, PME grid
    spacing <= 1 A (the size rule picks the grid), order 6, fnex 5.  A
    ligand whose substituents land on core atoms is rebuilt from the next
    derived seed (seed + attempt * 2^32)."""
    for attempt in range(64):
        spec = _batch_member_once(seed + (attempt << 32), natoms, nsubs, core)
        if spec is not None:
            return spec
    raise RuntimeError("batch member %d: no clash-free ligand in 64 seeds" % seed)


def _batch_member_once(seed, natoms, nsubs, core):
    rng = np.random.default_rng(seed)
    if core is None:
        core = int(rng.integers(30, 61))
    n_lig = core + 8 * nsubs
    n_water = (natoms - n_lig) // 3
    box = ((3 * n_water + n_lig) / WATER_DENSITY) ** (1 / 3)
    spec = Spec(name="batch-%d-%d" % (natoms, nsubs), box=(box,) * 3)
    spec._pos = []
    _water_params(spec)
    g = _Grid(spec.box, 3.0)
    radius = 6.0 * (n_lig / 60.0) ** (1 / 3) + 2.0
    _, subs = _polymer(spec, rng, n_lig, np.array(spec.box) / 2, radius, g, site_at=0.5, nsubs=nsubs,
                       typed_subs=True)
    block_of = np.zeros(spec.n, dtype=np.int32)
    for s, atoms in enumerate(subs):
        block_of[atoms] = s + 1
    if _ligand_clash(spec, block_of):
        return None
    _angles_dihes(spec, block_of)
    _fill_water(spec, rng, n_water, g)
    spec.pos = np.array(spec._pos)
    _msld_site(spec, subs, fnex=5.0)
    spec.rcut, spec.rswitch = 10.0, 9.0
    spec.beta = 1.0 / 3.2
    spec.grid, spec.grid_space, spec.order = (-1, -1, -1), 1.0, 6
    spec.temperature = 298.0
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


def batch_set(n=32, seed=0xB1ADE004, lo=3000, hi=25000):
    """(natoms, nsubs, seed) of the n members: atom counts uniform in
    [lo, hi], 2-6 substituents, seeded"""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(lo, hi + 1)), int(rng.integers(2, 7)), seed + 1 + k) for k in range(n)]


def msld_small(seed=0xB1ADE0A0, n_water=150, n_poly=60, box=20.0, grid=20):
    """A small MSLD system (water + short polymer + one 3-block site) whose
    oracle evaluation takes well under a second."""
    rng = np.random.default_rng(seed)
    spec = Spec(name="msld-small", box=(box,) * 3)
    spec._pos = []
    _water_params(spec)
    g = _Grid(spec.box, 3.0)
    _, subs = _polymer(spec, rng, n_poly, np.array(spec.box) / 2, 6.0, g, site_at=0.5)
    block_of = np.zeros(spec.n, dtype=np.int32)
    for s, atoms in enumerate(subs):
        block_of[atoms] = s + 1
    _angles_dihes(spec, block_of)
    _fill_water(spec, rng, n_water, g, lattice=7)
    spec.pos = np.array(spec._pos)
    _msld_site(spec, subs, fnex=5.0)
    spec.rcut, spec.rswitch = 9.0, 8.0
    spec.beta = 0.34
    spec.grid, spec.order = (grid, grid, grid), 6
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    return spec


def add_cmaps(spec, count=24, ngrid=24, seed=7, noise=0.05):
    """CMAP cross terms over `count` paths a-b-c-d-e of bonded atoms (phi on
    a-b-c-d, psi on b-c-d-e), substituent atoms first so that terms carry
    lambda scaling; one grid per distinct 8-type key, a smooth periodic
    surface of a few Fourier modes plus grid noise (kcal/mol)"""
    rng = np.random.default_rng(seed)
    nbr = {}
    for a, b in spec.bonds:
        nbr.setdefault(a, []).append(b)
        nbr.setdefault(b, []).append(a)
    ab = spec.atom_block if spec.atom_block is not None else np.zeros(spec.n, int)
    starts = sorted(nbr, key=lambda a: (-int(ab[a]), a))
    paths, used = [], set()
    for a in starts:
        if len(paths) >= count:
            break
        walk = [[a]]
        for _ in range(4):
            walk = [w + [x] for w in walk for x in nbr[w[-1]] if x not in w]
        for p in walk:
            if tuple(p) not in used and tuple(p[::-1]) not in used:
                used.add(tuple(p))
                paths.append(p)
                break
    ang = np.arange(ngrid) * 2 * np.pi / ngrid - np.pi
    P, Q = np.meshgrid(ang, ang, indexing="ij")
    for a, b, c, d, e in paths:
        q = (a, b, c, d, b, c, d, e)
        key = tuple(spec.atoms[x][4] for x in q)
        if key not in spec.cmap_par:
            k = rng.uniform(0.2, 0.8, 4)
            spec.cmap_par[key] = (k[0] * np.cos(P + 0.3) + k[1] * np.sin(2 * Q) + k[2] * np.cos(P - Q)
                                  + k[3] * np.sin(P + 2 * Q) + noise * rng.standard_normal((ngrid, ngrid)))
        spec.cmaps.append(q)
    return spec


def ensemble_probe(seed=0xB1ADE0E0, nsites=6, nsubs=3, npairs=24, nfree=12, box=30.0, k_pair=3510.0,
                   bias=(0.0, 0.6, -0.4), var_bias=None, theta_inde=None, msld_gamma_ps=5.0, gamma_ps=5.0):
    """A system whose equilibrium averages are known in closed form or by
    quadrature, for checking the integrator's sampling:
      * `nsites` lambda sites of `nsubs` substituent blocks, each block one
        non-interacting dummy atom, with fixed biases `bias` per substituent
        (and optional variable and independent theta biases): the lambda
        averages are ratios of integrals over theta in [-pi, pi)^nsubs;
      * `npairs` environment atom pairs held together by atom restraints of
        constant k_pair (E = k/4 |x1 - x2|^2 per pair, omega^2 = k/m): each
        Cartesian component of the pair separation is a harmonic oscillator
        with <r^2> = 2 kT / k, which the VRORV (BAOAB) splitting samples
        exactly at any stable step;
      * `nfree` free environment atoms and the pairs' centres of mass:
        Ornstein-Uhlenbeck velocities with <v^2> = kT / m per component.
    No charges, no Lennard-Jones (epsilon 0), force-switched electrostatics,
    no constraints."""
    rng = np.random.default_rng(seed)
    spec = Spec(name="ensemble-probe", box=(box,) * 3)
    spec.types["DUM"] = 12.0
    spec.nb_par["DUM"] = (0.0, 1.0, 0.0, 1.0, 1.0, 0)
    spec._pos = []

    def atom(p):
        spec.atoms.append(("PRB", str(len(spec.atoms) + 1), "DUM", "D%d" % len(spec.atoms), "DUM", 0.0, 12.0))
        spec._pos.append(np.asarray(p, float))
        return len(spec.atoms) - 1

    pairs = []
    for _ in range(npairs):
        c = rng.uniform(0, box, 3)
        u = rng.normal(size=3)
        u *= 0.02 / np.linalg.norm(u)
        pairs.append([atom(c + u), atom(c - u)])
    for _ in range(nfree):
        atom(rng.uniform(0, box, 3))
    sites = []
    for _ in range(nsites):
        sites.append([[atom(rng.uniform(0, box, 3))] for _ in range(nsubs)])
    spec.pos = np.array(spec._pos)
    _msld_sites(spec, sites, fnex=5.5)
    spec.fixed_bias = [0.0] + list(bias[:nsubs]) * nsites
    spec.theta0 = [0.0] + list(rng.uniform(-np.pi, np.pi, spec.nblocks - 1))
    spec.theta_v = [0.0] * spec.nblocks
    spec.var_bias = list(var_bias or [])
    spec.theta_inde = theta_inde
    spec.atom_restraints = pairs
    spec.k_restraint = k_pair
    spec.shake = False
    spec.elec_method, spec.vdw_method = 0, 0
    spec.rcut, spec.rswitch = 10.0, 9.0
    spec.temperature = 298.0
    spec.gamma = gamma_ps / PICOSECOND
    spec.msld_gamma = msld_gamma_ps / PICOSECOND
    spec.vel = _maxwell(rng, [a[6] for a in spec.atoms], spec.temperature)
    spec.seed = seed
    return spec


def term_variety(spec):
    """A copy of `spec` whose parameters exercise the term forms the default
    synthetic systems leave out: Urey-Bradley terms on every polymer angle
    type (kub 35 kcal/mol/A^2, s0 2.45 A), periodic impropers (k 1.5, n 2,
    90 degrees) in place of the harmonic ones, NBFIX overrides for two type
    pairs and the geometric combination rule for two types."""
    import copy
    sp = copy.deepcopy(spec)
    for key, (k, t0, kub, s0) in list(sp.angle_par.items()):
        if "OT" not in key:
            sp.angle_par[key] = (k, t0, 35.0, 2.45)
    for key in list(sp.impr_par):
        sp.impr_par[key] = (1.5, 2, 90.0)
    if "CT" in sp.nb_par and "OT" in sp.nb_par:
        sp.nbfix_par[("OT", "CT")] = (-0.12, 3.55, -0.08, 3.40)
    if "NT" in sp.nb_par and "HT" in sp.nb_par:
        sp.nbfix_par[("NT", "HT")] = (-0.05, 2.10, -0.05, 2.00)
    for t in ("CT", "NT"):
        if t in sp.nb_par:
            e, r, e14, r14, f, _ = sp.nb_par[t]
            sp.nb_par[t] = (e, r, e14, r14, f, 1)
    return sp


def add_restraints(spec, seed=11):
    """A copy of `spec` with a few terms of every restraint form
    (restrain/restrain.cu) on its polymer atoms: NOE pairs inside, below and
    beyond their flat band (one past the switch to the power-law tail) and
    one to a fixed point; harmonic position restraints of powers 2 and 4;
    bond, angle and dihedral restraints (periodic, harmonic and flat-bottom),
    some scaled by a substituent block's lambda; two RESD terms."""
    import copy
    sp = copy.deepcopy(spec)
    rng = np.random.default_rng(seed)
    P = np.asarray(sp.pos, float)
    L = np.array(sp.box, float)
    poly = [i for i, a in enumerate(sp.atoms) if a[0] in ("POLY", "LIG")]
    pick = lambda k: [int(x) for x in rng.choice(poly, size=k, replace=False)]
    dist = lambda i, j: float(np.linalg.norm((P[i] - P[j]) - L * np.round((P[i] - P[j]) / L)))
    blk = [int(b) for b in (sp.atom_block if sp.atom_block is not None else np.zeros(sp.n, int))]
    scaled = next((b for b in blk if b), 0)
    r = {"noe": [], "harm": [], "borest": [], "anrest": [], "direst": [], "resd": []}
    for k in range(4):
        i, j = pick(2)
        d = dist(i, j)
        bands = [(d - 1.0, d + 1.0), (d + 0.5, d + 2.0), (d - 2.0, d - 0.6), (d - 2.5, d - 1.5)]
        rmin, rmax = bands[k]
        rsw = 0.4 if k == 3 else 0.0
        r["noe"].append((i, j, rmin, 5.0, rmax, 3.0, 0.8, rsw, 1.0, 0.0, 0.0, 0.0, False))
    i = pick(1)[0]
    r["noe"].append((i, i, 0.0, 0.0, 1.0, 2.0, 0.0, 0.0, 1.0, *(P[i] + rng.uniform(1.0, 2.0, 3)), True))
    for n in (2.0, 4.0):
        i = pick(1)[0]
        r["harm"].append((i, 1.5, *(P[i] + rng.uniform(-0.8, 0.8, 3)), n))
    for k in range(3):
        i, j = pick(2)
        r["borest"].append((i, j, 4.0, dist(i, j) + rng.uniform(-0.5, 0.5), scaled if k == 0 else 0))
    for k in range(3):
        a = sp.angles[int(rng.integers(len(sp.angles)))]
        r["anrest"].append((a[0], a[1], a[2], 20.0, 100.0 + 10 * k, scaled if k == 1 else 0))
    for k, (n, w) in enumerate(((2, 0.0), (0, 0.0), (0, 15.0), (3, 0.0))):
        q = sp.dihes[int(rng.integers(len(sp.dihes)))]
        r["direst"].append((q[0], q[1], q[2], q[3], 3.0, n, 40.0 + 30 * k, w, scaled if k == 2 else 0))
    for k in range(2):
        i1, i2, j1, j2 = pick(4)
        r["resd"].append((i1, i2, j1, j2, 1.0, -1.0 if k else 0.5, 0.3, 6.0))
    sp.restraints = r
    return sp
