/*
 * blade_mi355x.h - kernel-layer C ABI of the MI355X implementation of
 * BLaDE's per-step force and integration path.
 *
 * Two layers are exported from libblade_mi355x.so:
 *
 *   1. The reference-compatible `blade_*` library API (include/blade.h),
 *      the one CHARMM binds (reference: system/system.h:86-91,
 *      run/run.h:146-158, msld/msld.h:147-165, system/structure.h:111-133,
 *      system/coordinates.h:57-63, system/parameters.h:202-230).
 *
 *   2. This header: the per-term kernel layer that the reference reaches
 *      through `getforce_X(System*,bool)` and friends.  Each entry point
 *      takes an opaque system handle (device buffers owned by the
 *      library) and a flag word.  No torch types, no C++ types.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative BM_E* code on failure; the library
 *     never calls exit();
 *   - forces are dU/dx (reference sign convention, update/update.cu:52),
 *     accumulated as signed 64-bit fixed point with BM_FIXED_SCALE counts
 *     per kcal/mol/A, so sums are independent of accumulation order;
 *   - energies and dU/dlambda are 64-bit fixed point with the same scale;
 *   - positions, velocities and theta are FP64, stored structure-of-arrays:
 *     x[0..N) y[N..2N) z[2N..3N) theta[3N..3N+nL);
 *   - units are AKMA (kcal/mol, A, e, amu; time in AKMA units, one
 *     picosecond = 20.454828284385908).
 */
#ifndef BLADE_MI355X_H
#define BLADE_MI355X_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_LAYOUT_VERSION 2

/* error codes */
#define BM_OK 0
#define BM_EINVAL (-1)     /* invalid argument or inconsistent sizes          */
#define BM_EHIP (-2)       /* a HIP runtime call failed                        */
#define BM_ENAN (-3)       /* NaN/Inf detected in the integrator               */
#define BM_EOVERFLOW (-4)  /* a fixed-capacity buffer would overflow           */
#define BM_ENOTSUP (-5)    /* feature outside the supported set                */
#define BM_ENOGPU (-6)     /* no HIP device available                          */

/* 64-bit fixed point: 2^32 counts per unit */
#define BM_FIXED_SCALE 4294967296.0

/* energy terms, same order as the reference EETerm (system/potential.h:20-44) */
enum bm_eterm {
  BM_EBOND = 0, BM_EANGLE, BM_EUREY, BM_EDIHE, BM_EIMPR, BM_ECMAP, BM_ENB14,
  BM_ENBDIRECT, BM_ENBRECIP, BM_ENBRECIPSELF, BM_ENBRECIPEXCL, BM_ELAMBDA,
  BM_ETHETA, BM_ECATS, BM_ENOE, BM_EHARMONIC, BM_EMMFP, BM_ERESD, BM_EMLP,
  BM_EBIAS, BM_EPOTENTIAL, BM_EKINETIC, BM_ETOTAL, BM_NTERMS
};

/* nonbonded methods (run/run.h EVdw / EElec) */
enum bm_vdw { BM_VFSWITCH = 0, BM_VSWITCH = 1, BM_VSHIFT = 2 };
enum bm_elec { BM_FSWITCH = 0, BM_PME = 1, BM_FSHIFT = 2 };

/* flags for force entry points */
#define BM_CALC_ENERGY 1u   /* accumulate energies                          */
#define BM_CALC_ALCH 2u     /* system carries lambda sites                  */
#define BM_SOFTCORE 4u      /* soft-core nonbonded (msld.h useSoftCore)     */
#define BM_SOFTCORE14 8u    /* soft-core 1-4 pairs (msld.h useSoftCore14)   */

/*
 * Derived topology: the term tables the reference's Potential::initialize
 * builds (system/potential.cxx:525-1837), as host arrays.  siteBlock codes
 * are (site<<16)|block exactly as merge_site_block (msld/msld.cu:470-476).
 */
typedef struct bm_topology {
  int32_t n_atoms;
  const double *charge;      /* [N]   e                                  */
  const double *mass;        /* [N]   amu                                */
  const int32_t *type_idx;   /* [N]   vdW type                           */
  const int32_t *site_block; /* [N]   siteBlock code of the atom         */
  int32_t n_types;
  const double *vdw_c12;     /* [T*T] eps*Rmin^12                        */
  const double *vdw_c6;      /* [T*T] 2*eps*Rmin^6                       */

  int32_t n_bonds, n_bonds12;      /* 1-2 bonds first, then Urey-Bradley */
  const int32_t *bond_idx;         /* [2n] */
  const int32_t *bond_sb;          /* [2n] siteBlock codes */
  const double *bond_kb, *bond_b0; /* E = kb/2 (r-b0)^2 */

  int32_t n_angles;
  const int32_t *angle_idx;        /* [3n] */
  const int32_t *angle_sb;         /* [2n] */
  const double *angle_k, *angle_t0;

  int32_t n_dihes;
  const int32_t *dihe_idx;         /* [4n] */
  const int32_t *dihe_sb;          /* [2n] */
  const double *dihe_k, *dihe_d0;  /* E = k (1+cos(n phi - d0)) */
  const int32_t *dihe_n;

  int32_t n_imprs;
  const int32_t *impr_idx;         /* [4n] */
  const int32_t *impr_sb;          /* [2n] */
  const double *impr_k, *impr_d0;  /* n>0: periodic; n==0: k/2 dphi^2 */
  const int32_t *impr_n;

  int32_t n_nb14;
  const int32_t *nb14_idx;         /* [2n] */
  const int32_t *nb14_sb;          /* [2n] */
  const double *nb14_qq, *nb14_c12, *nb14_c6, *nb14_e14fac;

  int32_t n_nbex;                  /* PME exclusion corrections */
  const int32_t *nbex_idx;         /* [2n] */
  const int32_t *nbex_sb;          /* [2n] */
  const double *nbex_qq;

  int32_t n_excl;                  /* unordered excluded pairs, i<j */
  const int32_t *excl_idx;         /* [2n] */

  int32_t n_tri, n_br1, n_br2, n_br3;  /* holonomic constraint groups */
  const int32_t *tri_idx;  const double *tri_b0;  /* [3n],[3n] b0: 01,02,12 */
  const int32_t *br1_idx;  const double *br1_b0;  /* [2n],[n]  */
  const int32_t *br2_idx;  const double *br2_b0;  /* [3n],[2n] */
  const int32_t *br3_idx;  const double *br3_b0;  /* [4n],[3n] */

  /* MSLD (block 0 is the environment, site 0) */
  int32_t n_blocks, n_sites;
  const int32_t *block_site;       /* [nL] site of each block            */
  const int32_t *site_bound;       /* [nSites+1] first block of each site */
  const double *theta_mass;        /* [nL] amu A^2 (0 => not integrated) */
  const int32_t *block_fixed;      /* [nL] 1 => lambda fixed at theta    */
  const double *fixed_bias;        /* [nL] lambdaBias                    */
  const double *theta_friction;    /* [nL] per-block gamma (AKMA); NULL => msld gamma */
  int32_t n_varbias;
  const int32_t *vb_i, *vb_j, *vb_type, *vb_n;   /* [n] (msld.h:13-19) */
  const double *vb_l0, *vb_k;

  /* MSLD atom restraints (msld.h:73-79, msld.cu:1328-1364): each group's
     atoms are pulled to their centroid, E = k/2 sum |x_i - centroid|^2,
     not lambda-scaled (energy term BM_ECATS) */
  int32_t n_cats;                  /* groups                              */
  const int32_t *cats_bound;       /* [n+1] first member of each group    */
  const int32_t *cats_idx;         /* [cats_bound[n]] member atoms        */
  double cats_k;                   /* kRestraint, kcal/mol/A^2            */

  /* MSLD charge restraints on Q = sum_b lambda_b q_b (msld.cu:1400-1551);
     energy term BM_EBIAS.  cr_k1 (PME only): E = -(2 pi/3) kE k1 Q / V;
     cr_k2 (PME only, 0 or 1): E = -(pi/2) kE k2 Q^2 / (V beta^2);
     cr_k3: E = k3/2 dQ^2, dQ = distance of Q - q0 from the band [-w, w] */
  const double *block_charge;      /* [nL] q_b: the sum of the charges of the
                                      block's atoms (Msld::initialize,
                                      msld.cu:684-690)                  */
  double cr_k1, cr_k2, cr_k3, cr_q0, cr_w;

  /* CMAP cross terms (bonded/bonded.cu:531-720): E = lambda-scaled bicubic
     interpolation of a periodic grid in (phi, psi), phi the dihedral of
     atoms 0-3 and psi of atoms 4-7; grid point (i, j) at
     (-180 + i 360/ngrid, -180 + j 360/ngrid) degrees, kcal/mol */
  int32_t n_cmaps;
  const int32_t *cmap_idx;         /* [8n] */
  const int32_t *cmap_sb;          /* [3n] */
  const int32_t *cmap_type;        /* [n] grid of each term */
  int32_t n_cmap_types;
  const int32_t *cmap_ngrid;       /* [types] points per 360 degrees */
  const int32_t *cmap_off;         /* [types] first value in cmap_grid */
  const double *cmap_grid;         /* row-major [ngrid(phi)][ngrid(psi)] */

  /* MSLD theta biases (msld/msld.cu:1155-1274; setters msld.cu:1675-1705),
     energy term BM_ETHETA, forces on theta directly:
       collective, per site s (k, n): with h_j = (1 + sin theta_j)/2 and
         l_j = (1 - sin theta_j)/2 over the site's nb blocks,
         A = sum_j h_j^n - 1, B = sum_j l_j^n - (nb - 1),
         E = k/2 (A^2 + B^2) and dU/dtheta_j = k n (A h_j^(n-1) -
         B l_j^(n-1)) cos theta_j, the reference's force (twice the theta
         derivative of its reported energy);
       independent, per block b of site s (k): with s_b = l_b,
         E = k (1 - s_b^4), dU/dtheta_b = 2 k s_b^3 cos theta_b.
     NULL => none. */
  const double *tcoll_k, *tcoll_n; /* [nSites] */
  const double *tinde_k;           /* [nSites] */
  /* piecewise lambda constraint (msld.cu:785-861, setter msld.cu:1706):
     lambda_i = max(theta_i - t_s, 0)^3 with t_s solved per site by Newton
     so the site's lambdas sum to 1; every block but the environment feels
     a flat-bottom harmonic pw_k/2 d^2 beyond theta = +-nb*pw_width
     (msld.cu:1189-1214); with |edge_k| > 1e-4 a bias edge_k/2 (N - edge_n)^2
     on the smooth substituent count N = sum_j 1/(1 + exp(-edge_alpha
     (theta_j - t_s + edge_phi))) above edge_n (msld.cu:1216-1252,
     blade_set_msld_thetaedgebias).  Fixed blocks are not allowed with it. */
  int32_t piecewise;
  double pw_width, pw_k;
  double edge_k, edge_n, edge_alpha, edge_phi;
  /* msld fix flag (msld.cu:913-935): every lambda held at its initial
     theta value, no theta force from dU/dlambda */
  int32_t fix_lambda;

  /* Restraints (restrain/restrain.cu:15-598; setters
     system/structure.cxx:1093-1196), minimum-image geometry:
     NOE, energy BM_ENOE: distance r of atoms i-j, or of atom i from the
       point (c0x, c0y, c0z) when noe_idx[2k+1] < 0; below rmin
       kmin/2 (r - rmin)^2, above rmax kmax/2 (r - rmax)^2, or with
       rswitch > 0 beyond rmax + rswitch kmax (a + b d^-nswitch + rpeak d),
       d = r - rmax, continuous with the harmonic piece;
     harmonic position, energy BM_EHARMONIC: E = k |x_i - x0|^n;
     bond / angle / dihedral restraints, energy BM_EMMFP, each scaled by
       lambda of its block (0: unscaled): kr/2 (r - r0)^2, kt/2 (t - t0)^2,
       and for the dihedral kphi (1 - cos(n (phi - phi0))) when n > 0, else
       kphi/2 d^2 with d the distance of phi - phi0 (wrapped) from the
       flat band [-width, width];
     RESD, energy BM_ERESD: kdist/2 (ci r1 + cj r2 - rdist)^2, r1 = |i1 -
       i2|, r2 = |j1 - j2|.  Angles in radians, AKMA units. */
  int32_t n_noe;
  const int32_t *noe_idx;          /* [2n]                                 */
  const double *noe_par;           /* [10n] rmin kmin rmax kmax rpeak rswitch nswitch c0x c0y c0z */
  int32_t n_harm;
  const int32_t *harm_idx;         /* [n]                                  */
  const double *harm_par;          /* [5n] k x0 y0 z0 n                    */
  int32_t n_borest;
  const int32_t *borest_idx, *borest_block;   /* [2n], [n]                 */
  const double *borest_par;        /* [2n] kr r0                           */
  int32_t n_anrest;
  const int32_t *anrest_idx, *anrest_block;   /* [3n], [n]                 */
  const double *anrest_par;        /* [2n] kt t0                           */
  int32_t n_direst;
  const int32_t *direst_idx, *direst_block, *direst_n;  /* [4n], [n], [n]  */
  const double *direst_par;        /* [3n] kphi phi0 width                 */
  int32_t n_resd;
  const int32_t *resd_idx;         /* [4n] i1 i2 j1 j2                     */
  const double *resd_par;          /* [4n] ci cj rdist kdist               */
  /* MSLD soft bonds (msld/msld.cu:404-428,478-495; bonded/bonded.cu:81-107,
     239-256, 408-425; setter msld.cu:1714): a bond, Urey-Bradley, angle,
     dihedral or improper term holding both atoms of a soft bond is soft.
     It keeps its blocks' lambda scaling where removescaling names its
     type, and with P = lambda_0 lambda_1 a soft bond or Urey-Bradley term
     is E = P^p E0 / (1 + (1 - P^p) alpha (r - b0)^2) (runparams
     soft_bond_alpha, soft_bond_power), a soft angle, dihedral or improper
     E = P^q E0 (runparams soft_bond_exp).  NULL when no term is soft. */
  const int32_t *bond_soft, *angle_soft, *dihe_soft, *impr_soft;   /* [n] 0 or 1 */
  /* Virtual sites / lone pairs (holonomic/virtual.cu:11-254; setters
     system/structure.cxx:1059-1085, PSF NUMLP section structure.cxx:
     734-800): massless atoms placed from their hosts before each force
     evaluation; the force on a site is carried to its hosts, so the
     energy's gradient with respect to the hosts is exact, and the site's
     own force entry is left zero.  Two hosts (site v, hosts h0 h1):
     x_v = x_h0 + (dist / |d| + scale) d with d = x_h0 - x_h1 (minimum
     image).  Three hosts (v, h0 h1 h2): with b = x_h2 - x_h1 (h1 moved to
     the midpoint of h1 and h2 when dist < 0), a = unit(x_h1 - x_h0),
     c = unit(b x a) and b' = a x c, x_v = x_h0 + |dist| (cos(theta) a +
     sin(theta) (cos(phi) b' + sin(phi) c)).  A site is excluded from its
     first host and from that host's bond partners. */
  int32_t n_virt2;
  const int32_t *virt2_idx;        /* [3n] v h0 h1 */
  const double *virt2_par;         /* [2n] dist (A) scale */
  int32_t n_virt3;
  const int32_t *virt3_idx;        /* [4n] v h0 h1 h2 */
  const double *virt3_par;         /* [3n] dist (A) theta phi (radians) */
} bm_topology;

/* run parameters (run/run.cu:27-110 defaults) */
typedef struct bm_runparams {
  double rcut, rswitch;            /* A */
  double beta_ewald;               /* 1/A */
  int32_t vdw_method, elec_method; /* bm_vdw, bm_elec */
  int32_t grid[3];                 /* PME grid */
  int32_t order;                   /* B-spline order (4, 6, 8) */
  double dt;                       /* AKMA time */
  double temperature;              /* K */
  double gamma;                    /* atom friction, 1/AKMA time */
  double msld_gamma;               /* theta friction, 1/AKMA time */
  double shake_tol;
  double fnex;
  int32_t soft_core, soft_core14;
  double soft_bond_exp;            /* softNotBondExponent q of soft angles, dihedrals, impropers */
  double box[3];                   /* orthorhombic box edges, A */
  int32_t freq_domdec;             /* list rebuild cadence (steps) */
  double cull_pad;                 /* list padding, A */
  uint64_t seed;                   /* Philox key */
  int32_t fft_mode;                /* PME FFT: 0 FP32 Stockham; 1 DFT contractions on
                                      MFMA in split-BF16x3 (every grid edge a multiple
                                      of 16 from 16 to 128, else BM_ENOTSUP) */
  int32_t tile_chunk;              /* directed tiles per nonbonded work unit (one
                                      wave), 1..16; 0 => 16 */
  double soft_bond_alpha;          /* 1/softBondRadius^2, A^-2 */
  double soft_bond_power;          /* softBondExponent p of soft bonds */
} bm_runparams;

/* ---------------------------------------------------------------------
 * Context: one per device and host thread (bm_ctx owns a System image).
 * ------------------------------------------------------------------- */
typedef struct bm_system bm_system;

/* Create a system on `device` from a derived topology, run parameters and
 * initial FP64 positions/velocities (SoA, [3N]) and theta state ([nL]).
 * Replaces the allocation half of Potential::initialize
 * (system/potential.cxx:525-1837), State::State (system/state.cxx:20-153)
 * and Domdec::initialize (domdec/domdec.cu:83-199). */
int bm_system_create(int device, const bm_topology *top, const bm_runparams *rp,
                     const double *pos, const double *vel,
                     const double *theta, const double *theta_vel,
                     bm_system **out);
int bm_system_destroy(bm_system *sys);

/* Domdec::update_domdec(System*,bool) (domdec/domdec.cu:213-222):
 * rebuild = 1 wraps positions, sorts atoms into blocks, builds the tile
 * list and exclusion masks; rebuild = 0 only repacks block positions. */
int bm_domdec_update(bm_system *sys, int rebuild);

/* Potential::calc_force(long step,System*) (system/potential.cxx:1851-1941):
 * all force terms at the current positions; flags = BM_CALC_ENERGY. */
int bm_calc_force(bm_system *sys, long step, unsigned flags);

/* Per-term entry points (nbdirect/nbdirect.h:7-8, nbrecip/nbrecip.h:7-8,
 * bonded/bonded.h:7-11, bonded/pair.h:7-8, restrain/restrain.h:7-12): each
 * adds its term's forces into the fixed-point force array, its dU/dlambda
 * and (flags & BM_CALC_ENERGY) its energy; the caller zeroes the
 * accumulators first (bm_reset_force).  Calling every term once gives the
 * same force array as bm_calc_force, bit for bit. */
int bm_reset_force(bm_system *sys, unsigned flags);
int bm_getforce_nbdirect(bm_system *sys, unsigned flags);
int bm_getforce_ewaldself(bm_system *sys, unsigned flags);  /* nbrecip.cu:19-77 */
int bm_getforce_ewald(bm_system *sys, unsigned flags);
int bm_getforce_bonded(bm_system *sys, unsigned flags);   /* bond,urey,angle,dihe,impr,cmap */
int bm_getforce_nb14(bm_system *sys, unsigned flags);
int bm_getforce_nbex(bm_system *sys, unsigned flags);
/* fixed, variable and theta biases, charge and atom restraints (msld.cu:1000-1551) */
int bm_getforce_msldbias(bm_system *sys, unsigned flags);
/* NOE, harmonic, bond/angle/dihedral restraint and RESD terms (restrain.cu:15-598) */
int bm_getforce_restrain(bm_system *sys, unsigned flags);

/* State::update(int step,System*) (update/update.cu:233-305): one VRORV
 * step including holonomic constraints and theta->lambda. */
int bm_update(bm_system *sys, long step, int energy_step);

/* holonomic_position / holonomic_velocity (holonomic/holonomic.h:14-15):
 * every constraint group solved from the backup positions the update keeps
 * (position form, which also sets v = dx/(dt/2)), or its velocities
 * projected onto the constraints (velocity form). */
int bm_holonomic_position(bm_system *sys);
int bm_holonomic_velocity(bm_system *sys);

/* holonomic_rectify (holonomic/rectify.cu:190-197), run by
 * Run::dynamics_initialize (run/run.cu:606) before the first step: each
 * constraint group's atoms other than the first put at exactly their
 * constrained distances from the first (a branch atom along its present
 * direction; a triangle's third atom in the plane of the three).  The
 * blade_* layer calls it once, when it creates the device system from
 * supplied coordinates, and not for a state read from a checkpoint. */
int bm_holonomic_rectify(bm_system *sys);

/* Msld::calc_lambda_from_theta / calc_thetaForce_from_lambdaForce
 * (msld/msld.h:134-136); the chain rule writes dU/dtheta (with the theta
 * biases of the last force evaluation) into the theta tail of the force
 * array */
int bm_msld_theta_to_lambda(bm_system *sys);
int bm_msld_chain_rule(bm_system *sys);

/* Run::dynamics step loop (run/run.cu:545-570) for nsteps steps starting
 * at `step0`; energies are computed on steps with step%freq_nrg==0.
 * use_graph = 1 replays captured hipGraphs, one per (rebuild, energy)
 * step variant. */
int bm_dynamics(bm_system *sys, long step0, long nsteps, int freq_nrg, int use_graph);
/* Steps run on the device without host synchronisation; the call waits for
   them at its end.  A capacity overflow (tile list, PME brick bins) halts
   the state updates on the device at the step that raised it; the call
   then grows that buffer x1.5 and runs again from that step, as the
   reference's cull retries (domdec/cull.cu:337-375, up to 12x the first
   size, domdec.cu:129-130), so the trajectory is the one an ample
   capacity gives.  A non-finite position or velocity halts the state the
   same way and the call returns BM_ENAN (bm_get_nan_flag names the first
   atom or DOF); a column beyond its capacity returns BM_EOVERFLOW.
   Members of a batch are not regrown: bm_batch_sync reports the overflow. */

/* Capture the step graphs of all four (rebuild, energy) variants now, so
   that no later bm_dynamics call captures one (a timed run starts with
   every graph instantiated).  No counterpart in the reference. */
int bm_capture_graphs(bm_system *sys);
/* Set the list capacities (directed tiles; spline records per PME brick;
   a value <= 0 keeps that one).  The list is emptied: the next step must
   rebuild it (a step s with s % freq_domdec == 0).  Tests use it to start
   from capacities too small for the system.  No counterpart in the
   reference. */
int bm_set_list_capacity(bm_system *sys, long long max_tiles, int brick_atoms);
int bm_get_list_capacity(bm_system *sys, long long *max_tiles, int *brick_atoms);

/* ---- batches: independent systems advanced together.  One captured
 * graph per (rebuild, energy) step variant advances every member: on
 * non-energy steps one launch per kernel type covers all members (members
 * are grouped only by spline order, PME grid size and nonbonded methods);
 * list rebuilds and energy steps run each member's kernels as parallel
 * branches.  Members stay bit-identical to the same systems run alone.
 * Members must live on one device and share the list rebuild cadence; the
 * device copies of their state descriptors are refreshed at each
 * bm_batch_dynamics call.  No counterpart in the reference (it runs one
 * system per GPU and process). */
typedef struct bm_batch bm_batch;
int bm_batch_create(int n, bm_system *const *systems, bm_batch **out);
int bm_batch_dynamics(bm_batch *batch, long step0, long nsteps, int freq_nrg);
/* as bm_capture_graphs, for the batch step graphs */
int bm_batch_capture_graphs(bm_batch *batch);
/* synchronise every member; BM_ENAN / BM_EOVERFLOW as bm_sync reports them */
int bm_batch_sync(bm_batch *batch);
/* Replica exchange between neighbouring members (update/rex.cxx:15-125,
 * with the MPI ranks replaced by the members of one batch): members k and
 * k+1 pair up for k even (parity 0) or odd (parity 1); each pair's
 * positions and theta are swapped on the device and
 *   dW = (E_k(x_k+1) - E_k(x_k)) / kT_k + (E_k+1(x_k) - E_k+1(x_k+1)) / kT_k+1
 * from potential energies evaluated before and after the swap; the swap is
 * kept with probability min(1, exp(-dW)) (uniform from Philox keyed by
 * (seed, step, k)), with the velocities swapped too and replica[k],
 * replica[k+1] exchanged; otherwise the positions are swapped back.  Each
 * member's list is rebuilt afterwards.  accepted[k] and dW[k] are written
 * for the lower member of each pair (either may be NULL).  Members need
 * equal atom and block counts.  No counterpart in the reference's
 * single-system library (its exchange runs over MPI, one rank per system). */
int bm_batch_exchange(bm_batch *batch, long step, int parity, uint64_t seed,
                      int32_t *replica, int32_t *accepted, double *dW);
/* bm_batch_exchange without any device allocation: the same exchange and
 * results, with the swap staging in a static device array of 2^21 doubles
 * (16 MiB per device, kept until the process ends) instead of memory
 * allocated per call.  No counterpart in the reference; it is for callers
 * that cannot allocate (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Members of more than 2^21 degrees of
 * freedom (3N + nL) are rejected with BM_EINVAL.  Calls that overlap on one
 * device share that device's array and race. */
int bm_batch_exchange_noalloc(bm_batch *batch, long step, int parity, uint64_t seed,
                              int32_t *replica, int32_t *accepted, double *dW);
int bm_batch_destroy(bm_batch *batch);

/* synchronise the system's stream */
int bm_sync(bm_system *sys);

/* ---- host readback (State::recv_* equivalents, system/state.cxx:330-365) */
int bm_get_positions(bm_system *sys, double *pos /*[3N]*/, double *theta /*[nL]*/);
int bm_set_positions(bm_system *sys, const double *pos, const double *theta);

/* ---- MC barostat volume moves (update/pressure.cu:63-118) ------------
 * bm_volume_move backs up the positions, scales every atom by scale[3]
 * (per axis), puts each constraint group's other atoms back at its scaled
 * first atom plus their old minimum-image offsets
 * (holonomic/rectify.cu:200-320) and scales the box; bm_volume_restore
 * puts the backed-up positions and the given box back (a rejected move).
 * Either invalidates the list: call bm_domdec_update(sys, 1) before the
 * next force evaluation. */
int bm_volume_move(bm_system *sys, const double *scale);
/* a uniform in (0, 1) from Philox4x32-10 keyed by (seed, step, k): the
   draws of the runtime's Monte Carlo decisions (barostat, replica exchange) */
double bm_uniform(uint64_t seed, long step, uint32_t k);
int bm_volume_restore(bm_system *sys, const double *box);
int bm_get_box(bm_system *sys, double *box);
int bm_get_velocities(bm_system *sys, double *vel /*[3N]*/, double *theta_vel /*[nL]*/);
int bm_set_velocities(bm_system *sys, const double *vel /*[3N] or NULL*/, const double *theta_vel /*[nL] or NULL*/);
int bm_get_forces(bm_system *sys, double *force /*[3N]*/, double *theta_force /*[nL]*/);
int bm_get_dudl(bm_system *sys, double *dudl /*[nL]*/);
int bm_get_lambda(bm_system *sys, double *lambda /*[nL]*/);
int bm_get_energy(bm_system *sys, double *energy /*[BM_NTERMS]*/);
int bm_get_nan_flag(bm_system *sys, int *flag);

/* ---- domdec readback for bit-exact list checks.  Tiles are directed:
   every (I,J) whose rebuild boxes lie within rcut+pad, both orders and
   (I,I), sorted by I then J; mask word i of tile (I,J) has bit j set when
   slot i of block I and slot j of block J are an excluded pair. */
int bm_get_domdec_sizes(bm_system *sys, int *n_blocks, int *n_tiles, int *n_excl_tiles);
int bm_get_domdec(bm_system *sys, int *local_to_global /*[32*nb], -1 = empty*/,
                  int *block_count /*[nb]*/, int *tile_i /*[nt]*/, int *tile_j /*[nt]*/,
                  unsigned *tile_mask /*[nt*32]*/);

/* ---- autotuner (no counterpart in the reference, which fixes these) --
 * Times candidate step configurations from one starting state and reports
 * the fastest: the PME FFT form (fft_mode 0, and 1 where every grid edge
 * allows it), the list rebuild cadence (freq_domdec 5, 10, 15, 20; the
 * list pad grows with it as 2 sqrt(30 kB T) freq dt, domdec/domdec.cu:138)
 * and the tiles per nonbonded work unit (tile_chunk 4, 8, 16), by
 * coordinate descent in that order from the values in `rp`.  Each
 * candidate is a fresh device system advanced `steps` graph-replayed
 * steps (rounded up to whole rebuild periods) after a warmup that captures
 * its graphs, timed on the host around synchronised launches; no energy
 * steps.  The caller's state is only read.  A candidate that fails (no
 * room, a NaN) is reported with its error code and never chosen. */
#define BM_TUNE_MAX_TRIALS 32
typedef struct bm_tune_trial {
  int32_t fft_mode, freq_domdec, tile_chunk, status;
  double ms_per_step;
} bm_tune_trial;
typedef struct bm_tune_result {
  int32_t fft_mode, freq_domdec, tile_chunk;    /* the fastest candidate */
  double ms_per_step;
  int32_t n_trials;
  bm_tune_trial trial[BM_TUNE_MAX_TRIALS];
} bm_tune_result;
int bm_autotune(int device, const bm_topology *top, const bm_runparams *rp,
                const double *pos, const double *vel,
                const double *theta, const double *theta_vel,
                int steps, bm_tune_result *out);
/* The same for n systems advanced together (bm_batch_dynamics): the list
 * cadence (10, 15, 20) and tile chunk (16, 8, 4) shared by every member,
 * tile chunk first; each member keeps its FFT form (fft_mode is reported
 * as -1).  Each candidate is a fresh batch of fresh device systems. */
int bm_batch_autotune(int device, int n, const bm_topology *const *tops,
                      const bm_runparams *const *rps,
                      const double *const *pos, const double *const *vel,
                      const double *const *theta, const double *const *theta_vel,
                      int steps, bm_tune_result *out);

/* ---- timing of the hot kernels over the last bm_dynamics call */
int bm_get_kernel_times(bm_system *sys, double *ms_nbdirect, double *ms_pme,
                        double *ms_total, long *n_measured_steps);

/* Per-phase timing: nsteps eager steps with HIP events on the system's
 * stream around the nbdirect and PME launches; out[3] = average ms of
 * nbdirect, PME (spread+FFT+gather) and the whole step. */
int bm_profile_steps(bm_system *sys, long step0, long nsteps, double *out);

/* Average duration of the nonbonded kernel: `reps` back-to-back launches
   (non-energy variant, current list and positions) between two HIP events on
   the system's stream; out[0] = ms per launch.  The launches rewrite the
   chunk partials and add dU/dlambda contributions that the next step's pack
   clears.  Measurement support, no counterpart in the reference. */
int bm_time_nbdirect(bm_system *sys, int reps, double *out);

/* Average durations of the PME phases: `reps` rounds of bins + spread,
   FFT with convolution, and the gather (the PME part of the force
   assembly), each between HIP events on the system's stream; out[3] = ms
   per round of spread (bins included), FFT and gather.  The gather adds
   the PME forces to the force array and empties the bins, as a step's
   assembly does.  BM_ENOTSUP without PME.  Measurement support, no
   counterpart in the reference. */
int bm_time_pme(bm_system *sys, int reps, double *out);

/* HBM calibration: `reps` launches of a plain 16-byte-vector copy kernel
   (bm_calibration_copy_kernel) over `bytes` bytes on `device`; out[0] = mean
   ms per launch.  Measurement support, no counterpart in the reference. */
int bm_calibration_copy(int device, size_t bytes, int reps, double *out);

/* raw fixed-point force accumulators [3N+nL] (bit-reproducibility checks) */
int bm_get_force_fixed(bm_system *sys, long long *force);

/* library build information and the last error message of this thread */
const char *bm_version(void);
const char *bm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
