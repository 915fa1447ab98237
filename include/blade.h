/*
 * blade.h - the reference-compatible BLaDE library API, as CHARMM binds it.
 *
 * Every function keeps the reference's name, argument list, argument
 * meaning (AKMA units, 1-based atom/block/site indices) and void return.
 * Where the reference calls fatal() and exits, this build records the
 * message (blade_last_error), reports it through blade_log, and makes later
 * calls on that System no-ops.
 *
 * Reference declarations replaced:
 *   system/system.h:86-91        blade_init_system .. blade_set_verbose
 *   system/parameters.h:202-230  blade_init_parameters .. blade_add_parameter_nbfixs
 *   system/structure.h:111-133   blade_init_structure .. blade_add_shake (+ restraint adders)
 *   system/coordinates.h:57-63   blade_init_coordinates .. blade_add_coordinates_box
 *   msld/msld.h:147-165          blade_init_msld .. blade_set_msld_block_fixed
 *   run/run.h:146-158            blade_init_run .. blade_run_energy
 *   io/io.h:69-83                blade_interpretter (subset) and blade_log
 *
 * Extensions (no reference counterpart) are grouped at the end: they give a
 * C caller the dynamics loop and host readback that the reference reaches
 * through its script interpreter and CHARMM-side copies.
 */
#ifndef BLADE_H
#define BLADE_H

#include <stdbool.h>
#include <stdio.h>

#include "blade_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct System System;

/* system/system.h:86-91 */
System *blade_init_system(int ngpus, int *gpus);
void blade_dest_system(System *system);
void blade_set_device(System *system);
void blade_set_verbose(System *system, int v);

/* system/parameters.h:202-230 */
void blade_init_parameters(System *system);
void blade_dest_parameters(System *system);
void blade_add_parameter_atoms(System *system, const char *name, double mass);
void blade_add_parameter_bonds(System *system, const char *t1, const char *t2, double kb, double b0);
void blade_add_parameter_angles(System *system, const char *t1, const char *t2, const char *t3,
                                double kangle, double angle0, double kureyb, double ureyb0);
void blade_add_parameter_dihes(System *system, const char *t1, const char *t2, const char *t3,
                               const char *t4, double kdih, int ndih, double dih0);
void blade_add_parameter_imprs(System *system, const char *t1, const char *t2, const char *t3,
                               const char *t4, double kimp, int nimp, double imp0);
void blade_add_parameter_cmaps(System *system, const char *t1, const char *t2, const char *t3,
                               const char *t4, const char *t5, const char *t6, const char *t7,
                               const char *t8, int ngrid);
void blade_add_parameter_cmaps_fill(System *system, const char *t1, const char *t2, const char *t3,
                                    const char *t4, const char *t5, const char *t6, const char *t7,
                                    const char *t8, int i, int j, double kcmapij);
void blade_add_parameter_nbonds(System *system, const char *t1, double eps, double sig,
                                double eps14, double sig14, double e14fac, int combine);
void blade_add_parameter_nbfixs(System *system, const char *t1, const char *t2, double eps,
                                double sig, double eps14, double sig14);

/* system/structure.h:111-133 */
void blade_init_structure(System *system);
void blade_dest_structure(System *system);
void blade_add_atom(System *system, int atomIdx, const char *segName, const char *resIdx,
                    const char *resName, const char *atomName, const char *atomTypeName,
                    double charge, double mass);
void blade_add_bond(System *system, int i, int j);
void blade_add_angle(System *system, int i, int j, int k);
void blade_add_dihe(System *system, int i, int j, int k, int l);
void blade_add_impr(System *system, int i, int j, int k, int l);
void blade_add_cmap(System *system, int i1, int j1, int k1, int l1, int i2, int j2, int k2, int l2);
void blade_add_shake(System *system, int shakeHbond);
/* virtual sites (system/structure.cxx:1059-1085), 1-based atoms: site v at
   host h1 plus (dist/|e| + scale) e, e = x_h1 - x_h2; or, with three hosts,
   at distance |dist| from h1 at angle theta and dihedral phi (degrees) in
   the frame of h1, h2, h3 (dist < 0: the bisector of h2 and h3 replaces
   h2).  Sites are placed before every force evaluation and their forces
   move onto the hosts; a site may host no other site, belong to no
   constraint group, and a system with sites cannot join a batch. */
void blade_add_virt2(System *system, int v, int h1, int h2, double dist, double scale);
void blade_add_virt3(System *system, int v, int h1, int h2, int h3, double dist, double theta, double phi);
/* restraints (system/structure.h:124-130; forms in bm_topology): 1-based
   atoms, lambdaBlock 1-based (1 = the environment, unscaled), angles in
   degrees */
void blade_add_noe(System *system, int i, int j, double rmin, double kmin, double rmax, double kmax,
                   double rpeak, double rswitch, double nswitch, double c0x, double c0y, double c0z,
                   bool is_pnoe);
void blade_add_harmonic(System *system, int i, double k, double x0, double y0, double z0, double n);
void blade_add_borest(System *system, int i, int j, double kr, double r0, int lambdaBlock);
void blade_add_anrest(System *system, int i, int j, int k, double kt, double t0, int lambdaBlock);
void blade_add_direst(System *system, int i, int j, int k, int l, double kphi, int nphi, double phi0,
                      double width, int lambdaBlock);
void blade_add_resd(System *system, int i1, int i2, int j1, int j2, double ci, double cj, double rdist,
                    double kdist);
/* system/structure.h:131-132: machine-learned potentials need LibTorch; the
   system is marked unsupported (BM_ENOTSUP when prepared) */
void blade_tani_internal_setup(System *system, const int is_tani, const char *ptname, const int ptnml,
                               const int *mlatomidx, const int *mlZidx, const int *mlmaskid,
                               const int mlnatoms);

/* system/coordinates.h:57-63 */
void blade_init_coordinates(System *system, int n);
void blade_dest_coordinates(System *system);
void blade_add_coordinates_position(System *system, int i, double x, double y, double z);
void blade_add_coordinates_velocity(System *system, int i, double vx, double vy, double vz);
void blade_add_coordinates_box(System *system, int name, double a, double b, double c,
                               double alpha, double beta, double gamma);

/* msld/msld.h:147-165 */
void blade_init_msld(System *system, int nblocks);
void blade_dest_msld(System *system);
void blade_add_msld_atomassignment(System *system, int atomIdx, int blockIdx);
void blade_add_msld_initialconditions(System *system, int blockIdx, int siteIdx, double theta0,
                                      double thetaVelocity, double thetaMass, double fixBias,
                                      double blockCharge);
void blade_add_msld_termscaling(System *system, int scaleBond, int scaleUrey, int scaleAngle,
                                int scaleDihe, int scaleImpr, int scaleCmap);
void blade_add_msld_flags(System *system, double gamma, double fnex, int useSoftCore,
                          int useSoftCore14, int msldEwaldType, double kRestraint,
                          double softBondRadius, double softBondExponent,
                          double softNotBondExponent, int fix);
void blade_add_msld_bias(System *system, int i, int j, int type, double l0, double k, int n);
void blade_add_msld_block_friction(System *system, int blockIdx, double friction);
void blade_set_msld_block_fixed(System *system, int blockIdx, int fixed);
/* theta biases (msld/msld.h:155-157): per-site collective and independent
   biases (the first call sizes the table with `sites` entries, site i
   1-based), and the edge bias of the piecewise constraint */
void blade_add_msld_thetacollbias(System *system, int sites, int i, double k, double n);
void blade_add_msld_thetaindebias(System *system, int sites, int i, double k);
void blade_set_msld_thetaedgebias(System *system, double k, double N, double alpha, double phi);
/* msld/msld.h:158: lambda_i = max(theta_i - t, 0)^3 per site instead of the
   softmax, with a flat-bottom harmonic of half-width nsubs*width on theta */
void blade_set_msld_piecewise_constraint(System *system, int do_imp, double width, double k);
/* msld/msld.h:159: soft bonds are recorded; a system with any is rejected
   with BM_EINVAL when its terms are derived (not supported) */
void blade_add_msld_softbond(System *system, int i, int j);
/* MSLD atom restraints (msld/msld.h:160-161): a new group, then its atoms
   (1-based); the force constant is blade_add_msld_flags' kRestraint */
void blade_add_msld_atomrestraint(System *system);
void blade_add_msld_atomrestraint_element(System *system, int atomIdx);
/* MSLD charge restraints (msld/msld.h:154) on the lambda-weighted net
   charge of the blocks' blockCharge values */
void blade_add_msld_charges(System *system, double kChargeRestraint1, double kChargeRestraint2,
                            double kChargeRestraint3, double q0ChargeRestraint3, double wChargeRestraint3);

/* run/run.h:146-158 */
void blade_init_run(System *system);
void blade_dest_run(System *system);
void blade_add_run_flags(System *system, double gamma, double betaEwald, double rCut,
                         double rSwitch, int vdWfSwitch, int elecPME, double gridSpace,
                         int gridx, int gridy, int gridz, int orderEwald, double shakeTolerance);
void blade_add_run_dynopts(System *system, int step, int step0, int nsteps, double dt, double T,
                           int freqNPT, double volumeFluctuation, double pressure);
void blade_run_energy(System *system);

/* io/io.h:69-70: runs an input script (the reference's language: variables,
   if/elseif/else/endif, while/endwhile, stream, selections, and the
   parameters, structure, coordinates, msld and run classes) on the system.
   An error is recorded on the system (blade_system_error) and ends the
   script. */
void blade_interpretter(const char *fnm, System *system);

/* io/io.h:82 - user-replaceable logger (weak default prints to stderr) */
void blade_log(const char *message);

/* ---- extensions ----------------------------------------------------- */
/* last error message of this thread ("" when none) */
const char *blade_last_error(void);
/* the first error of this system (empty when it has none); no counterpart
   in the reference, whose errors end the process */
const char *blade_system_error(System *system);
/* Run::dynamics: nsteps steps from the current step (run/run.cu:545-570);
 * energies on steps with step % freqNRG == 0; use_graph replays hipGraphs. */
int blade_run_dynamics(System *system, int nsteps, int freqNRG, int use_graph);
/* host readback after blade_run_energy / blade_run_dynamics (SoA [3N]) */
int blade_get_energy(System *system, double *energy /*[BM_NTERMS]*/);
int blade_get_forces(System *system, double *force);
int blade_get_positions(System *system, double *pos);
int blade_set_positions(System *system, const double *pos);
int blade_get_lambda(System *system, double *lambda);
int blade_get_dudl(System *system, double *dudl);
/* derived term tables and run parameters (host memory owned by System) */
const bm_topology *blade_get_topology(System *system);
const bm_runparams *blade_get_runparams(System *system);
/* the device system behind this System (NULL before initialisation) */
bm_system *blade_get_device_system(System *system);
/* derive terms without touching a GPU (returns 0 on success) */
int blade_prepare(System *system);
/* Philox seed of the Langevin noise */
void blade_set_seed(System *system, unsigned long long seed);
/* steps between neighbour-list rebuilds (default 10, domdec/domdec.cu:138);
   1 rebuilds every step, for starts whose atoms outrun the thermal list pad */
void blade_set_domdec_frequency(System *system, int steps);
/* Output files of blade_run_dynamics, as the reference's run module writes
   them (run/run.cu:572-590, io/io.cxx:350-387): every `freq` steps one
   .nrg line (step, then every energy term, " %12.4f") or one text .lmd
   line (step, then lambda of every block after the environment, " %8.6f").
   Energies are computed on the .nrg steps.  path NULL or freq 0 closes. */
int blade_set_output_nrg(System *system, const char *path, int freq);
int blade_set_output_lmd(System *system, const char *path, int freq);
/* GROMACS XTC coordinate frames (run/run.cu:299-317, io/io.cxx:314-345):
   every `freq` steps the positions and box in nm, the time in ps, at
   precision 1000 per nm.  path NULL or freq 0 closes. */
int blade_set_output_xtc(System *system, const char *path, int freq);
/* one XTC frame (positions AoS in nm, 3x3 box in nm row-major) appended
   to an open file, or written to a file by name (append nonzero: added at
   its end); 0 on success.  The frames are byte-identical to the
   reference's vendored xdrfile writer (src/xdr/xdrfile_xtc.cxx). */
int blade_xtc_frame(FILE *fp, int natoms, int step, float time, const float *box9, const float *xyz,
                    float precision);
int blade_write_xtc_frame(const char *path, int append, int natoms, int step, float time, const float *box9,
                          const float *xyz, float precision);
/* .lmd frame format, the reference's "run setvariable hrlmd" (run/run.cu:324):
   nonzero (the default) writes text lines, zero XDR frames of the block
   count and single-precision lambdas (io/io.cxx:347-370) */
void blade_set_lmd_format(System *system, int humanReadable);
/* Binary checkpoint of the exact FP64 state (step, box, positions,
   velocities, theta, theta velocities), replacing the reference's "%f"
   text checkpoint (io/io.cxx:416-536).  A run continued from a checkpoint
   written on a list rebuild step (a multiple of the domdec cadence) is
   bit-identical to the uninterrupted run.  Read it after setup and before
   the first energy or dynamics call. */
int blade_write_checkpoint(System *system, const char *path);
/* barostat counters of blade_run_dynamics with freqNPT > 0 (volume moves
   tried and kept) and the current box; no counterpart in the reference,
   which prints them to its log */
int blade_get_npt_stats(System *system, long *moves, long *accepted, double *box);
int blade_read_checkpoint(System *system, const char *path);
/* directed tiles per nonbonded work unit, 1..16 (default 16); no
   counterpart in the reference */
void blade_set_tile_chunk(System *system, int tiles);
/* Time candidate FFT forms, list rebuild cadences and tile chunks from the
   current coordinates (bm_autotune) and keep the fastest for the runs that
   follow.  Call after setup and before the first energy or dynamics call;
   `result` may be NULL.  No counterpart in the reference. */
int blade_autotune(System *system, int steps, struct bm_tune_result *result);
/* The same for n systems to be advanced together (bm_batch_autotune): the
   cadence and tile chunk chosen for the batch are set on every system. */
int blade_batch_autotune(int n, System *const *systems, int steps, struct bm_tune_result *result);
/* PME FFT form: 0 FP32 Stockham (default); 1 DFT contractions on the matrix
   cores in split-BF16x3 with FP32 accumulation (grid edges multiples of 16
   from 16 to 128).  No counterpart in the reference (cuFFT). */
void blade_set_fft_mode(System *system, int mode);

/* CHARMM input files, read through the setters above (the script commands
   "parameters file", "structure file psf", "coordinates file crd|pdb":
   system/parameters.cxx:129-484, system/structure.cxx:566-826,
   system/coordinates.cxx:152-274).  0 on success. */
int blade_read_parameter_file(System *system, const char *path);
int blade_read_psf(System *system, const char *path);
int blade_read_crd(System *system, const char *path);
int blade_read_pdb(System *system, const char *path);
/* structure queries: atom count, and (segment, residue id, atom name) of
   a 1-based atom into buffers of len bytes */
int blade_structure_atom_count(System *system);
int blade_structure_atom_names(System *system, int atomIdx, char *seg, char *res, char *name, int len);

#ifdef __cplusplus
}
#endif

#endif
