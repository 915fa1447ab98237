/*
 * oracle.h - FP64 host restatement of BLaDE's per-step force and
 * integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X kernels and the CPU baseline timed by bench.py; nothing in the
 * product path links, loads or calls it.  Every function names the
 * reference file:line it restates.  Inputs are the same derived topology
 * (bm_topology) and run parameters (bm_runparams) the GPU library takes.
 *
 * Arrays are structure-of-arrays: pos[d*N+i]; force likewise (dU/dx).
 */
#ifndef BLADE_ORACLE_H
#define BLADE_ORACLE_H

#include "../include/blade_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* term selection bits for orc_forces */
#define ORC_NBDIRECT   (1u << 0)
#define ORC_EWALDSELF  (1u << 1)
#define ORC_PME        (1u << 2)
#define ORC_BONDED     (1u << 3)
#define ORC_NB14       (1u << 4)
#define ORC_NBEX       (1u << 5)
#define ORC_MSLDBIAS   (1u << 6)
#define ORC_RESTRAIN   (1u << 7)
#define ORC_ALL        0xFFu

/* All selected force terms at positions `pos` and lambdas `lambda`
 * (Potential::calc_force, system/potential.cxx:1851-1941).  force[3N] and
 * dudl[nL] are overwritten; energy[BM_NTERMS] is overwritten.  Returns
 * the number of non-excluded pairs with r<rcut (the pair-interaction count
 * of one nbdirect evaluation) when ORC_NBDIRECT is set, else 0. */
long orc_forces(const bm_topology *top, const bm_runparams *rp,
                const double *pos, const double *lambda, unsigned terms,
                double *force, double *dudl, double *energy);

/* One PME exclusion correction at distance r (bonded/pair.cu:181-193):
 * e = -kqq erf(beta r)/r and fr = (de/dr)/r, by Taylor series about r = 0
 * when beta r < 0.01 (coincident atoms give the finite limit). */
void orc_nbex_pair(double kqq, double beta, double r, double *fr, double *e);

/* Reciprocal-space Ewald energy/forces by direct structure-factor
 * summation over |m_i| <= kmax (validation of the PME path). */
double orc_ewald_recip_direct(const bm_topology *top, const bm_runparams *rp,
                              const double *pos, const double *lambda, int kmax,
                              double *force, double *dudl);

/* Real-space Ewald sum over all images within rcut with plain erfc
 * (no switching, no exclusions): used with the reciprocal sum and the
 * self term to reproduce a known Madelung constant. */
double orc_ewald_real_images(int n, const double *q, const double *pos,
                             const double *box, double beta, double rcut);

/* Cardinal B-spline value M_n(u) and the PME moduli 1/|b(m)|^2
 * (system/potential.cxx:1245-1287). */
double orc_bspline(int order, double u);
void orc_bspline_moduli(int order, int K, double *out /*[K]*/);

/* In-place complex 3D FFT (unnormalised), sign = -1 forward, +1 inverse,
 * data interleaved re/im, index (i*Ky+j)*Kz+k. */
void orc_fft3d(int kx, int ky, int kz, double *data, int sign);

/* Domain decomposition semantics (domdec/assign_domain.cu:14-45,
 * domdec/assign_blocks.cu:24-468, domdec/cull.cu:44-375,
 * domdec/assign_excl.cu:17-371) restated on integer fractional
 * coordinates.  Wraps pos in place.  Outputs (caller-sized):
 *   local_to_global[32*max_blocks] (-1 = empty slot), block_count[max_blocks],
 *   tile_i/tile_j[max_tiles]: directed tiles, every (I,J) whose rebuild
 *   boxes lie within rcut+pad (both orders, and (I,I)), by I then J;
 *   tile_mask[32*max_tiles]: bit j of word i set when slot i of block I and
 *   slot j of block J are an excluded pair.
 * Returns the number of tiles, and *n_blocks; -1 when a buffer is too small. */
long orc_domdec(const bm_topology *top, const bm_runparams *rp, double *pos,
                int ncol_x, int ncol_y, int max_blocks, long max_tiles,
                int *local_to_global, int *block_count, int *n_blocks,
                int *tile_i, int *tile_j, unsigned *tile_mask);

/* Column counts used by the list build (domdec/domdec.cu:110-118). */
void orc_domdec_columns(const bm_runparams *rp, int n_atoms, int *ncol_x, int *ncol_y);

/* MSLD theta -> lambda (msld/msld.cu:863-921; the piecewise constraint
 * msld.cu:785-861 and the fix flag msld.cu:913-935 as the topology selects)
 * and the chain rule dU/dlambda -> dU/dtheta (msld/msld.cu:937-998). */
void orc_theta_to_lambda(const bm_topology *top, const bm_runparams *rp,
                         const double *theta, double *lambda);
void orc_chain_rule(const bm_topology *top, const bm_runparams *rp,
                    const double *theta, const double *lambda,
                    const double *dudl, double *dudtheta);

/* MSLD theta biases (msld/msld.cu:1155-1264): collective and independent
 * biases, and with the piecewise constraint the flat-bottom harmonic and the
 * edge bias; adds dU/dtheta into dudtheta[nL] and the energy into
 * energy[BM_ETHETA]. */
void orc_theta_bias(const bm_topology *top, const bm_runparams *rp, const double *theta,
                    double *dudtheta, double *energy);

/* Holonomic constraints (holonomic/holonomic.cu:253-600 position,
 * :15-211 velocity).  x/xprev/v are [3N] SoA; halfdt in AKMA. */
void orc_holonomic_position(const bm_topology *top, const bm_runparams *rp,
                            double *x, const double *xprev, double *v, double halfdt);
/* holonomic_rectify: constraint groups put at exactly their lengths */
void orc_rectify(const bm_topology *t, const bm_runparams *rp, double *x);
void orc_holonomic_velocity(const bm_topology *top, const bm_runparams *rp,
                            const double *x, double *v);
/* Virtual sites (holonomic/virtual.cu:11-131): every site placed at its
 * first host plus the offset its hosts fix.  orc_forces moves each site's
 * force onto its hosts as J^T F (J the site's derivative in its hosts'
 * positions) and leaves the site's entry zero; orc_step places the sites
 * before its forces.  A site goes to the periodic image nearest its
 * previous position.  Massless particles are not integrated. */
void orc_virtual_place(const bm_topology *t, const bm_runparams *rp, double *x);

/* Philox4x32-10 block (Salmon et al., SC'11) and the standard normal the
 * integrator draws for degree of freedom `dof` at step `step`. */
void orc_philox4x32(const unsigned ctr[4], const unsigned key[2], unsigned out[4]);
double orc_normal(uint64_t seed, uint64_t step, uint64_t dof);

/* One VRORV Langevin step (update/update.cu:233-305) in FP64.  state:
 * x[3N+nL] (theta last), v[3N+nL], lambda[nL].  Forces are evaluated at
 * the incoming x.  energy[BM_NTERMS] receives potential terms and, on
 * energy steps, the kinetic energy.  Returns 0, or -1 on NaN. */
int orc_step(const bm_topology *top, const bm_runparams *rp, long step, int energy_step,
             double *x, double *v, double *lambda, double *energy);

/* Thread count used by the OpenMP loops. */
int orc_threads(void);

#ifdef __cplusplus
}
#endif

#endif
