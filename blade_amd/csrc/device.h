// device.h - shared device-side definitions for the MI355X kernels.
//
// Everything here is compiled for gfx950 only (wave64).  Kernels take a
// `Dev` view by value: plain device pointers and sizes owned by the host
// runtime (runtime.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blade_mi355x.h"

namespace blade {

constexpr double kElectricD = 332.0716;            // main/defines.h:55
constexpr double kBoltzD = 0.001987191;            // main/defines.h:53
constexpr float kElectricF = 332.0716f;
constexpr float kSoftCoreRadius = 4.0f;            // main/defines.h:48
constexpr float kTwoOverSqrtPi = 1.1283791670955126f;
constexpr float kFixScaleF = 4294967296.0f;
constexpr double kFixScaleD = 4294967296.0;
constexpr double kFixInvD = 1.0 / 4294967296.0;

// force assembly components (force_assemble mask)
constexpr unsigned kAsmNb = 1u, kAsmBonded = 2u, kAsmPme = 4u;

// term families of the bonded launch (launch_bonded's mask); a family left
// out of the mask writes zero forces to its slots
constexpr unsigned kFamBond = 1u << 0, kFamAngle = 1u << 1, kFamDihe = 1u << 2, kFamImpr = 1u << 3,
                   kFamNb14 = 1u << 4, kFamNbex = 1u << 5, kFamCats = 1u << 6, kFamCmap = 1u << 7,
                   kFamRestrain = 1u << 8, kFamSelf = 1u << 9, kFamBias = 1u << 10, kFamAll = 0x7FFu;

// tile list capacity helpers
constexpr int kBlockAtoms = 32;      // atoms per block (one half-wave)
constexpr int kTileChunk = 16;       // most directed tiles in one nonbonded work unit
constexpr int kMaxColumnAtoms = 2048;

// Device view of one system.
struct Dev {
  int n, nl, nsites;             // atoms, lambda blocks, sites
  double box[3], invbox[3];
  float boxf[3];
  float fracscale[3];            // box / 2^32: fractional count -> A
  // state (SoA FP64, [x|y|z|theta])
  double *pos, *posb, *vel, *ism;
  double *xi;                    // [3N+nL] standard normals of the current step
  long long *force;              // [3N+nL] fixed point
  long long *dudl;               // [nL]
  long long *energy;             // [BM_NTERMS]
  double *lambda;                // [nL]
  float *lambdaf;                // [nL]
  int *nanflag;
  // per atom
  float *q;                      // [N]
  int *type;                     // [N]
  int *sb;                       // [N] siteBlock code
  double *qd;                    // [N] charge, FP64 (energy sums, spread)
  float *qlam;                   // [N] scratch: lambda*q (PME)
  float2 *vdw;                   // [T*T] (c12, c6)
  double2 *vdwd;                 // [T*T] (c12, c6) FP64 for energy sums
  double *blockq2;               // [nL] sum of q^2 over each block (Ewald self term)
  float *slo;                    // [32*maxB] slot charge low part: q - (float)q
  int ntypes;
  // domdec
  int ncx, ncy, ncol, maxblocks;
  long long maxtiles;
  unsigned *ufrac;               // [3][N] fractional coordinates at rebuild
  int *colcount, *colstart, *colfill, *bucket;   // [ncol],[ncol+1],[ncol],[N]
  int *blkstart;                 // [ncol+1] first block of each column
  int *nblocks;                  // [1]
  int *l2g;                      // [32*maxB] slot -> atom (-1 empty)
  int *g2s;                      // [N] atom -> slot
  unsigned *bblo, *bbhi;         // [3*maxB] rebuild boxes (exact intervals)
  uint4 *bbc;                    // [maxB] per-step box centre (frac counts)
  uint4 *bbh;                    // [maxB] per-step half extent (frac counts), .w = alch flag
  uint4 *spos;                   // [32*maxB] slot (ux,uy,uz,q bits)
  int2 *spar;                    // [32*maxB] slot (type or -1, siteBlock)
  int *tcount, *tstart;          // [maxB],[maxB+1] directed tiles per row
  int *ntiles;                   // [1]
  int *tile_i, *tile_j, *tile_ex;// [maxtiles]
  int *tile_flag;                // [maxtiles]
  int *ccount, *cstart;          // [maxB],[maxB+1] chunks per row
  int *nchunks;                  // [1]
  int maxchunks;
  int4 *chunk;                   // [maxchunks] (I, first tile, tiles, 0)
  float *ichunk;                 // [maxchunks][3][32] i-force partials of a chunk (FP32, fixed
                                 // order); the force assembly converts them to fixed point
  int nalchat;                   // atoms of lambda blocks (site block != 0)
  int *alchat;                   // [nalchat] their indices
  int *ecount, *estart;          // [maxB],[maxB+1] tiles with exclusions per row
  int *nexcltiles;               // [1]
  unsigned *emask;               // [32*2*n_excl]
  int *excl_tile;                // [2*n_excl] tiles (I,J) and (J,I) of each pair
  int nexcl;
  int *excl;                     // [2*n_excl]
  int *workctr;                  // [4] dynamic work counters
  int *errflag;                  // [1] overflow / capacity errors
  double rlist2;                 // (rcut+pad)^2, as the oracle computes it
  double ddscale[3];             // box * 2^-32, as the oracle computes it
  // nonbonded parameters
  float rc, rs, beta;
  float rc2;
  double rcd, rsd, betad;        // FP64 copies for energy-step sums
  int vdwmethod, elecmethod;
  // PME
  int K[3], order;
  float *rgrid;                  // [K0*K1*K2] real charge grid (spread output)
  int pnb[3];                    // spread bricks per axis (each >= order-1 points)
  int pcap;                      // atoms per brick bin
  int *pbcount;                  // [nbricks] atoms binned per brick, zeroed by the gather
  unsigned *pbhdr;               // [nbricks*pcap] each record's first stencil point, 10 bits per axis
  double *pbrec;                 // [nbricks*pcap] spline records by home brick
                                 // (pme_spline_rec_words(order) doubles each)
  float2 *cgrid;                 // [K0*K1*K2] complex work grid
  float2 *twiddle;               // [K0+K1+K2] e^{-2 pi i k/K} per axis
  float *bmod[3];                // [K] B-spline moduli per axis
  float pmefac;                  // 0.5 kE/(pi V)
  float pmeexp;                  // pi^2/beta^2
  double *bmodd[3];              // FP64 moduli (energy sum)
  double pmefacd, pmeexpd;
  int fft_mode;                  // 0 FP32 Stockham, 1 MFMA split-BF16x3 DFT
  int tchunk;                    // directed tiles per nonbonded work unit (<= kTileChunk)
  const void *mdft[3][2];        // per axis, hi/lo: bf16 A fragments of the forward DFT matrix (fft_mfma)
  // bonded terms (float parameters, int indices)
  int nbond, nbond12, nangle, ndihe, nimpr, nnb14, nnbex;
  int2 *bond_i; int2 *bond_sb; double2 *bond_p;           // (kb, b0)
  int4 *angle_i; int2 *angle_sb; double2 *angle_p;        // (k, t0)
  int4 *dihe_i; int2 *dihe_sb; double4 *dihe_p;           // (k, n, d0)
  int4 *impr_i; int2 *impr_sb; double4 *impr_p;           // (k, n, d0, soft)
  // MSLD soft terms: per-term flags of bonds (with Urey-Bradley) and
  // angles (null when none is soft; dihedrals and impropers carry theirs in
  // p.w), soft-bond alpha and power, and the power of other soft terms
  int *bond_soft, *angle_soft;
  double sb_alpha, sb_pow, snb_exp;
  int2 *nb14_i; int2 *nb14_sb; double4 *nb14_p;           // (qq, c12, c6, e14fac)
  int2 *nbex_i; int2 *nbex_sb; double *nbex_p;            // qq
  // bonded force slots: every term stores its per-atom forces (fixed
  // point) in slots of its own, laid out atom by atom (a static CSR), so
  // the force assembly reads each atom's slots as one contiguous run
  int sl_bond, sl_angle, sl_dihe, sl_impr, sl_nb14, sl_nbex, sl_cats, nslot;
  long long *bslot;              // [3][nslot] in atom order
  int *bcsr_start;               // [N+1] first slot position of each atom
  int *bspos;                    // [nslot] term slot -> position in atom order
  // constraints
  int ntri, nbr1, nbr2, nbr3;
  int4 *tri_i; double4 *tri_b;   // b0 (01,02,12), .w unused
  int2 *br1_i; double *br1_b;
  int4 *br2_i; double2 *br2_b;
  int4 *br3_i; double4 *br3_b;
  double *mass;                  // [N]
  // virtual sites: placed from their hosts before the forces, their
  // forces moved onto the hosts after them (no site hosts another)
  int nvirt2, nvirt3;
  int4 *virt2_i; double2 *virt2_p;   // (site, h0, h1, -); (dist, scale)
  int4 *virt3_i; double4 *virt3_p;   // (site, h0, h1, h2); [2n] (dist, cos t, sin t, -), (cos p, sin p, -, -)
  int nfree;                     // atoms in no constraint group
  int *freeatom;                 // [nfree]
  // msld
  int *block_site, *site_bound, *block_fixed;
  double *fixed_bias;
  double *lfric, *lnoise;        // [nL] per-block theta friction/noise
  int ncats;                     // MSLD atom restraint groups
  int *cats_bound, *cats_idx;    // [ncats+1] first member, [members] atoms
  double cats_k;                 // kRestraint
  int ncmap;                     // CMAP cross terms
  int4 *cmap_i;                  // [2n] atoms of phi, atoms of psi
  int4 *cmap_s;                  // [n] (sb0, sb1, sb2, grid)
  int2 *cmap_g;                  // [grids] (ngrid, first point)
  double4 *cmap_tab;             // per point (E, dE/dphi, dE/dpsi, d2E/dphi dpsi), grid units
  int sl_cmap;
  // restraints: NOE, harmonic positions, bond/angle/dihedral restraints, RESD
  int nnoe, nharm, nbor, nanr, ndir, nresd;
  int2 *noe_i; double *noe_p;    // (i, j or -1); 10 parameters each
  int *harm_i; double *harm_p;   // atom; 5 parameters each (k, x0, y0, z0, n)
  int4 *bor_i; double2 *bor_p;   // (i, j, block); (kr, r0)
  int4 *anr_i; int *anr_b; double2 *anr_p;   // atoms; block; (kt, t0)
  int4 *dir_i; int2 *dir_bn; double4 *dir_p; // atoms; (block, n); (kphi, phi0, width)
  int4 *resd_i; double4 *resd_p; // (i1, i2, j1, j2); (ci, cj, rdist, kdist)
  int sl_noe, sl_harm, sl_bor, sl_anr, sl_dir, sl_resd;
  int ncr;                       // charge restraints on Q = sum_b lambda_b q_b
  double *blockcharge;           // [nL] q_b
  double crk1, crk2, crk3, crq0, crw;
  int nvb;
  int4 *vb_i;                    // (i, j, type, n)
  double2 *vb_p;                 // (l0, k)
  // theta biases and the piecewise lambda constraint
  const double *tcoll_k, *tcoll_n, *tinde_k;   // [nsites] or null
  int piecewise, fixlam;         // lambda = max(theta - t_s, 0)^3; lambda held at its initial theta
  double pw_width, pw_k;         // flat-bottom harmonic on theta beyond +-nb*pw_width
  double edge_k, edge_n, edge_alpha, edge_phi;   // edge bias on the smooth substituent count
  double *theta0;                // [nsites] piecewise root t_s of the current theta
  long long *dtheta;             // [nL] dU/dtheta of the theta biases (fixed point), zeroed by the pack
  double fnex;
  int softcore, softcore14;
  // integrator
  double halfdt, fric, noise, lfric0, lnoise0, shaketol;
  unsigned long long seed;
  long long *step;               // [1] step counter (RNG counter), advanced on device
};

// One atom's PME spread record: lambda*q, the B-spline weights along each
// axis and the first stencil point per axis (as doubles: exact small ints).
__host__ __device__ constexpr int pme_spline_rec_words(int order) { return 1 + 3 * order + 3; }

__device__ __forceinline__ long long fix_f(float v) {
  return (long long)__float2ll_rn(v * kFixScaleF);
}
__device__ __forceinline__ long long fix_d(double v) {
  return (long long)__double2ll_rn(v * kFixScaleD);
}
__device__ __forceinline__ void atomic_fix(long long *p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}
__device__ __forceinline__ void atomic_fix_f(long long *p, float v) { atomic_fix(p, fix_f(v)); }

// 32-bit fractional coordinate of x in a periodic box of edge L:
// floor(x*(1/L)*2^32) taken modulo 2^32.  Must round exactly like the host
// oracle: one rounded multiply, an exact power-of-two scale, floor.
__device__ __forceinline__ unsigned frac_u32(double x, double invL) {
  double f = __dmul_rn(x, invL);
  double s = f * 4294967296.0;
  return (unsigned)(long long)floor(s);
}

// first non-finite value wins the NaN flag (atom or DOF index)
__device__ __forceinline__ void nan_check(const Dev &d, double v, int i) {
  if (!isfinite(v)) atomicCAS(d.nanflag, -1, i);
}

// Integrator updates with their rounding fixed by explicit fused
// multiply-adds, so the fused step kernels and the one-term-at-a-time
// kernels produce the same bits.
// V: v - k halfdt ism^2 f (k = 2 for a full kick, 1 for a half kick)
__device__ __forceinline__ double kick_v(double v, double k, double halfdt, double ism, double f) {
  const double c = __dmul_rn(__dmul_rn(__dmul_rn(k, halfdt), ism), ism);
  return __fma_rn(-c, f, v);
}
// O: a v + noise ism xi
__device__ __forceinline__ double o_v(double a, double v, double noise, double ism, double xi) {
  return __fma_rn(a, v, __dmul_rn(__dmul_rn(noise, ism), xi));
}
// R: x + halfdt v
__device__ __forceinline__ double drift_x(double x, double halfdt, double v) { return __fma_rn(halfdt, v, x); }

// A raised capacity overflow or a non-finite value halts the state updates
// of the step that raised it and of every later step: the positions,
// velocities, theta and the step counter stay as they were at the start of
// that step, so the host can grow the buffers and run the step again.
__device__ __forceinline__ bool halted(const Dev &d) { return *d.errflag != 0 || *d.nanflag >= 0; }
// the same test taken once per workgroup (the flags may be raised while the
// kernel runs): every thread of a workgroup returns together or none does
__device__ __forceinline__ bool halted_wg(const Dev &d) {
  __shared__ int h;
  if (threadIdx.x == 0) h = halted(d) ? 1 : 0;
  __syncthreads();
  return h != 0;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -----------------
__host__ __device__ inline uint4 philox(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    uint4 o;
    o.x = (unsigned)(p1 >> 32) ^ c.y ^ k.x;
    o.y = (unsigned)p1;
    o.z = (unsigned)(p0 >> 32) ^ c.w ^ k.y;
    o.w = (unsigned)p0;
    c = o;
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// minimum-image displacement from fractional coordinates (int wrap)
__device__ __forceinline__ float fdisp(unsigned a, unsigned b, float scale) {
  return (float)(int)(a - b) * scale;
}

// Reduce a long long across a wave64 and add it with one atomic.
__device__ __forceinline__ void wave_atomic_fix(long long *p, long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomic_fix(p, v);
}

// Wave-aggregated counter claim: every lane of the wave must call it.  Lanes
// with equal keys share one atomicAdd; each active lane gets the counter
// value before the add plus its rank among the lanes of its key.  The key
// groups are found first (ballots only), then every group's leader issues
// its atomic at once: one round trip per wave, whatever the number of keys.
__device__ __forceinline__ int wave_claim(int *counters, int key, bool active) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(active);
  unsigned long long grp = 0;
  int leader = lane;
  while (pending) {
    const int l = __ffsll((long long)pending) - 1;
    const int lkey = __shfl(key, l, 64);
    const bool mine = ((pending >> lane) & 1ull) && key == lkey;
    const unsigned long long g = __ballot(mine);
    if (mine) {
      grp = g;
      leader = l;
    }
    pending &= ~g;
  }
  int base = 0;
  if (active && lane == leader) base = atomicAdd(&counters[key], __popcll(grp));
  base = __shfl(base, leader, 64);
  return active ? base + __popcll(grp & ((1ull << lane) - 1ull)) : 0;
}

// Several systems in one launch (batches): the workgroups of system k are
// [wg0[k], wg0[k+1]) of the grid; devs holds device copies of the systems'
// Dev.  A kernel body runs with its system's Dev and its workgroup index
// and count within that system, exactly as a launch of that system alone.
struct Multi {
  const Dev *devs;   // [nsys]
  const int *wg0;    // [nsys+1]
  int nsys, nwg;     // systems, workgroups in all
};

struct MultiBlk {
  int sys, bid, nblk;
};

__device__ __forceinline__ MultiBlk multi_block(const Multi &m) {
  const int b = blockIdx.x;
  int lo = 0, hi = m.nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.wg0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return MultiBlk{lo, b - m.wg0[lo], m.wg0[lo + 1] - m.wg0[lo]};
}

// Block-wide fixed-point energy sum (any blockDim multiple of 64).
__device__ __forceinline__ void block_energy(long long *p, float e) {
  wave_atomic_fix(p, fix_f(e));
}

}  // namespace blade
