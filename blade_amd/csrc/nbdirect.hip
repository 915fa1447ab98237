// nbdirect.hip - real-space nonbonded tile kernel
// (restates nbdirect/nbdirect.cu:87-490 pair semantics).
//
// Work unit: a chunk of at most d.tchunk directed tiles (I, J) of one row
// I of the rebuild list.  One wave64 owns a chunk: both 32-lane halves hold
// the 32 i-atoms of block I in registers, so i-forces accumulate without
// any cross-lane or memory traffic, and no j-side force is ever written
// (every pair is visited from both sides; energies take half per visit).
//
// Per tile and step: the tile is skipped when the current block boxes are
// farther apart than rcut; otherwise the j-atoms within rcut of the current
// i-box form a wave-uniform 32-bit mask, and the two halves take the set
// bits in turns (j data broadcast from LDS).  Tiles without alchemical atoms
// run a path with no lambda arithmetic.
//
// Accumulation: each lane sums its i-atom's pair forces over the chunk's
// tiles in FP32, in a fixed order; at the end of a chunk the two halves'
// sums are added and stored to the chunk's own slot.  The force assembly
// (pme.hip) converts the chunks of every atom's block to 64-bit fixed point
// and adds them to the force array, one thread per atom: no global
// atomics, and the result does not depend on the order in which the
// chunks ran.
#include "kernels.h"
#include "erfc_coef.h"

namespace blade {

// erfc(x) = exp(-x^2) * poly(1/(1+P x)) given ex = exp(-x^2).  The
// hardware reciprocal (1 ulp) replaces the IEEE division sequence.
__device__ __forceinline__ float erfc_fit(float x, float ex) {
  float t = __builtin_amdgcn_rcpf(fmaf(BM_ERFC_P, x, 1.0f));
  float p = bm_erfc_c[BM_ERFC_DEG];
#pragma unroll
  for (int k = BM_ERFC_DEG - 1; k >= 0; k--) p = fmaf(p, t, bm_erfc_c[k]);
  return ex * p;
}

struct NbConst {
  float rc, rs, rc2, beta, beta2log2e;
  // VFSWITCH
  float dv6, k6, k12, rcinv3, rcinv6;
  // FSWITCH
  float A, B, C, D, dvc, rcinv;
  // VSWITCH
  float c2of, c2on, rul3inv, rul12;
  // VSHIFT
  float rcinv12, rcinv18;
};

__device__ __forceinline__ NbConst make_const(const Dev &d) {
  NbConst c;
  c.rc = d.rc; c.rs = d.rs; c.rc2 = d.rc2; c.beta = d.beta;
  c.beta2log2e = d.beta * d.beta * 1.4426950408889634f;
  float rc3 = d.rc * d.rc * d.rc, rs3 = d.rs * d.rs * d.rs;
  c.dv6 = 1.0f / (rc3 * rs3);
  c.k6 = rc3 / (rc3 - rs3);
  c.k12 = rc3 * rc3 / (rc3 * rc3 - rs3 * rs3);
  c.rcinv3 = 1.0f / rc3;
  c.rcinv6 = c.rcinv3 * c.rcinv3;
  float roff2 = d.rc * d.rc, ron2 = d.rs * d.rs;
  float ginv = 1.0f / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
  c.A = roff2 * roff2 * (roff2 - 3 * ron2) * ginv;
  c.B = 6 * roff2 * ron2 * ginv;
  c.C = -(ron2 + roff2) * ginv;
  c.D = 2 * ginv / 5;
  c.dvc = 8 * (ron2 * roff2 * (d.rc - d.rs) - (roff2 * roff2 * d.rc - ron2 * ron2 * d.rs) / 5) * ginv;
  c.rcinv = 1.0f / d.rc;
  c.c2of = roff2; c.c2on = ron2;
  c.rul3inv = ginv;
  c.rul12 = 12 * ginv;
  c.rcinv12 = c.rcinv6 * c.rcinv6;
  c.rcinv18 = c.rcinv12 * c.rcinv6;
  return c;
}

// One pair at distance r (1/r = rinv); qq already x kE.  f = dE/dr.  The
// energy is returned factored, E = qq*eq + c12*eA - c6*eB, so an energy
// sum can apply the parameters in FP64 (the factors carry only the
// distance-dependent FP32 rounding, which does not bias the sum).
template <int ELEC, int VDW>
__device__ __forceinline__ void pair_ef(const NbConst &c, float qq, float c12, float c6, float r,
                                        float rinv, float &f, float &eq, float &eA, float &eB) {
  if (ELEC == BM_PME) {
    float br = c.beta * r;
    // argument >= -(beta rcut)^2 log2(e), far above the denormal range
    float ex = __builtin_amdgcn_exp2f(-c.beta2log2e * r * r);
    float erfcrinv = erfc_fit(fminf(br, BM_ERFC_XMAX), ex) * rinv;
    f = -qq * (erfcrinv + kTwoOverSqrtPi * c.beta * ex) * rinv;
    eq = erfcrinv;
  } else if (ELEC == BM_FSWITCH) {
    if (r <= c.rs) {
      f = -qq * rinv * rinv;
      eq = rinv + c.dvc;
    } else {
      // A/r^2 + B + 3C r^2 + 5D r^4 = g (roff^2-r^2)^2 (roff^2+2r^2-3ron^2)/r^2
      // with g = 1/(roff^2-ron^2)^3: the product form has no cancellation
      const float r2 = r * r, du = c.c2of - r2;
      f = -qq * rinv * rinv * c.rul3inv * du * du * (c.c2of + 2 * r2 - 3 * c.c2on);
      // the energy polynomial cancels to a few digits: FP64 (it is only
      // formed for the lambda derivative of alchemical pairs)
      const double rcd = c.rc, rsd = c.rs, roff2 = rcd * rcd, ron2 = rsd * rsd;
      const double g = 1.0 / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
      const double A = roff2 * roff2 * (roff2 - 3 * ron2) * g, B = 6 * roff2 * ron2 * g;
      const double C = -(ron2 + roff2) * g, D = 0.4 * g;
      const double rd = r, r3 = rd * rd * rd, r5 = r3 * rd * rd;
      eq = (float)(A * (1.0 / rd - 1.0 / rcd) + B * (rcd - rd) + C * (roff2 * rcd - r3) +
                   D * (roff2 * roff2 * rcd - r5));
    }
  } else {
    f = -qq * (rinv * rinv - c.rcinv * c.rcinv);
    const double rd = r, rci = 1.0 / (double)c.rc;
    eq = (float)(1.0 / rd - 2 * rci + rd * rci * rci);
  }
  float rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  if (VDW == BM_VFSWITCH) {
    // branch-free: the lanes of a wave straddle rswitch
    const bool in = r < c.rs;
    const float a3 = rinv3 - c.rcinv3, a6 = rinv6 - c.rcinv6;
    const float A = in ? rinv6 : c.k6 * a3 * rinv3;
    const float B = in ? rinv6 * rinv6 : c.k12 * a6 * rinv6;
    f += (6 * c6 * A - 12 * c12 * B) * rinv;
    eA = in ? rinv6 * rinv6 - c.dv6 * c.dv6 : c.k12 * a6 * a6;
    eB = in ? rinv6 - c.dv6 : c.k6 * a3 * a3;
  } else if (VDW == BM_VSWITCH) {
    if (r < c.rs) {
      f += (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv;
      eA = rinv6 * rinv6;
      eB = rinv6;
    } else {
      float rijl = c.c2on - r * r, riju = c.c2of - r * r;
      float fsw = riju * riju * (riju - 3 * rijl) * c.rul3inv, dfsw = rijl * riju * c.rul12;
      f += fsw * (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv + dfsw * (c12 * rinv6 - c6) * rinv6;
      eA = fsw * rinv6 * rinv6;
      eB = fsw * rinv6;
    }
  } else {
    float r2 = r * r, r5 = r2 * r2 * r;
    float rinv7 = rinv6 * rinv, rinv12 = rinv6 * rinv6, rinv13 = rinv12 * rinv;
    f += 6 * c6 * (rinv7 - r5 * c.rcinv12) - 12 * c12 * (rinv13 + r5 * c.rcinv18);
    // shifted energies vanish as (rc-r)^2: FP64 (alchemical pairs only)
    const double ri = 1.0 / (double)r, ri6 = ri * ri * ri * ri * ri * ri, rd = r;
    const double r6d = rd * rd * rd * rd * rd * rd;
    eA = (float)(ri6 * ri6 + 2 * r6d * (double)c.rcinv18 - 3 * (double)c.rcinv12);
    eB = (float)(ri6 + r6d * (double)c.rcinv12 - 2 * (double)c.rcinv6);
  }
}

// FP64 pair energy at effective distance r (energy steps only): the total
// potential is a small difference of large pair sums, so the sum is built
// from FP64 pair energies (same pair set and parameters as the forces).
template <int ELEC, int VDW>
__device__ double pair_energy_d(const Dev &d, double qq, double c12, double c6, double r) {
  const double rc = d.rcd, rs = d.rsd, rinv = 1.0 / r;
  double e;
  if (ELEC == BM_PME) {
    e = qq * erfc(d.betad * r) * rinv;
  } else if (ELEC == BM_FSWITCH) {
    const double roff2 = rc * rc, ron2 = rs * rs;
    const double ginv = 1 / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
    const double A = roff2 * roff2 * (roff2 - 3 * ron2) * ginv, B = 6 * roff2 * ron2 * ginv;
    const double C = -(ron2 + roff2) * ginv, D = 2 * ginv / 5;
    const double dvc = 8 * (ron2 * roff2 * (rc - rs) - (roff2 * roff2 * rc - ron2 * ron2 * rs) / 5) * ginv;
    const double r3 = r * r * r, r5 = r3 * r * r;
    e = (r <= rs) ? qq * (rinv + dvc)
                  : qq * (A * (rinv - 1 / rc) + B * (rc - r) + C * (roff2 * rc - r3) + D * (roff2 * roff2 * rc - r5));
  } else {
    e = qq * (rinv - 2 / rc + r / (rc * rc));
  }
  const double rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  if (VDW == BM_VFSWITCH) {
    const double rc3 = rc * rc * rc, rs3 = rs * rs * rs;
    if (r < rs) {
      const double dv6 = 1 / (rc3 * rs3);
      e += c12 * (rinv6 * rinv6 - dv6 * dv6) - c6 * (rinv6 - dv6);
    } else {
      const double k6 = rc3 / (rc3 - rs3), k12 = rc3 * rc3 / (rc3 * rc3 - rs3 * rs3);
      const double a3 = rinv3 - 1 / rc3, a6 = rinv6 - 1 / (rc3 * rc3);
      e += c12 * k12 * a6 * a6 - c6 * k6 * a3 * a3;
    }
  } else if (VDW == BM_VSWITCH) {
    if (r < rs) {
      e += c12 * rinv6 * rinv6 - c6 * rinv6;
    } else {
      const double c2of = rc * rc, c2on = rs * rs;
      const double rul3 = (c2of - c2on) * (c2of - c2on) * (c2of - c2on);
      const double rijl = c2on - r * r, riju = c2of - r * r;
      e += riju * riju * (riju - 3 * rijl) / rul3 * (c12 * rinv6 - c6) * rinv6;
    }
  } else {
    const double r6 = r * r * r * r * r * r, rc6 = rc * rc * rc * rc * rc * rc;
    e += c12 * (rinv6 * rinv6 + 2 * r6 / (rc6 * rc6 * rc6) - 3 / (rc6 * rc6)) - c6 * (rinv6 + r6 / (rc6 * rc6) - 2 / rc6);
  }
  return e;
}

// cull test between current block boxes (centre/half extent, fractional
// counts): true when the boxes may hold a pair closer than rcut
__device__ __forceinline__ bool boxes_within(const Dev &d, int I, int J, float rc2) {
  uint4 ci = d.bbc[I], cj = d.bbc[J], hi = d.bbh[I], hj = d.bbh[J];
  float g2 = 0;
  {
    long long g = llabs((long long)(int)(ci.x - cj.x)) - (long long)hi.x - (long long)hj.x;
    float gf = g > 0 ? (float)g * d.fracscale[0] : 0.f;
    g2 += gf * gf;
  }
  {
    long long g = llabs((long long)(int)(ci.y - cj.y)) - (long long)hi.y - (long long)hj.y;
    float gf = g > 0 ? (float)g * d.fracscale[1] : 0.f;
    g2 += gf * gf;
  }
  {
    long long g = llabs((long long)(int)(ci.z - cj.z)) - (long long)hi.z - (long long)hj.z;
    float gf = g > 0 ? (float)g * d.fracscale[2] : 0.f;
    g2 += gf * gf;
  }
  return g2 <= rc2 * 1.0001f;
}

// j-atom (fractional coordinates pj) within rcut of the current box of
// block I (centre/half extent in fractional counts)
__device__ __forceinline__ bool near_box(const Dev &d, uint4 ci, uint4 hi, uint4 pj, float rc2) {
  float g2 = 0;
  {
    long long g = llabs((long long)(int)(pj.x - ci.x)) - (long long)hi.x;
    float gf = g > 0 ? (float)g * d.fracscale[0] : 0.f;
    g2 += gf * gf;
  }
  {
    long long g = llabs((long long)(int)(pj.y - ci.y)) - (long long)hi.y;
    float gf = g > 0 ? (float)g * d.fracscale[1] : 0.f;
    g2 += gf * gf;
  }
  {
    long long g = llabs((long long)(int)(pj.z - ci.z)) - (long long)hi.z;
    float gf = g > 0 ? (float)g * d.fracscale[2] : 0.f;
    g2 += gf * gf;
  }
  return g2 <= rc2 * 1.0001f;
}

// i-atom of a lane, held for a whole chunk
struct IAtom {
  uint4 p;       // fractional coordinates, charge bits
  int t, b;      // LJ type (-1: empty slot), siteBlock
  float q;       // kE * charge
  float lam;     // lambda of its block (1 outside sites)
  double qd;     // kE * charge in FP64 (energy steps)
};

// LJ pair table staged in LDS (dynamic size ntypes^2); tables beyond
// kVdwLdsMax entries are read from global memory by a separate variant
constexpr int kVdwLdsMax = 6144;

// The pairs of one tile for the lane's i-atom: the nj j-atoms near the
// i-box, compacted into sp/spr/slo (sjx: their slot index in block J),
// shared out between the halves (half 0 takes entries 0, 2, ..., half 1
// entries 1, 3, ...).  CHK: the tile has exclusions or is the diagonal
// tile (I == J), so the exclusion bit and j == i are tested.
template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT, bool CHK>
__device__ __forceinline__ void tile_pairs(const Dev &d, const NbConst &c, const IAtom &ia, int li,
                                           int half, const uint4 *sp, const int2 *spr, const float *slo,
                                           const unsigned char *sjx, const float2 *svdw, int nj,
                                           unsigned em, bool self, float &fx, float &fy, float &fz,
                                           double &etile) {
  const float sx = d.fracscale[0], sy = d.fracscale[1], sz = d.fracscale[2];
  for (int jl = half; jl < nj; jl += 2) {
    const uint4 pj = sp[jl];
    const float dx = fdisp(ia.p.x, pj.x, sx);
    const float dy = fdisp(ia.p.y, pj.y, sy);
    const float dz = fdisp(ia.p.z, pj.z, sz);
    const float r2 = dx * dx + dy * dy + dz * dz;
    bool ok = ia.t >= 0 && r2 < c.rc2;
    if (CHK) {
      const int jo = sjx[jl];
      ok = ok && !((em >> jo) & 1u) && !(self && jo == li);
    }
    if (!ok) continue;
    const int2 prj = spr[jl];
    const int vi = ia.t * d.ntypes + prj.x;
    const float2 lj = svdw[vi];
    float rinv = __frsqrt_rn(r2);
    float r = r2 * rinv;
    const float qq = ia.q * __uint_as_float(pj.w);
    float s = 1.f, dredr = 1.f, lamj = 1.f;
    const int bi = ia.b, bj = prj.y;
    const bool alchPair = ALCH && (bi | bj);
    if (ALCH && alchPair) {
      if (bj & 0xFFFF) lamj = d.lambdaf[bj & 0xFFFF];
      const bool same = (bi & 0xFFFF0000) == (bj & 0xFFFF0000);
      s = same ? (bi == bj ? ia.lam : 0.f) : ia.lam * lamj;
      if (SOFT) {
        const float rSoft = kSoftCoreRadius * (1.f - s);
        if (r < rSoft) {
          const float q = r / rSoft;
          float re = 1.f - 0.5f * q;
          re = re * q * q * q + 0.5f;
          dredr = (3.f - 2.f * q) * q * q;
          r = re * rSoft;
          rinv = __builtin_amdgcn_rcpf(r);
        }
      }
    }
    float f, eq, eA, eB;
    pair_ef<ELEC, VDW>(c, qq, lj.x, lj.y, r, rinv, f, eq, eA, eB);
    if (ALCH) f *= s;
    if (ENERGY) {
      // FP64 pair energy from the exact integer displacement and FP64
      // charges / LJ parameters; half per visit
      const double2 ljd = d.vdwd[vi];
      const double qjd = (double)__uint_as_float(pj.w) + (double)slo[jl];
      const double ddx = (double)(int)(ia.p.x - pj.x) * d.ddscale[0];
      const double ddy = (double)(int)(ia.p.y - pj.y) * d.ddscale[1];
      const double ddz = (double)(int)(ia.p.z - pj.z) * d.ddscale[2];
      double rd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
      double sd = 1.0;
      if (ALCH && alchPair) {
        const double li_d = (bi & 0xFFFF) ? d.lambda[bi & 0xFFFF] : 1.0;
        const double lj_d = (bj & 0xFFFF) ? d.lambda[bj & 0xFFFF] : 1.0;
        const bool same = (bi & 0xFFFF0000) == (bj & 0xFFFF0000);
        sd = same ? (bi == bj ? li_d : 0.0) : li_d * lj_d;
        if (SOFT) {
          const double rSoft = 4.0 * (1.0 - sd);
          if (rd < rSoft) {
            const double q = rd / rSoft;
            rd = ((1.0 - 0.5 * q) * q * q * q + 0.5) * rSoft;
          }
        }
      }
      etile += 0.5 * sd * pair_energy_d<ELEC, VDW>(d, ia.qd * qjd, ljd.x, ljd.y, rd);
    }
    (void)eq; (void)eA; (void)eB;
    float fr;
    if (SOFT && alchPair) fr = f * dredr * __frsqrt_rn(r2);
    else fr = f * rinv;
    fx += fr * dx;
    fy += fr * dy;
    fz += fr * dz;
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One chunk of the list: the tiles of row I it holds, against the 32
// i-atoms in registers; the i-force partials go to the chunk's slot.
template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT>
__device__ __forceinline__ void chunk_pairs(const Dev &d, const NbConst &c, const float2 *s_vdw, int ch, int lane,
                                            int half, int li, uint4 *sp, int2 *spr, float *slo,
                                            unsigned char *sjx, double &eacc) {
  const int4 cr = d.chunk[ch];
  const int I = cr.x, t0 = cr.y, cnt = cr.z;
  IAtom ia;
  {
    const int si = I * kBlockAtoms + li;
    ia.p = d.spos[si];
    const int2 pr = d.spar[si];
    ia.t = pr.x;
    ia.b = pr.y;
    ia.q = __uint_as_float(ia.p.w) * kElectricF;
    ia.qd = (ENERGY || ALCH) ? kElectricD * ((double)__uint_as_float(ia.p.w) + (double)d.slo[si]) : 0.0;
    ia.lam = (ALCH && (ia.b & 0xFFFF)) ? d.lambdaf[ia.b & 0xFFFF] : 1.f;
  }
  const uint4 ci = d.bbc[I], hi = d.bbh[I];
  int tJ = 0, tE = -1;
  bool keep = false;
  if (lane < cnt) {
    tJ = d.tile_j[t0 + lane];
    tE = d.tile_ex[t0 + lane];
    keep = boxes_within(d, I, tJ, c.rc2);
  }
  unsigned long long act = __ballot(keep);
  // i-partials of the whole chunk in FP32 (at most 16 tiles, a fixed
  // order), stored once per chunk
  float fx = 0.f, fy = 0.f, fz = 0.f;
  // software pipeline: the j slots, exclusion word and alchemical flag of
  // the next tile are loaded while the current tile computes
  int k = act ? __ffsll((long long)act) - 1 : -1;
  if (k >= 0) act &= act - 1;
  int J = 0;
  uint4 pj = make_uint4(0, 0, 0, 0);
  int2 prj = make_int2(-1, 0);
  float lo = 0.f;
  unsigned em = 0u;
  unsigned alchJ = 0u;
  int tEc = -1;
  auto fetch = [&](int kk, int &J_, uint4 &pj_, int2 &prj_, float &lo_, unsigned &em_, unsigned &al_,
                   int &ex_) {
    J_ = __builtin_amdgcn_readlane(tJ, kk);
    const int ex = __builtin_amdgcn_readlane(tE, kk);
    ex_ = ex;
    const int sj = J_ * kBlockAtoms + li;
    pj_ = d.spos[sj];
    prj_ = d.spar[sj];
    if (ENERGY || ALCH) lo_ = d.slo[sj];
    em_ = ex >= 0 ? d.emask[ex * kBlockAtoms + li] : 0u;
    al_ = ALCH ? d.bbh[J_].w : 0u;
  };
  if (k >= 0) fetch(k, J, pj, prj, lo, em, alchJ, tEc);
  while (k >= 0) {
    // the j-atoms near the i-box, compacted to the front of the staging
    const bool near = half == 0 && prj.x >= 0 && near_box(d, ci, hi, pj, c.rc2);
    const unsigned long long nm = __ballot(near);
    const int nj = __popcll(nm);
    if (near) {
      const int e = __popcll(nm & ((1ull << li) - 1ull));
      sp[e] = pj;
      spr[e] = prj;
      sjx[e] = (unsigned char)li;
      if (ENERGY || ALCH) slo[e] = lo;
    }
    const int Jc = J, ex = tEc;
    const unsigned emc = em, alchJc = alchJ;
    const int kn = act ? __ffsll((long long)act) - 1 : -1;
    if (kn >= 0) {
      act &= act - 1;
      fetch(kn, J, pj, prj, lo, em, alchJ, tEc);
    }
    wave_lds_sync();
    double etile = 0.0;
    const bool alchTile = ALCH && (hi.w | alchJc);
    const bool self = I == Jc, chk = self || ex >= 0;
#define BM_TILE(AL, SO, CK)                                                                        \
tile_pairs<ELEC, VDW, ENERGY, AL, SO, CK>(d, c, ia, li, half, sp, spr, slo, sjx, s_vdw, nj, emc, \
                                          self, fx, fy, fz, etile)
    if (alchTile) {
      if (chk) BM_TILE(ALCH, SOFT, true);
      else BM_TILE(ALCH, SOFT, false);
    } else {
      if (chk) BM_TILE(false, false, true);
      else BM_TILE(false, false, false);
    }
#undef BM_TILE
    if (ENERGY) eacc += etile;
    wave_lds_sync();
    k = kn;
  }
  // both halves hold partials of the same i-atoms: (own + other) is the
  // same sum in either half
  fx += __shfl_xor(fx, 32, 64);
  fy += __shfl_xor(fy, 32, 64);
  fz += __shfl_xor(fz, 32, 64);
  if (half == 0) {
    float *dst = d.ichunk + (long long)ch * 3 * kBlockAtoms;
    dst[li] = fx;
    dst[kBlockAtoms + li] = fy;
    dst[2 * kBlockAtoms + li] = fz;
  }

}

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT, bool GTABLE>
__device__ __forceinline__ void nbdirect_kernel_body(const Dev &d, int bid, int nblk) {
  __shared__ uint4 s_pos[4][kBlockAtoms];
  __shared__ int2 s_par[4][kBlockAtoms];
  __shared__ float s_lo[4][kBlockAtoms];
  __shared__ unsigned char s_jx[4][kBlockAtoms];
  extern __shared__ float2 s_vdw_dyn[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int half = lane >> 5, li = lane & 31;
  if (!GTABLE) {
    const int nt2 = d.ntypes * d.ntypes;
    for (int k = threadIdx.x; k < nt2; k += blockDim.x) s_vdw_dyn[k] = d.vdw[k];
    __syncthreads();
  }
  const float2 *s_vdw = GTABLE ? d.vdw : s_vdw_dyn;
  const NbConst c = make_const(d);
  const bool fits = *d.ntiles <= d.maxtiles && *d.nchunks <= d.maxchunks;
  const int nchunks = fits ? *d.nchunks : 0;
  double eacc = 0.0;
  uint4 *sp = s_pos[w];
  int2 *spr = s_par[w];
  float *slo = s_lo[w];
  unsigned char *sjx = s_jx[w];

  (void)nblk;
  // one chunk per wave; waves past this step's chunk count exit
  const int k = bid * 4 + w;
  if (k >= nchunks) return;
  chunk_pairs<ELEC, VDW, ENERGY, ALCH, SOFT>(d, c, s_vdw, k, lane, half, li, sp, spr, slo, sjx, eacc);
  if (ENERGY) {
    long long v = (long long)__double2ll_rn(eacc * kFixScaleD);
    wave_atomic_fix(&d.energy[BM_ENBDIRECT], v);
  }
}

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT, bool GTABLE>
__global__ void __launch_bounds__(256) nbdirect_kernel(Dev d) { nbdirect_kernel_body<ELEC, VDW, ENERGY, ALCH, SOFT, GTABLE>(d, blockIdx.x, gridDim.x); }

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT, bool GTABLE>
__global__ void __launch_bounds__(256) nbdirect_kernel_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  nbdirect_kernel_body<ELEC, VDW, ENERGY, ALCH, SOFT, GTABLE>(d, b.bid, b.nblk);
}

// dU/dlambda of the nonbonded pairs (the lambda derivative of
// nbdirect/nbdirect.cu:244-275,400-415) in FP64, by alchemical atom: the
// tile kernels leave it out, since FP32 pair energies summed over a site's
// thousands of pairs drift past the 1e-3 kcal/mol dU/dlambda bar.  Wave w
// takes atom alchat[w / kDudlWaves] and every kDudlWaves-th pair of tiles
// of its block's row (a half-wave per tile, a lane per j slot); each pair
// within rcut that is not excluded gives d(s E)/ds (E in FP64 at the FP64
// effective distance; the soft-core term s dE/dr dr/ds with dE/dr in FP32)
// times the lambda of the other block, or half of it within one block
// (the pair is met again from the other atom), to the atom's block.
constexpr int kDudlWaves = 32;
template <int ELEC, int VDW, bool SOFT>
__device__ __forceinline__ void nb_dudl_body(const Dev &d, int bid) {
  const int lane = threadIdx.x & 63, half = lane >> 5, lj = lane & 31;
  const int wv = bid * 4 + (threadIdx.x >> 6);
  const int ai = wv / kDudlWaves, sub = wv % kDudlWaves;
  if (ai >= d.nalchat) return;
  if (*d.ntiles > d.maxtiles || *d.nchunks > d.maxchunks) return;
  const int a = d.alchat[ai], sa = d.g2s[a], I = sa / kBlockAtoms, li = sa % kBlockAtoms;
  const uint4 pa = d.spos[sa];
  const int2 pra = d.spar[sa];
  const int ba = pra.y;
  const double qa = kElectricD * ((double)__uint_as_float(pa.w) + (double)d.slo[sa]);
  const double la = d.lambda[ba & 0xFFFF];
  const NbConst c = make_const(d);
  double acc = 0.0;
  const int t1 = d.tstart[I + 1];
  for (int t = d.tstart[I] + 2 * sub + half; t < t1; t += 2 * kDudlWaves) {
    const int J = d.tile_j[t], ex = d.tile_ex[t];
    const int sj = J * kBlockAtoms + lj;
    const int2 prj = d.spar[sj];
    if (prj.x < 0 || sj == sa) continue;
    if (ex >= 0 && ((d.emask[ex * kBlockAtoms + li] >> lj) & 1u)) continue;
    const int bj = prj.y;
    const bool same = (ba & 0xFFFF0000) == (bj & 0xFFFF0000);
    if (same && ba != bj) continue;   // two blocks of one site do not interact
    const uint4 pj = d.spos[sj];
    const double dx = (double)(int)(pa.x - pj.x) * d.ddscale[0];
    const double dy = (double)(int)(pa.y - pj.y) * d.ddscale[1];
    const double dz = (double)(int)(pa.z - pj.z) * d.ddscale[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r >= d.rcd) continue;
    const double lo = (bj & 0xFFFF) ? d.lambda[bj & 0xFFFF] : 1.0;
    const double sd = same ? la : la * lo;
    double reff = r, dredll = 0.0;
    if (SOFT) {
      const double rSoft = 4.0 * (1.0 - sd);
      if (r < rSoft) {
        const double q = r / rSoft, re = (1.0 - 0.5 * q) * q * q * q + 0.5;
        dredll = (re - (3.0 - 2.0 * q) * q * q * q) * (-4.0);
        reff = re * rSoft;
      }
    }
    const double2 ljd = d.vdwd[pra.x * d.ntypes + prj.x];
    const double qq = qa * ((double)__uint_as_float(pj.w) + (double)d.slo[sj]);
    double fl = pair_energy_d<ELEC, VDW>(d, qq, ljd.x, ljd.y, reff);
    if (SOFT && dredll != 0.0) {
      float f, eq, eA, eB;
      const float rf = (float)reff;
      pair_ef<ELEC, VDW>(c, (float)qq, (float)ljd.x, (float)ljd.y, rf, 1.f / rf, f, eq, eA, eB);
      fl += sd * (double)f * dredll;
    }
    acc += same ? 0.5 * fl : lo * fl;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0 && acc != 0.0) atomic_fix(&d.dudl[ba & 0xFFFF], fix_d(acc));
}

template <int ELEC, int VDW, bool SOFT>
__global__ void __launch_bounds__(256) nb_dudl_kernel(Dev d) { nb_dudl_body<ELEC, VDW, SOFT>(d, blockIdx.x); }

template <int ELEC, int VDW, bool SOFT>
__global__ void __launch_bounds__(256) nb_dudl_kernel_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  nb_dudl_body<ELEC, VDW, SOFT>(d, b.bid);
}

int grid_nb_dudl(const Dev &d) { return (d.nalchat * kDudlWaves + 3) / 4; }

template <int ELEC, int VDW>
static void launch_dudl_ev(const Dev &d, bool soft, hipStream_t st) {
  const unsigned g = (unsigned)grid_nb_dudl(d);
  if (!g) return;
  if (soft) nb_dudl_kernel<ELEC, VDW, true><<<g, 256, 0, st>>>(d);
  else nb_dudl_kernel<ELEC, VDW, false><<<g, 256, 0, st>>>(d);
}

template <int ELEC, int VDW>
static void launch_dudl_mev(const Multi &m, bool soft, hipStream_t st) {
  if (soft) nb_dudl_kernel_multi<ELEC, VDW, true><<<m.nwg, 256, 0, st>>>(m);
  else nb_dudl_kernel_multi<ELEC, VDW, false><<<m.nwg, 256, 0, st>>>(m);
}

#define BM_METHODS(F, ...)                                                         \
  if (elec == BM_PME) {                                                            \
    if (vdw == BM_VFSWITCH) F<BM_PME, BM_VFSWITCH>(__VA_ARGS__);                   \
    else if (vdw == BM_VSWITCH) F<BM_PME, BM_VSWITCH>(__VA_ARGS__);                \
    else F<BM_PME, BM_VSHIFT>(__VA_ARGS__);                                        \
  } else if (elec == BM_FSWITCH) {                                                 \
    if (vdw == BM_VFSWITCH) F<BM_FSWITCH, BM_VFSWITCH>(__VA_ARGS__);               \
    else if (vdw == BM_VSWITCH) F<BM_FSWITCH, BM_VSWITCH>(__VA_ARGS__);            \
    else F<BM_FSWITCH, BM_VSHIFT>(__VA_ARGS__);                                    \
  } else {                                                                         \
    if (vdw == BM_VFSWITCH) F<BM_FSHIFT, BM_VFSWITCH>(__VA_ARGS__);                \
    else if (vdw == BM_VSWITCH) F<BM_FSHIFT, BM_VSWITCH>(__VA_ARGS__);             \
    else F<BM_FSHIFT, BM_VSHIFT>(__VA_ARGS__);                                     \
  }

void launch_nb_dudl(const Dev &d, bool soft, hipStream_t st) {
  const int elec = d.elecmethod, vdw = d.vdwmethod;
  BM_METHODS(launch_dudl_ev, d, soft, st)
}

void launch_nb_dudl_multi(const Multi &m, int elec, int vdw, bool soft, hipStream_t st) {
  BM_METHODS(launch_dudl_mev, m, soft, st)
}
#undef BM_METHODS

// one wave per chunk slot of the list capacity; waves past this step's
// chunk count exit at once
int grid_nbdirect(const Dev &d) { return (d.maxchunks + 3) / 4; }

// the LJ table in LDS (0: the table is read from global memory)
size_t lds_nbdirect(const Dev &d) {
  const int nt2 = d.ntypes * d.ntypes;
  return nt2 <= kVdwLdsMax ? (size_t)nt2 * sizeof(float2) : 0;
}

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT>
static void launch_t(const Dev &d, hipStream_t st) {
  const unsigned grid = (unsigned)grid_nbdirect(d);
  const size_t lds = lds_nbdirect(d);
  if (lds)
    nbdirect_kernel<ELEC, VDW, ENERGY, ALCH, SOFT, false><<<grid, 256, lds, st>>>(d);
  else
    nbdirect_kernel<ELEC, VDW, ENERGY, ALCH, SOFT, true><<<grid, 256, 0, st>>>(d);
}

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT>
static void launch_mt(const Multi &m, size_t lds, hipStream_t st) {
  if (lds)
    nbdirect_kernel_multi<ELEC, VDW, ENERGY, ALCH, SOFT, false><<<m.nwg, 256, lds, st>>>(m);
  else
    nbdirect_kernel_multi<ELEC, VDW, ENERGY, ALCH, SOFT, true><<<m.nwg, 256, 0, st>>>(m);
}

// systems with lambda sites run the kernel with the lambda arithmetic
template <int ELEC, int VDW>
static void launch_ev(const Dev &d, bool energy, bool alch, bool soft, hipStream_t st) {
  if (!alch) {
    if (energy) launch_t<ELEC, VDW, true, false, false>(d, st);
    else launch_t<ELEC, VDW, false, false, false>(d, st);
  } else if (!soft) {
    if (energy) launch_t<ELEC, VDW, true, true, false>(d, st);
    else launch_t<ELEC, VDW, false, true, false>(d, st);
  } else {
    if (energy) launch_t<ELEC, VDW, true, true, true>(d, st);
    else launch_t<ELEC, VDW, false, true, true>(d, st);
  }
}

// one launch over several members sharing the methods and flags
template <int ELEC, int VDW>
static void launch_mev(const Multi &m, bool energy, bool alch, bool soft, size_t lds, hipStream_t st) {
  if (!alch) {
    if (energy) launch_mt<ELEC, VDW, true, false, false>(m, lds, st);
    else launch_mt<ELEC, VDW, false, false, false>(m, lds, st);
  } else if (!soft) {
    if (energy) launch_mt<ELEC, VDW, true, true, false>(m, lds, st);
    else launch_mt<ELEC, VDW, false, true, false>(m, lds, st);
  } else {
    if (energy) launch_mt<ELEC, VDW, true, true, true>(m, lds, st);
    else launch_mt<ELEC, VDW, false, true, true>(m, lds, st);
  }
}

// systems sharing the methods and flags; lds: the largest LJ table of the
// group (0 when any of them is read from global memory)
void launch_nbdirect_multi(const Multi &m, int elec, int vdw, bool energy, bool alch, bool soft, size_t lds,
                           hipStream_t st) {
  if (elec == BM_PME) {
    if (vdw == BM_VFSWITCH) launch_mev<BM_PME, BM_VFSWITCH>(m, energy, alch, soft, lds, st);
    else if (vdw == BM_VSWITCH) launch_mev<BM_PME, BM_VSWITCH>(m, energy, alch, soft, lds, st);
    else launch_mev<BM_PME, BM_VSHIFT>(m, energy, alch, soft, lds, st);
  } else if (elec == BM_FSWITCH) {
    if (vdw == BM_VFSWITCH) launch_mev<BM_FSWITCH, BM_VFSWITCH>(m, energy, alch, soft, lds, st);
    else if (vdw == BM_VSWITCH) launch_mev<BM_FSWITCH, BM_VSWITCH>(m, energy, alch, soft, lds, st);
    else launch_mev<BM_FSWITCH, BM_VSHIFT>(m, energy, alch, soft, lds, st);
  } else {
    if (vdw == BM_VFSWITCH) launch_mev<BM_FSHIFT, BM_VFSWITCH>(m, energy, alch, soft, lds, st);
    else if (vdw == BM_VSWITCH) launch_mev<BM_FSHIFT, BM_VSWITCH>(m, energy, alch, soft, lds, st);
    else launch_mev<BM_FSHIFT, BM_VSHIFT>(m, energy, alch, soft, lds, st);
  }
}

void launch_nbdirect(const Dev &d, bool energy, bool alch, bool soft, int nwg, hipStream_t st) {
  (void)nwg;
  if (d.elecmethod == BM_PME) {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_PME, BM_VFSWITCH>(d, energy, alch, soft, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_PME, BM_VSWITCH>(d, energy, alch, soft, st);
    else launch_ev<BM_PME, BM_VSHIFT>(d, energy, alch, soft, st);
  } else if (d.elecmethod == BM_FSWITCH) {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_FSWITCH, BM_VFSWITCH>(d, energy, alch, soft, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_FSWITCH, BM_VSWITCH>(d, energy, alch, soft, st);
    else launch_ev<BM_FSWITCH, BM_VSHIFT>(d, energy, alch, soft, st);
  } else {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_FSHIFT, BM_VFSWITCH>(d, energy, alch, soft, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_FSHIFT, BM_VSWITCH>(d, energy, alch, soft, st);
    else launch_ev<BM_FSHIFT, BM_VSHIFT>(d, energy, alch, soft, st);
  }
}

}  // namespace blade
