// nbdirect.hip - real-space nonbonded tile kernel
// (restates nbdirect/nbdirect.cu:87-490 pair semantics).
//
// Work unit: a tile = (i-block I, j-block J>=I) of 32x32 slots from the
// rebuild list.  One wave64 processes two tiles at once, one per 32-lane
// half.  Lane li of a half holds i-atom li in registers; the 32 j-atoms of
// its tile sit in LDS.  At step k lane li meets j-atom (li+k)&31, so the 32
// lanes of a half touch 32 distinct j slots: j-forces accumulate in LDS
// without atomics or cross-lane reductions.  A tile whose current block
// boxes are farther apart than rcut is skipped (per-step cull).
//
// Accumulation: per-tile FP32 partial sums are converted to 64-bit fixed
// point; i-forces stay in registers while consecutive tiles share I, and
// are added to the global fixed-point force array with one atomic per
// component when I changes.  Energies accumulate in FP64 per lane.
#include "kernels.h"
#include "erfc_coef.h"

namespace blade {

// erfc(x) = exp(-x^2) * poly(1/(1+P x)); returns erfc and exp(-x^2)
__device__ __forceinline__ float erfc_fit(float x, float ex) {
  float t = __frcp_rn(fmaf(BM_ERFC_P, x, 1.0f));
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
    float ex = exp2f(-c.beta2log2e * r * r);
    float erfcrinv = erfc_fit(fminf(br, BM_ERFC_XMAX), ex) * rinv;
    f = -qq * (erfcrinv + kTwoOverSqrtPi * c.beta * ex) * rinv;
    eq = erfcrinv;
  } else if (ELEC == BM_FSWITCH) {
    if (r <= c.rs) {
      f = -qq * rinv * rinv;
      eq = rinv + c.dvc;
    } else {
      float r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
      f = -qq * rinv * (c.A * rinv + c.B * r + 3 * c.C * r3 + 5 * c.D * r5);
      eq = c.A * (rinv - c.rcinv) + c.B * (c.rc - r) + c.C * (c.c2of * c.rc - r3) +
           c.D * (c.c2of * c.c2of * c.rc - r5);
    }
  } else {
    f = -qq * (rinv * rinv - c.rcinv * c.rcinv);
    eq = rinv - 2 * c.rcinv + r * c.rcinv * c.rcinv;
  }
  float rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  if (VDW == BM_VFSWITCH) {
    if (r < c.rs) {
      f += (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv;
      eA = rinv6 * rinv6 - c.dv6 * c.dv6;
      eB = rinv6 - c.dv6;
    } else {
      float a3 = rinv3 - c.rcinv3, a6 = rinv6 - c.rcinv6;
      f += (6 * c6 * c.k6 * a3 * rinv3 - 12 * c12 * c.k12 * a6 * rinv6) * rinv;
      eA = c.k12 * a6 * a6;
      eB = c.k6 * a3 * a3;
    }
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
    float r2 = r * r, r5 = r2 * r2 * r, r6 = r2 * r2 * r2;
    float rinv7 = rinv6 * rinv, rinv12 = rinv6 * rinv6, rinv13 = rinv12 * rinv;
    f += 6 * c6 * (rinv7 - r5 * c.rcinv12) - 12 * c12 * (rinv13 + r5 * c.rcinv18);
    eA = rinv12 + 2 * r6 * c.rcinv18 - 3 * c.rcinv12;
    eB = rinv6 + r6 * c.rcinv12 - 2 * c.rcinv6;
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

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT>
__global__ void __launch_bounds__(256) nbdirect_kernel(Dev d) {
  __shared__ uint4 s_pos[4][2][kBlockAtoms];
  __shared__ int2 s_par[4][2][kBlockAtoms];
  __shared__ float4 s_f[4][2][kBlockAtoms];
  __shared__ float s_lo[4][2][kBlockAtoms];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int half = lane >> 5, li = lane & 31;
  const NbConst c = make_const(d);
  const float sx = d.fracscale[0], sy = d.fracscale[1], sz = d.fracscale[2];
  const int ntiles = *d.ntiles;
  const int nchunks = (ntiles + kTileChunk - 1) / kTileChunk;
  double eacc = 0.0;
  uint4 *sp = s_pos[w][half];
  int2 *spr = s_par[w][half];
  float4 *sf = s_f[w][half];
  float *slo = s_lo[w][half];

  for (;;) {
    int chunk = 0;
    if (lane == 0) chunk = atomicAdd(&d.workctr[0], 1);
    chunk = __shfl(chunk, 0, 64);
    if (chunk >= nchunks) break;
    const int t0 = chunk * kTileChunk;
    const int cnt = min(kTileChunk, ntiles - t0);
    bool keep = false;
    if (lane < cnt) keep = boxes_within(d, d.tile_i[t0 + lane], d.tile_j[t0 + lane], c.rc2);
    unsigned long long act = __ballot(keep);

    int curI = -1, ai = -1;
    uint4 pi = make_uint4(0, 0, 0, 0);
    int ti = -1, bi = 0;
    float qi = 0.f, lami = 1.f;
    double qid = 0.0;
    long long fix_x = 0, fix_y = 0, fix_z = 0;
    float fli = 0.f;

    while (true) {
      int tA = -1, tB = -1;
      if (act) { tA = __ffsll((long long)act) - 1; act &= act - 1; }
      if (act) { tB = __ffsll((long long)act) - 1; act &= act - 1; }
      const int myt = half ? tB : tA;
      const bool hv = myt >= 0;
      const int I = hv ? d.tile_i[t0 + myt] : -2;
      // flush the i accumulators when this half moves to another i-block
      if (I != curI) {
        if (curI >= 0 && ai >= 0) {
          atomic_fix(&d.force[ai], fix_x);
          atomic_fix(&d.force[d.n + ai], fix_y);
          atomic_fix(&d.force[2 * d.n + ai], fix_z);
          if (ALCH && (bi & 0xFFFF) && fli != 0.f) atomic_fix_f(&d.dudl[bi & 0xFFFF], fli);
        }
        fix_x = fix_y = fix_z = 0;
        fli = 0.f;
        curI = I;
        ai = -1;
        if (hv) {
          const int si = I * kBlockAtoms + li;
          pi = d.spos[si];
          int2 pr = d.spar[si];
          ti = pr.x;
          bi = pr.y;
          ai = ti >= 0 ? d.l2g[si] : -1;
          qi = __uint_as_float(pi.w) * kElectricF;
          if (ENERGY) qid = kElectricD * ((double)__uint_as_float(pi.w) + (double)d.slo[si]);
          lami = 1.f;
          if (ALCH && (bi & 0xFFFF)) lami = d.lambdaf[bi & 0xFFFF];
        }
      }
      if (tA < 0) break;  // both halves done (tB < 0 implies tA < 0 next)

      int J = -1, ex = -1;
      if (hv) {
        J = d.tile_j[t0 + myt];
        ex = d.tile_ex[t0 + myt];
        const int sj = J * kBlockAtoms + li;
        sp[li] = d.spos[sj];
        spr[li] = d.spar[sj];
        if (ENERGY) slo[li] = d.slo[sj];
      }
      sf[li] = make_float4(0.f, 0.f, 0.f, 0.f);
      const unsigned emask = (ex >= 0) ? d.emask[ex * kBlockAtoms + li] : 0u;
      const bool self = (I == J);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      float fx = 0.f, fy = 0.f, fz = 0.f;
      double etile = 0.0;
      const bool ivalid = hv && ti >= 0;
      for (int k = 0; k < kBlockAtoms; k++) {
        const int jl = (li + k) & (kBlockAtoms - 1);
        const uint4 pj = sp[jl];
        const int2 prj = spr[jl];
        bool ok = ivalid && prj.x >= 0 && !((emask >> jl) & 1u);
        if (self) ok = ok && k > 0 && (li + k) < kBlockAtoms;
        const float dx = fdisp(pi.x, pj.x, sx);
        const float dy = fdisp(pi.y, pj.y, sy);
        const float dz = fdisp(pi.z, pj.z, sz);
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (ok && r2 < c.rc2) {
          const float2 lj = d.vdw[ti * d.ntypes + prj.x];
          float rinv = __frsqrt_rn(r2);
          float r = r2 * rinv;
          const float qq = qi * __uint_as_float(pj.w);
          float s = 1.f, dredr = 1.f, dredll = 0.f, lamj = 1.f;
          const int bj = prj.y;
          const bool alchPair = ALCH && (bi | bj);
          if (ALCH && alchPair) {
            if (bj & 0xFFFF) lamj = d.lambdaf[bj & 0xFFFF];
            const bool same = (bi & 0xFFFF0000) == (bj & 0xFFFF0000);
            s = same ? (bi == bj ? lami : 0.f) : lami * lamj;
            if (SOFT) {
              const float rSoft = kSoftCoreRadius * (1.f - s);
              if (r < rSoft) {
                const float q = r / rSoft;
                float re = 1.f - 0.5f * q;
                re = re * q * q * q + 0.5f;
                dredr = (3.f - 2.f * q) * q * q;
                dredll = (re - dredr * q) * (-kSoftCoreRadius);
                r = re * rSoft;
                rinv = 1.f / r;
              }
            }
          }
          float f, eq, eA, eB;
          pair_ef<ELEC, VDW>(c, qq, lj.x, lj.y, r, rinv, f, eq, eA, eB);
          const float e = qq * eq + lj.x * eA - lj.y * eB;
          if (ALCH) f *= s;
          if (ALCH && alchPair) {
            const float fl = SOFT ? e + f * dredll : e;
            const bool same = (bi & 0xFFFF0000) == (bj & 0xFFFF0000);
            if (same) {
              if (bi == bj) fli += fl;
            } else {
              fli += lamj * fl;
              atomicAdd(&sf[jl].w, lami * fl);
            }
          }
          if (ENERGY) {
            // FP64 pair energy from the exact integer displacement and FP64
            // charges / LJ parameters
            const double2 ljd = d.vdwd[ti * d.ntypes + prj.x];
            const double qjd = (double)__uint_as_float(pj.w) + (double)slo[jl];
            const double ddx = (double)(int)(pi.x - pj.x) * d.ddscale[0];
            const double ddy = (double)(int)(pi.y - pj.y) * d.ddscale[1];
            const double ddz = (double)(int)(pi.z - pj.z) * d.ddscale[2];
            double rd = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            double sd = 1.0;
            if (ALCH && alchPair) {
              const int bj = prj.y;
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
            (void)eq; (void)eA; (void)eB;
            etile += sd * pair_energy_d<ELEC, VDW>(d, qid * qjd, ljd.x, ljd.y, rd);
          }
          float fr;
          if (SOFT && alchPair) fr = f * dredr * __frsqrt_rn(r2);
          else fr = f * rinv;
          const float gx = fr * dx, gy = fr * dy, gz = fr * dz;
          fx += gx; fy += gy; fz += gz;
          atomicAdd(&sf[jl].x, -gx);
          atomicAdd(&sf[jl].y, -gy);
          atomicAdd(&sf[jl].z, -gz);
        }
      }
      if (ENERGY) eacc += etile;
      fix_x += fix_f(fx);
      fix_y += fix_f(fy);
      fix_z += fix_f(fz);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (hv) {
        const int sj = J * kBlockAtoms + li;
        const int aj = d.l2g[sj];
        if (aj >= 0) {
          const float4 fj = sf[li];
          atomic_fix_f(&d.force[aj], fj.x);
          atomic_fix_f(&d.force[d.n + aj], fj.y);
          atomic_fix_f(&d.force[2 * d.n + aj], fj.z);
          const int bj = spr[li].y;
          if (ALCH && (bj & 0xFFFF) && fj.w != 0.f) atomic_fix_f(&d.dudl[bj & 0xFFFF], fj.w);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (ENERGY) {
    long long v = (long long)__double2ll_rn(eacc * kFixScaleD);
    wave_atomic_fix(&d.energy[BM_ENBDIRECT], v);
  }
}

template <int ELEC, int VDW, bool ENERGY, bool ALCH, bool SOFT>
static void launch_t(const Dev &d, int nwg, hipStream_t st) {
  nbdirect_kernel<ELEC, VDW, ENERGY, ALCH, SOFT><<<nwg, 256, 0, st>>>(d);
}

template <int ELEC, int VDW>
static void launch_ev(const Dev &d, bool energy, bool alch, bool soft, int nwg, hipStream_t st) {
  if (!alch) {
    if (energy) launch_t<ELEC, VDW, true, false, false>(d, nwg, st);
    else launch_t<ELEC, VDW, false, false, false>(d, nwg, st);
  } else if (!soft) {
    if (energy) launch_t<ELEC, VDW, true, true, false>(d, nwg, st);
    else launch_t<ELEC, VDW, false, true, false>(d, nwg, st);
  } else {
    if (energy) launch_t<ELEC, VDW, true, true, true>(d, nwg, st);
    else launch_t<ELEC, VDW, false, true, true>(d, nwg, st);
  }
}

void launch_nbdirect(const Dev &d, bool energy, bool alch, bool soft, int nwg, hipStream_t st) {
  if (d.elecmethod == BM_PME) {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_PME, BM_VFSWITCH>(d, energy, alch, soft, nwg, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_PME, BM_VSWITCH>(d, energy, alch, soft, nwg, st);
    else launch_ev<BM_PME, BM_VSHIFT>(d, energy, alch, soft, nwg, st);
  } else if (d.elecmethod == BM_FSWITCH) {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_FSWITCH, BM_VFSWITCH>(d, energy, alch, soft, nwg, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_FSWITCH, BM_VSWITCH>(d, energy, alch, soft, nwg, st);
    else launch_ev<BM_FSWITCH, BM_VSHIFT>(d, energy, alch, soft, nwg, st);
  } else {
    if (d.vdwmethod == BM_VFSWITCH) launch_ev<BM_FSHIFT, BM_VFSWITCH>(d, energy, alch, soft, nwg, st);
    else if (d.vdwmethod == BM_VSWITCH) launch_ev<BM_FSHIFT, BM_VSWITCH>(d, energy, alch, soft, nwg, st);
    else launch_ev<BM_FSHIFT, BM_VSHIFT>(d, energy, alch, soft, nwg, st);
  }
}

}  // namespace blade
