// bonded.hip - bonds (+Urey-Bradley), angles, dihedrals, impropers, 1-4
// pairs, PME exclusion corrections and MSLD lambda biases in one launch
// (restates bonded/bonded.cu:44-527, bonded/pair.cu:56-387 and
// msld/msld.cu:1000-1130).  Each workgroup handles one term family so the
// energy slot is uniform.  Every term is FP64 (geometry from the FP64
// positions, minimum image by rounding): bonded energies enter dU/dlambda
// directly, and FP32 term energies summed over a site's terms miss the
// 1e-3 kcal/mol dU/dlambda bar.  Forces go to per-term slots, not to the
// force array: no global atomics (the force assembly sums them per atom).
#include "kernels.h"
#include "erfc_coef.h"

namespace blade {

struct D3v { double x, y, z; };
// FP64 minimum-image displacement r_i - r_j
__device__ __forceinline__ D3v dispd(const Dev &d, int i, int j) {
  D3v r;
  const double *p = d.pos;
  r.x = p[i] - p[j];
  r.y = p[d.n + i] - p[d.n + j];
  r.z = p[2 * d.n + i] - p[2 * d.n + j];
  r.x -= d.box[0] * floor(r.x * d.invbox[0] + 0.5);
  r.y -= d.box[1] * floor(r.y * d.invbox[1] + 0.5);
  r.z -= d.box[2] * floor(r.z * d.invbox[2] + 0.5);
  return r;
}
// per-atom force of a term into its slot (fixed point); the force
// assembly adds every atom's slots
__device__ __forceinline__ void put_d(const Dev &d, int slot, double s, D3v r) {
  const int p = d.bspos[slot];
  d.bslot[p] = fix_d(s * r.x);
  d.bslot[d.nslot + p] = fix_d(s * r.y);
  d.bslot[2 * d.nslot + p] = fix_d(s * r.z);
}
__device__ __forceinline__ D3v dsub(D3v a, D3v b) { return D3v{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3v dadd(D3v a, D3v b) { return D3v{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3v dscale(double s, D3v a) { return D3v{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double ddot(D3v a, D3v b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3v dcross(D3v a, D3v b) {
  return D3v{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ void put_v(const Dev &d, int slot, D3v f) { put_d(d, slot, 1.0, f); }

// lambda scaling of a term carrying two siteBlock codes (bonded.cu:80-89),
// and dU/dlambda of its energy e (FP64)
struct Sd { int b0, b1; double l0, l1; };
__device__ __forceinline__ Sd scale2d(const Dev &d, int2 sb) {
  Sd s{sb.x & 0xFFFF, sb.y & 0xFFFF, 1.0, 1.0};
  if (s.b0) {
    s.l0 = d.lambda[s.b0];
    if (s.b1) s.l1 = d.lambda[s.b1];
  }
  return s;
}
__device__ __forceinline__ void scale_dudl_d(const Dev &d, const Sd &s, double e) {
  if (s.b0) {
    atomic_fix(&d.dudl[s.b0], fix_d(s.l1 * e));
    if (s.b1) atomic_fix(&d.dudl[s.b1], fix_d(s.l0 * e));
  }
}
// lambda weight of a term's energy E0, E = w E0 with P = l0 l1: w = P, or
// P^q for a soft angle, dihedral or improper (bonded.cu:239-256,408-425);
// dw = dw/dP
struct Wd { double w, dw; };
__device__ __forceinline__ Wd weight_d(const Dev &d, const Sd &s, bool soft) {
  const double P = s.l0 * s.l1;
  if (!soft) return Wd{P, 1.0};
  return Wd{pow(P, d.snb_exp), d.snb_exp * pow(P, d.snb_exp - 1.0)};
}

__device__ __forceinline__ float erfc_pair(float x) {
  float ex = __expf(-x * x);
  float t = __frcp_rn(fmaf(BM_ERFC_P, x, 1.0f));
  float p = bm_erfc_c[BM_ERFC_DEG];
#pragma unroll
  for (int k = BM_ERFC_DEG - 1; k >= 0; k--) p = fmaf(p, t, bm_erfc_c[k]);
  return ex * p;
}

// 1-4 pair energy/derivative (pair.cu:56-175); VFSWITCH energy below the
// switch is shifted by rc^-6/rc^-12 exactly as the reference does
__device__ void nb14_ef(const Dev &d, double4 p, double r, double &f, double &e) {
  const double rc = d.rc, rs = d.rs;
  f = 0.0; e = 0.0;
  if (!(r < rc)) return;
  const double qq = p.x, c12 = p.y, c6 = p.z, e14 = p.w;
  double rinv = 1.0 / r, rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  if (d.vdwmethod == BM_VFSWITCH) {
    double rc3 = rc * rc * rc, rc6 = rc3 * rc3, rs3 = rs * rs * rs, rs6 = rs3 * rs3;
    double k6 = rc3 / (rc3 - rs3), k12 = rc6 / (rc6 - rs6);
    bool in = r < rs;
    double A6 = in ? 1.0 : k6, A12 = in ? 1.0 : k12, B6 = in ? 0.0 : 1.0 / rc3, B12 = in ? 0.0 : 1.0 / rc6;
    f = 6 * c6 * A6 * (rinv3 - B6) * rinv3 * rinv - 12 * c12 * A12 * (rinv6 - B12) * rinv6 * rinv;
    double CC6 = in ? -1.0 / rc6 : 0.0, CC12 = in ? -1.0 / (rc6 * rc6) : 0.0;
    e = c12 * (A12 * (rinv6 - B12) * (rinv6 - B12) + CC12) - c6 * (A6 * (rinv3 - B6) * (rinv3 - B6) + CC6);
  } else if (d.vdwmethod == BM_VSWITCH) {
    double c2of = rc * rc, c2on = rs * rs;
    double rul3 = (c2of - c2on) * (c2of - c2on) * (c2of - c2on), rul12 = 12 / rul3;
    double rijl = c2on - r * r, riju = c2of - r * r;
    double fsw = (r < rs) ? 1.0 : riju * riju * (riju - 3 * rijl) / rul3;
    double dfsw = (r < rs) ? 0.0 : rijl * riju * rul12;
    f = fsw * (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv + dfsw * (c12 * rinv6 - c6) * rinv6;
    e = fsw * (c12 * rinv6 - c6) * rinv6;
  } else {
    double r2 = r * r, r5 = r2 * r2 * r, r6 = r2 * r2 * r2;
    double rinv7 = rinv6 * rinv, rinv12 = rinv6 * rinv6, rinv13 = rinv12 * rinv;
    double rc2 = rc * rc, rc6 = rc2 * rc2 * rc2, rc12 = rc6 * rc6, rc18 = rc12 * rc6;
    f = 6 * c6 * (rinv7 - r5 / rc12) - 12 * c12 * (rinv13 + r5 / rc18);
    e = c12 * (rinv12 + 2 * r6 / rc18 - 3 / rc12) - c6 * (rinv6 + r6 / rc12 - 2 / rc6);
  }
  const double kqq = kElectricD * qq;
  if (d.elecmethod == BM_FSWITCH) {
    double kqqe = kqq * e14;
    double roff2 = rc * rc, ron2 = rs * rs;
    double ginv = 1 / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
    double A = roff2 * roff2 * (roff2 - 3 * ron2) * ginv, B = 6 * roff2 * ron2 * ginv;
    double C = -(ron2 + roff2) * ginv, D = 2 * ginv / 5;
    double dvc = 8 * (ron2 * roff2 * (rc - rs) - (roff2 * roff2 * rc - ron2 * ron2 * rs) / 5) * ginv;
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    if (r <= rs) {
      f += -kqqe * rinv * rinv;
      e += kqqe * (rinv + dvc);
    } else {
      f += -kqqe * rinv * (A * rinv + B * r + 3 * C * r3 + 5 * D * r5);
      e += kqqe * (A * (rinv - 1 / rc) + B * (rc - r) + C * (roff2 * rc - r3) + D * (roff2 * roff2 * rc - r5));
    }
  } else if (d.elecmethod == BM_PME) {
    double br = d.betad * r;
    double erfcrinv = (erfc(br) + e14 - 1.0) * rinv;
    f += kqq * (-erfcrinv - 1.1283791670955126 * d.betad * exp(-br * br)) * rinv;
    e += kqq * erfcrinv;
  } else {
    double roffinv = 1 / rc, roffinv2 = roffinv * roffinv, kqqe = kqq * e14;
    f += -kqqe * (rinv * rinv - roffinv2);
    e += kqqe * (rinv - 2 * roffinv + r * roffinv2);
  }
}

// Theta biases of site s in FP64 (msld/msld.cu:1155-1264): dU/dtheta of
// each of the site's blocks into the dtheta accumulators, the energy
// returned.  Collective: A = sum h_j^n - 1, B = sum l_j^n - (nb - 1) with
// h = (1 + sin theta)/2, l = (1 - sin theta)/2, E = k/2 (A^2 + B^2) and the
// reference's force k n (A h^(n-1) - B l^(n-1)) cos theta (twice the theta
// derivative of that energy, kept as the reference has it).  Independent:
// E = k (1 - l^4), force 2 k l^3 cos theta.  Piecewise constraint (sites
// after the environment): a flat-bottom harmonic beyond +-nb*width, and
// with |edge_k| > 1e-4 the edge bias on N = sum sigmoid(alpha (theta_j -
// t_s + phi)) above its target, whose theta_b derivative carries the
// root's shift g_b / G (g = 3 max(theta - t_s, 0)^2).
__device__ double theta_bias_site(const Dev &d, int s) {
  const int b0 = d.site_bound[s], b1 = d.site_bound[s + 1], nb = b1 - b0;
  const double *th = d.pos + 3 * (size_t)d.n;
  double e = 0.0;
  if (d.tcoll_k && d.tcoll_k[s] != 0.0) {
    const double k = d.tcoll_k[s], n = d.tcoll_n[s];
    double A = -1.0, B = -(double)(nb - 1);
    for (int b = b0; b < b1; b++) {
      const double sn = sin(th[b]);
      A += pow(0.5 + 0.5 * sn, n);
      B += pow(0.5 - 0.5 * sn, n);
    }
    e += 0.5 * k * (A * A + B * B);
    for (int b = b0; b < b1; b++) {
      double sn, cs;
      sincos(th[b], &sn, &cs);
      const double f = k * n * (A * pow(0.5 + 0.5 * sn, n - 1) - B * pow(0.5 - 0.5 * sn, n - 1)) * cs;
      atomic_fix(&d.dtheta[b], fix_d(f));
    }
  }
  if (d.tinde_k && d.tinde_k[s] != 0.0) {
    const double k = d.tinde_k[s];
    for (int b = b0; b < b1; b++) {
      double sn, cs;
      sincos(th[b], &sn, &cs);
      const double l = 0.5 - 0.5 * sn, l3 = l * l * l;
      e += k * (1.0 - l3 * l);
      atomic_fix(&d.dtheta[b], fix_d(2.0 * k * l3 * cs));
    }
  }
  if (d.piecewise && s > 0) {
    const double w = nb * d.pw_width;
    for (int b = b0; b < b1; b++) {
      double x = 0.0;
      if (th[b] > w) x = th[b] - w;
      else if (th[b] < -w) x = th[b] + w;
      e += 0.5 * d.pw_k * x * x;
      if (x != 0.0) atomic_fix(&d.dtheta[b], fix_d(d.pw_k * x));
    }
    if (fabs(d.edge_k) > 1e-4) {
      const double t0 = d.theta0[s];
      double G = 0.0, N = 0.0, ssum = 0.0;
      for (int b = b0; b < b1; b++) {
        const double q = th[b] - t0;
        G += q > 0.0 ? 3.0 * q * q : 0.0;
        const double x = exp(-d.edge_alpha * (q + d.edge_phi));
        N += 1.0 / (1.0 + x);
        ssum += d.edge_alpha * x / ((1.0 + x) * (1.0 + x));
      }
      const double over = N > d.edge_n ? N - d.edge_n : 0.0;
      e += 0.5 * d.edge_k * over * over;
      if (over != 0.0)
        for (int b = b0; b < b1; b++) {
          const double q = th[b] - t0, gb = q > 0.0 ? 3.0 * q * q : 0.0;
          const double x = exp(-d.edge_alpha * (q + d.edge_phi));
          const double sg = d.edge_alpha * x / ((1.0 + x) * (1.0 + x));
          atomic_fix(&d.dtheta[b], fix_d(d.edge_k * over * (sg - (gb / G) * ssum)));
        }
    }
  }
  return e;
}

struct BondedLaunch {
  // first block of each family, then the bias workgroup
  int b_bond, b_angle, b_dihe, b_impr, b_nb14, b_nbex, b_cats, b_cmap;
  int b_noe, b_harm, b_bor, b_anr, b_dir, b_resd, b_bias;
  unsigned fam;   // families to evaluate (kFam*); the others zero their slots
};

// workgroups of 256 threads per term family, in family order; the last
// workgroup carries the Ewald self term and the lambda and theta biases
__host__ __device__ inline BondedLaunch bonded_layout(const Dev &d, unsigned fam = kFamAll) {
  auto nb = [](int n) { return (n + 255) / 256; };
  BondedLaunch L;
  L.b_bond = 0;
  L.b_angle = L.b_bond + nb(d.nbond);
  L.b_dihe = L.b_angle + nb(d.nangle);
  L.b_impr = L.b_dihe + nb(d.ndihe);
  L.b_nb14 = L.b_impr + nb(d.nimpr);
  L.b_nbex = L.b_nb14 + nb(d.nnb14);
  L.b_cats = L.b_nbex + nb(d.nnbex);
  L.b_cmap = L.b_cats + nb(d.ncats);
  L.b_noe = L.b_cmap + nb(d.ncmap);
  L.b_harm = L.b_noe + nb(d.nnoe);
  L.b_bor = L.b_harm + nb(d.nharm);
  L.b_anr = L.b_bor + nb(d.nbor);
  L.b_dir = L.b_anr + nb(d.nanr);
  L.b_resd = L.b_dir + nb(d.ndir);
  L.b_bias = L.b_resd + nb(d.nresd);
  L.fam = fam;
  return L;
}

__device__ __forceinline__ void zero_slots(const Dev &d, int first, int count) {
  for (int k = 0; k < count; k++) put_d(d, first + k, 0.0, D3v{0.0, 0.0, 0.0});
}

// Restraint terms in FP64 (restrain/restrain.cu:15-598): term t of the
// family starting at block b0 of the launch; returns the energy.
__device__ double restraint_term(const Dev &d, const BondedLaunch &L, int bid, int t) {
  double e = 0.0;
  if (bid < L.b_harm) {
    // NOE: flat-bottom harmonic on a pair distance (or the distance to a
    // fixed point), optionally switched to a power-law tail
    if (t >= d.nnoe) return 0.0;
    const int2 ij = d.noe_i[t];
    const double *p = d.noe_p + 10 * (size_t)t;
    D3v dr;
    if (ij.y >= 0) {
      dr = dispd(d, ij.x, ij.y);
    } else {
      dr = D3v{d.pos[ij.x] - p[7], d.pos[d.n + ij.x] - p[8], d.pos[2 * d.n + ij.x] - p[9]};
      dr.x -= d.box[0] * floor(dr.x * d.invbox[0] + 0.5);
      dr.y -= d.box[1] * floor(dr.y * d.invbox[1] + 0.5);
      dr.z -= d.box[2] * floor(dr.z * d.invbox[2] + 0.5);
    }
    const double r = sqrt(ddot(dr, dr));
    double f = 0.0;
    if (r < p[0]) {
      f = p[1] * (r - p[0]);
      e = 0.5 * f * (r - p[0]);
    } else if (r > p[2]) {
      const double dd = r - p[2], rpk = p[4], rsw = p[5], nsw = p[6];
      if (rsw > 0 && dd > rsw) {
        const double b = (rpk - rsw) / nsw * pow(rsw, nsw + 1);
        const double a = 0.5 * rsw * rsw - rpk * rsw - rsw * (rpk - rsw) / nsw;
        f = p[3] * (rpk - b * pow(dd, -nsw - 1));
        e = p[3] * (a + b * pow(dd, -nsw) + rpk * dd);
      } else {
        f = p[3] * dd;
        e = 0.5 * f * dd;
      }
    }
    const double g = r > 0 ? f / r : 0.0;
    put_d(d, d.sl_noe + 2 * t, g, dr);
    put_d(d, d.sl_noe + 2 * t + 1, ij.y >= 0 ? -g : 0.0, dr);
  } else if (bid < L.b_bor) {
    // harmonic position restraint k |x - x0|^n
    if (t >= d.nharm) return 0.0;
    const int i = d.harm_i[t];
    const double *p = d.harm_p + 5 * (size_t)t;
    D3v dr{d.pos[i] - p[1], d.pos[d.n + i] - p[2], d.pos[2 * d.n + i] - p[3]};
    dr.x -= d.box[0] * floor(dr.x * d.invbox[0] + 0.5);
    dr.y -= d.box[1] * floor(dr.y * d.invbox[1] + 0.5);
    dr.z -= d.box[2] * floor(dr.z * d.invbox[2] + 0.5);
    const double r2 = ddot(dr, dr);
    const double krn2 = r2 > 0 ? p[0] * pow(r2, 0.5 * p[4] - 1) : 0.0;
    put_d(d, d.sl_harm + t, p[4] * krn2, dr);
    e = krn2 * r2;
  } else if (bid < L.b_anr) {
    // bond restraint kr/2 (r - r0)^2 scaled by lambda of its block
    if (t >= d.nbor) return 0.0;
    const int4 ix = d.bor_i[t];
    const double2 p = d.bor_p[t];
    const double l = ix.z ? d.lambda[ix.z] : 1.0;
    const D3v dr = dispd(d, ix.x, ix.y);
    const double r = sqrt(ddot(dr, dr)), e0 = 0.5 * p.x * (r - p.y) * (r - p.y), f = p.x * (r - p.y) * l;
    if (ix.z) atomic_fix(&d.dudl[ix.z], fix_d(e0));
    put_d(d, d.sl_bor + 2 * t, f / r, dr);
    put_d(d, d.sl_bor + 2 * t + 1, -f / r, dr);
    e = e0 * l;
  } else if (bid < L.b_dir) {
    // angle restraint kt/2 (theta - t0)^2 scaled by lambda of its block
    if (t >= d.nanr) return 0.0;
    const int4 ix = d.anr_i[t];
    const int b = d.anr_b[t];
    const double2 p = d.anr_p[t];
    const double l = b ? d.lambda[b] : 1.0;
    const D3v rij = dispd(d, ix.x, ix.y), rkj = dispd(d, ix.z, ix.y), c = dcross(rij, rkj);
    const double mc = sqrt(ddot(c, c)), th = atan2(mc, ddot(rij, rkj));
    const double e0 = 0.5 * p.x * (th - p.y) * (th - p.y), f = p.x * (th - p.y) * l;
    if (b) atomic_fix(&d.dudl[b], fix_d(e0));
    const D3v fi = dscale(f / (mc * ddot(rij, rij)), dcross(rij, c));
    const D3v fk = dscale(-f / (mc * ddot(rkj, rkj)), dcross(rkj, c));
    put_v(d, d.sl_anr + 3 * t, fi);
    put_v(d, d.sl_anr + 3 * t + 2, fk);
    put_v(d, d.sl_anr + 3 * t + 1, dscale(-1.0, dadd(fi, fk)));
    e = e0 * l;
  } else if (bid < L.b_resd) {
    // dihedral restraint: periodic kphi (1 - cos(n (phi - phi0))), or a
    // flat-bottom harmonic in the wrapped phi - phi0; phi from
    // m = (j - i) x (k - j) and n = (k - j) x (l - k)
    if (t >= d.ndir) return 0.0;
    const int4 ix = d.dir_i[t];
    const int2 bn = d.dir_bn[t];
    const double4 p = d.dir_p[t];
    const double l = bn.x ? d.lambda[bn.x] : 1.0;
    const D3v rij = dispd(d, ix.y, ix.x), rjk = dispd(d, ix.z, ix.y), rkl = dispd(d, ix.w, ix.z);
    const D3v m = dcross(rij, rjk), nv = dcross(rjk, rkl);
    const double mm = sqrt(ddot(m, m)), nn = sqrt(ddot(nv, nv)), bb = sqrt(ddot(rjk, rjk));
    const double phi = atan2(ddot(dscale(1.0 / bb, rjk), dcross(dscale(1.0 / mm, m), dscale(1.0 / nn, nv))),
                             ddot(m, nv) / (mm * nn));
    const double twopi = 6.283185307179586, pi = 3.141592653589793;
    double fphi, e0;
    if (bn.y > 0) {
      double dp = bn.y * (phi - p.y);
      dp -= twopi * floor((dp + pi) / twopi);
      fphi = p.x * bn.y * sin(dp);
      e0 = p.x * (1.0 - cos(dp));
    } else {
      double dp = phi - p.y;
      dp -= twopi * floor((dp + pi) / twopi);
      if (p.z > 0) dp = fabs(dp) <= p.z ? 0.0 : dp - copysign(p.z, dp);
      fphi = p.x * dp;
      e0 = 0.5 * p.x * dp * dp;
    }
    if (bn.x) atomic_fix(&d.dudl[bn.x], fix_d(e0));
    fphi *= l;
    const D3v fi = dscale(-fphi * bb / (mm * mm), m), fk = dscale(-fphi * bb / (nn * nn), nv);
    const double pp = ddot(rij, rjk) / (bb * bb), qq = ddot(rkl, rjk) / (bb * bb);
    const D3v fj = dadd(dscale(-pp, fi), dscale(-qq, fk));
    put_v(d, d.sl_dir + 4 * t, fi);
    put_v(d, d.sl_dir + 4 * t + 3, dscale(-1.0, fk));
    put_v(d, d.sl_dir + 4 * t + 2, dsub(fk, fj));
    put_v(d, d.sl_dir + 4 * t + 1, dsub(fj, fi));
    e = e0 * l;
  } else {
    // RESD: kdist/2 (ci r1 + cj r2 - rdist)^2
    if (t >= d.nresd) return 0.0;
    const int4 ix = d.resd_i[t];
    const double4 p = d.resd_p[t];
    const D3v d1 = dispd(d, ix.x, ix.y), d2 = dispd(d, ix.z, ix.w);
    const double r1 = sqrt(ddot(d1, d1)), r2 = sqrt(ddot(d2, d2));
    const double del = p.x * r1 + p.y * r2 - p.z, f = del * p.w;
    put_d(d, d.sl_resd + 4 * t, f * p.x / r1, d1);
    put_d(d, d.sl_resd + 4 * t + 1, -f * p.x / r1, d1);
    put_d(d, d.sl_resd + 4 * t + 2, f * p.y / r2, d2);
    put_d(d, d.sl_resd + 4 * t + 3, -f * p.y / r2, d2);
    e = 0.5 * f * del;
  }
  return e;
}

// Family of workgroup bid: its mask bit, energy term, first slot and slots
// per term (for zeroing when the family is off)
struct FamInfo { unsigned bit; int eterm, first, width, count, b0; };
__device__ __forceinline__ FamInfo family_of(const Dev &d, const BondedLaunch &L, int bid) {
  if (bid < L.b_angle) return FamInfo{kFamBond, BM_EBOND, d.sl_bond, 2, d.nbond, L.b_bond};
  if (bid < L.b_dihe) return FamInfo{kFamAngle, BM_EANGLE, d.sl_angle, 3, d.nangle, L.b_angle};
  if (bid < L.b_impr) return FamInfo{kFamDihe, BM_EDIHE, d.sl_dihe, 4, d.ndihe, L.b_dihe};
  if (bid < L.b_nb14) return FamInfo{kFamImpr, BM_EIMPR, d.sl_impr, 4, d.nimpr, L.b_impr};
  if (bid < L.b_nbex) return FamInfo{kFamNb14, BM_ENB14, d.sl_nb14, 2, d.nnb14, L.b_nb14};
  if (bid < L.b_cats) return FamInfo{kFamNbex, BM_ENBRECIPEXCL, d.sl_nbex, 2, d.nnbex, L.b_nbex};
  if (bid < L.b_cmap) return FamInfo{kFamCats, BM_ECATS, d.sl_cats, 0, d.ncats, L.b_cats};
  if (bid < L.b_noe) return FamInfo{kFamCmap, BM_ECMAP, d.sl_cmap, 8, d.ncmap, L.b_cmap};
  if (bid < L.b_harm) return FamInfo{kFamRestrain, BM_ENOE, d.sl_noe, 2, d.nnoe, L.b_noe};
  if (bid < L.b_bor) return FamInfo{kFamRestrain, BM_EHARMONIC, d.sl_harm, 1, d.nharm, L.b_harm};
  if (bid < L.b_anr) return FamInfo{kFamRestrain, BM_EMMFP, d.sl_bor, 2, d.nbor, L.b_bor};
  if (bid < L.b_dir) return FamInfo{kFamRestrain, BM_EMMFP, d.sl_anr, 3, d.nanr, L.b_anr};
  if (bid < L.b_resd) return FamInfo{kFamRestrain, BM_EMMFP, d.sl_dir, 4, d.ndir, L.b_dir};
  if (bid < L.b_bias) return FamInfo{kFamRestrain, BM_ERESD, d.sl_resd, 4, d.nresd, L.b_resd};
  return FamInfo{kFamSelf | kFamBias, BM_ELAMBDA, 0, 0, 0, L.b_bias};
}

__device__ __forceinline__ void bonded_body(const Dev &d, const BondedLaunch &L, int bid, bool energy) {
  double e = 0.0;
  int slot = -1;
  const FamInfo fi = family_of(d, L, bid);
  if (!(fi.bit & L.fam)) {
    // a family left out writes zero forces to its slots (the assembly adds
    // every slot) and no energy
    const int t = (bid - fi.b0) * blockDim.x + threadIdx.x;
    if (fi.width == 0 && fi.bit == kFamCats) {
      if (t < d.ncats) zero_slots(d, d.sl_cats + d.cats_bound[t], d.cats_bound[t + 1] - d.cats_bound[t]);
    } else if (t < fi.count) {
      zero_slots(d, fi.first + fi.width * t, fi.width);
    }
    return;
  }
  if (bid >= L.b_noe && bid < L.b_bias) {
    const int t = (bid - fi.b0) * blockDim.x + threadIdx.x;
    e = restraint_term(d, L, bid, t);
    slot = fi.eterm;
  } else if (bid < L.b_angle) {
    const int t = (bid - L.b_bond) * blockDim.x + threadIdx.x;
    if (t < d.nbond) {
      int2 ij = d.bond_i[t];
      const D3v dr = dispd(d, ij.x, ij.y);
      const double r = sqrt(ddot(dr, dr));
      const Sd s = scale2d(d, d.bond_sb[t]);
      const double2 p = d.bond_p[t];
      const double db = r - p.y;
      const double et = 0.5 * p.x * db * db;
      double fb;
      if (d.bond_soft && d.bond_soft[t]) {
        // soft bond (bonded.cu:81-107): Q = P^p, E = Q E0 g with
        // g = 1 / (1 + (1 - Q) alpha db^2)
        const double P = s.l0 * s.l1, Q = pow(P, d.sb_pow);
        const double g = 1.0 / (1.0 + (1.0 - Q) * d.sb_alpha * db * db);
        const double dgdr = -2.0 * g * g * (1.0 - Q) * d.sb_alpha * db;
        const double dgdQ = g * g * d.sb_alpha * db * db;
        fb = Q * (p.x * db * g + et * dgdr);
        scale_dudl_d(d, s, et * d.sb_pow * pow(P, d.sb_pow - 1.0) * (g + Q * dgdQ));
        e = Q * et * g;
      } else {
        fb = p.x * db * s.l0 * s.l1;
        scale_dudl_d(d, s, et);
        e = et * s.l0 * s.l1;
      }
      put_d(d, d.sl_bond + 2 * t, fb / r, dr);
      put_d(d, d.sl_bond + 2 * t + 1, -fb / r, dr);
      // bonds and Urey-Bradley may share a workgroup: split by index
      if (t >= d.nbond12) {
        if (energy) atomic_fix(&d.energy[BM_EUREY], fix_d(e));
        e = 0.0;
      }
    }
    slot = BM_EBOND;
  } else if (bid < L.b_dihe) {
    slot = BM_EANGLE;
    const int t = (bid - L.b_angle) * blockDim.x + threadIdx.x;
    if (t < d.nangle) {
      int4 ix = d.angle_i[t];
      const D3v rij = dispd(d, ix.x, ix.y), rkj = dispd(d, ix.z, ix.y);
      const double dp = ddot(rij, rkj);
      const D3v cr = dcross(rij, rkj);
      const double mc = sqrt(ddot(cr, cr));
      const double th = atan2(mc, dp);
      const Sd s = scale2d(d, d.angle_sb[t]);
      const Wd wt = weight_d(d, s, d.angle_soft && d.angle_soft[t]);
      const double2 p = d.angle_p[t];
      const double dth = th - p.y;
      const double et = 0.5 * p.x * dth * dth;
      const double fa = p.x * dth * wt.w;
      scale_dudl_d(d, s, wt.dw * et);
      const D3v fi = dscale(fa / (mc * ddot(rij, rij)), dcross(rij, cr));
      const D3v fk = dscale(-fa / (mc * ddot(rkj, rkj)), dcross(rkj, cr));
      put_v(d, d.sl_angle + 3 * t, fi);
      put_v(d, d.sl_angle + 3 * t + 2, fk);
      put_v(d, d.sl_angle + 3 * t + 1, dscale(-1.0, dadd(fi, fk)));
      e = et * wt.w;
    }
  } else if (bid < L.b_nb14) {
    const bool impr = bid >= L.b_impr;
    slot = impr ? BM_EIMPR : BM_EDIHE;
    const int t = (bid - (impr ? L.b_impr : L.b_dihe)) * blockDim.x + threadIdx.x;
    if (t < (impr ? d.nimpr : d.ndihe)) {
      int4 ix = impr ? d.impr_i[t] : d.dihe_i[t];
      const double4 p = impr ? d.impr_p[t] : d.dihe_p[t];
      const D3v rij = dispd(d, ix.x, ix.y), rjk = dispd(d, ix.y, ix.z), rkl = dispd(d, ix.z, ix.w);
      const D3v m = dcross(rij, rjk), nv = dcross(rjk, rkl);
      const D3v mn = dcross(m, nv);
      double phi = atan2(sqrt(ddot(mn, mn)), ddot(m, nv));
      if (ddot(rij, nv) > 0.0) phi = -phi;  // bonded.cu:372-374
      double fphi, et;
      const int nper = (int)p.y;
      if (!impr || nper > 0) {
        const double dphi = nper * phi - p.z;
        double sn, cs;
        sincos(dphi, &sn, &cs);
        fphi = -p.x * nper * sn;
        et = p.x * (cs + 1.0);
      } else {
        double dphi = phi - p.z;
        dphi -= 6.283185307179586 * floor((dphi + 3.141592653589793) / 6.283185307179586);
        fphi = p.x * dphi;
        et = 0.5 * p.x * dphi * dphi;
      }
      const Sd s = scale2d(d, impr ? d.impr_sb[t] : d.dihe_sb[t]);
      const Wd wt = weight_d(d, s, p.w != 0.0);
      fphi *= wt.w;
      scale_dudl_d(d, s, wt.dw * et);
      const double rjk2 = ddot(rjk, rjk), r = sqrt(rjk2);
      const D3v fi = dscale(-fphi * r / ddot(m, m), m);
      const D3v fl = dscale(-fphi * r / ddot(nv, nv), nv);
      const double pp = ddot(rij, rjk) / rjk2, qq = ddot(rkl, rjk) / rjk2;
      const D3v fj = dadd(dscale(-pp, fi), dscale(-qq, fl));
      const int s0 = (impr ? d.sl_impr : d.sl_dihe) + 4 * t;
      put_v(d, s0, fi);
      put_v(d, s0 + 3, dscale(-1.0, fl));
      put_v(d, s0 + 2, dsub(fl, fj));
      put_v(d, s0 + 1, dsub(fj, fi));
      e = et * wt.w;
    }
  } else if (bid < L.b_nbex) {
    slot = BM_ENB14;
    const int t = (bid - L.b_nb14) * blockDim.x + threadIdx.x;
    // 1-4 pairs in FP64: few terms whose energies are large and cancel
    // against the PME terms, so FP32 rounding of charges would bias them
    if (t < d.nnb14) {
      int2 ij = d.nb14_i[t];
      D3v dr = dispd(d, ij.x, ij.y);
      double r = sqrt(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
      const Sd s = scale2d(d, d.nb14_sb[t]);
      const double ll = (s.b0 ? d.lambda[s.b0] : 1.0) * (s.b1 ? d.lambda[s.b1] : 1.0);
      double reff = r, dredr = 1.0, dredll = 0.0;
      if (d.softcore14 && s.b0) {
        double rSoft = 4.0 * (1.0 - ll);
        if (r < rSoft) {
          double q = r / rSoft;
          double re = 1.0 - 0.5 * q;
          re = re * q * q * q + 0.5;
          dredr = (3.0 - 2.0 * q) * q * q;
          dredll = (re - dredr * q) * (-4.0);
          reff = re * rSoft;
        }
      }
      double f, et;
      nb14_ef(d, d.nb14_p[t], reff, f, et);
      f *= ll;
      if (s.b0) {
        const double g = et + f * dredll;
        atomic_fix(&d.dudl[s.b0], fix_d((s.b1 ? d.lambda[s.b1] : 1.0) * g));
        if (s.b1) atomic_fix(&d.dudl[s.b1], fix_d(d.lambda[s.b0] * g));
      }
      f *= dredr;
      put_d(d, d.sl_nb14 + 2 * t, (f / r), dr);
      put_d(d, d.sl_nb14 + 2 * t + 1, (-f / r), dr);
      e = et * ll;
    }
  } else if (bid < L.b_cats) {
    slot = BM_ENBRECIPEXCL;
    const int t = (bid - L.b_nbex) * blockDim.x + threadIdx.x;
    // PME exclusion corrections in FP64 (they cancel the self term)
    if (t < d.nnbex) {
      int2 ij = d.nbex_i[t];
      D3v dr = dispd(d, ij.x, ij.y);
      const double r = sqrt(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
      const double kqq = kElectricD * d.nbex_p[t], beta = d.betad, x = beta * r;
      // fr = (dE/dr)/r; for beta r < 0.01 E and fr from their series about
      // r = 0 (substituent atoms of one site may coincide: 0/0 otherwise)
      double fr, et;
      if (x < 0.01) {
        const double x2 = x * x;
        et = -kqq * beta * 1.1283791670955126 * (1 - x2 / 3 + x2 * x2 / 10 - x2 * x2 * x2 / 42);
        fr = kqq * beta * beta * beta * 1.1283791670955126 * (2.0 / 3 - 2 * x2 / 5 + x2 * x2 / 7);
      } else {
        const double rinv = 1.0 / r, erfv = erf(x);
        et = -kqq * erfv * rinv;
        fr = kqq * (erfv * rinv - 1.1283791670955126 * beta * exp(-x * x)) * rinv * rinv;
      }
      const Sd s = scale2d(d, d.nbex_sb[t]);
      const double ll = (s.b0 ? d.lambda[s.b0] : 1.0) * (s.b1 ? d.lambda[s.b1] : 1.0);
      fr *= ll;
      if (s.b0) {
        atomic_fix(&d.dudl[s.b0], fix_d((s.b1 ? d.lambda[s.b1] : 1.0) * et));
        if (s.b1) atomic_fix(&d.dudl[s.b1], fix_d(d.lambda[s.b0] * et));
      }
      put_d(d, d.sl_nbex + 2 * t, fr, dr);
      put_d(d, d.sl_nbex + 2 * t + 1, -fr, dr);
      e = et * ll;
    }
  } else if (bid < L.b_cmap) {
    // MSLD atom restraints (msld.cu:1328-1364), one thread per group in
    // FP64: the centroid from minimum-image offsets to the first atom,
    // dU/dx_i = k (x_i - centroid), not lambda-scaled
    slot = BM_ECATS;
    const int t = (bid - L.b_cats) * blockDim.x + threadIdx.x;
    if (t < d.ncats) {
      const int m0 = d.cats_bound[t], m1 = d.cats_bound[t + 1];
      const int a0 = d.cats_idx[m0];
      D3v c{0.0, 0.0, 0.0};
      for (int m = m0 + 1; m < m1; m++) {
        const D3v r = dispd(d, d.cats_idx[m], a0);
        c.x += r.x;
        c.y += r.y;
        c.z += r.z;
      }
      const double inv = 1.0 / (m1 - m0);
      c.x *= inv;
      c.y *= inv;
      c.z *= inv;
      for (int m = m0; m < m1; m++) {
        D3v r = m == m0 ? D3v{0.0, 0.0, 0.0} : dispd(d, d.cats_idx[m], a0);
        r.x -= c.x;
        r.y -= c.y;
        r.z -= c.z;
        r.x -= d.box[0] * floor(r.x * d.invbox[0] + 0.5);
        r.y -= d.box[1] * floor(r.y * d.invbox[1] + 0.5);
        r.z -= d.box[2] * floor(r.z * d.invbox[2] + 0.5);
        put_d(d, d.sl_cats + m, d.cats_k, r);
        e += 0.5 * d.cats_k * (r.x * r.x + r.y * r.y + r.z * r.z);
      }
    }
  } else if (bid < L.b_noe) {
    // CMAP cross terms (bonded.cu:531-720), one thread per term: the two
    // dihedrals as in the dihedral family, then bicubic Hermite
    // interpolation of the cell's corner values and slopes, all FP64
    slot = BM_ECMAP;
    const int t = (bid - L.b_cmap) * blockDim.x + threadIdx.x;
    if (t < d.ncmap) {
      const int4 sv = d.cmap_s[t];
      const int2 g = d.cmap_g[sv.w];
      const int ng = g.x;
      D3v rij[2], rjk[2], rkl[2], m[2], nv[2];
      double ang[2];
      for (int a = 0; a < 2; a++) {
        const int4 ix = d.cmap_i[2 * t + a];
        rij[a] = dispd(d, ix.x, ix.y);
        rjk[a] = dispd(d, ix.y, ix.z);
        rkl[a] = dispd(d, ix.z, ix.w);
        m[a] = dcross(rij[a], rjk[a]);
        nv[a] = dcross(rjk[a], rkl[a]);
        const D3v mn = dcross(m[a], nv[a]);
        double phi = atan2(sqrt(ddot(mn, mn)), ddot(m[a], nv[a]));
        if (ddot(rij[a], nv[a]) > 0.0) phi = -phi;
        ang[a] = phi;
      }
      // grid point i sits at -pi + i 2 pi / ng
      const double gs = ng / 6.283185307179586;
      const double u = ang[0] * gs, v = ang[1] * gs;
      const double fu = floor(u), fv = floor(v);
      const double tt = u - fu, ss = v - fv;
      const int i0 = (((int)fu + ng / 2) % ng + ng) % ng, j0 = (((int)fv + ng / 2) % ng + ng) % ng;
      const int i1 = i0 + 1 == ng ? 0 : i0 + 1, j1 = j0 + 1 == ng ? 0 : j0 + 1;
      const double4 *tb = d.cmap_tab + g.y;
      const double4 c00 = tb[ng * i0 + j0], c10 = tb[ng * i1 + j0], c01 = tb[ng * i0 + j1], c11 = tb[ng * i1 + j1];
      // cubic Hermite basis: value weights H, slope weights G, derivatives
      const double H0t = (1 + 2 * tt) * (1 - tt) * (1 - tt), H1t = tt * tt * (3 - 2 * tt);
      const double G0t = tt * (1 - tt) * (1 - tt), G1t = tt * tt * (tt - 1);
      const double dH0t = 6 * tt * (tt - 1), dH1t = -dH0t, dG0t = (1 - tt) * (1 - 3 * tt), dG1t = tt * (3 * tt - 2);
      const double H0s = (1 + 2 * ss) * (1 - ss) * (1 - ss), H1s = ss * ss * (3 - 2 * ss);
      const double G0s = ss * (1 - ss) * (1 - ss), G1s = ss * ss * (ss - 1);
      const double dH0s = 6 * ss * (ss - 1), dH1s = -dH0s, dG0s = (1 - ss) * (1 - 3 * ss), dG1s = ss * (3 * ss - 2);
      auto cell = [](double4 c, double ha, double ga, double hb, double gb) {
        return c.x * ha * hb + c.y * ga * hb + c.z * ha * gb + c.w * ga * gb;
      };
      const double ec = cell(c00, H0t, G0t, H0s, G0s) + cell(c10, H1t, G1t, H0s, G0s) +
                        cell(c01, H0t, G0t, H1s, G1s) + cell(c11, H1t, G1t, H1s, G1s);
      const double dt = cell(c00, dH0t, dG0t, H0s, G0s) + cell(c10, dH1t, dG1t, H0s, G0s) +
                        cell(c01, dH0t, dG0t, H1s, G1s) + cell(c11, dH1t, dG1t, H1s, G1s);
      const double ds = cell(c00, H0t, G0t, dH0s, dG0s) + cell(c10, H1t, G1t, dH0s, dG0s) +
                        cell(c01, H0t, G0t, dH1s, dG1s) + cell(c11, H1t, G1t, dH1s, dG1s);
      const int b0 = sv.x & 0xFFFF, b1 = sv.y & 0xFFFF, b2 = sv.z & 0xFFFF;
      const double l0 = b0 ? d.lambda[b0] : 1.0, l1 = b1 ? d.lambda[b1] : 1.0, l2 = b2 ? d.lambda[b2] : 1.0;
      const double ll = l0 * l1 * l2;
      if (b0) atomic_fix(&d.dudl[b0], fix_d(l1 * l2 * ec));
      if (b1) atomic_fix(&d.dudl[b1], fix_d(l0 * l2 * ec));
      if (b2) atomic_fix(&d.dudl[b2], fix_d(l0 * l1 * ec));
      for (int a = 0; a < 2; a++) {
        const double fphi = (a ? ds : dt) * gs * ll;
        const double rjk2 = ddot(rjk[a], rjk[a]), r = sqrt(rjk2);
        const D3v fi = dscale(-fphi * r / ddot(m[a], m[a]), m[a]);
        const D3v fl = dscale(-fphi * r / ddot(nv[a], nv[a]), nv[a]);
        const double pp = ddot(rij[a], rjk[a]) / rjk2, qq = ddot(rkl[a], rjk[a]) / rjk2;
        const D3v fj = dadd(dscale(-pp, fi), dscale(-qq, fl));
        const int s0 = d.sl_cmap + 8 * t + 4 * a;
        put_v(d, s0, fi);
        put_v(d, s0 + 3, dscale(-1.0, fl));
        put_v(d, s0 + 2, dsub(fl, fj));
        put_v(d, s0 + 1, dsub(fj, fi));
      }
      e = ec * ll;
    }
  } else {
    // MSLD fixed and variable biases: one thread per block / bias
    slot = BM_ELAMBDA;
    const int t = threadIdx.x;
    if (d.elecmethod == BM_PME && t == 0 && (L.fam & kFamSelf)) {
      // Ewald self term per block (nbrecip.cu:19-50): -beta kE/sqrt(pi)
      // lambda_b^2 sum_{i in b} q_i^2, in FP64 from the block sums
      const double pref = -d.betad * kElectricD / 1.7724538509055159;
      double es = 0.0;
      for (int b = 0; b < d.nl; b++) {
        const double l = b ? d.lambda[b] : 1.0;
        const double eb = pref * d.blockq2[b];
        if (b && eb != 0.0) atomic_fix(&d.dudl[b], fix_d(2.0 * l * eb));
        es += eb * l * l;
      }
      if (energy) atomic_fix(&d.energy[BM_ENBRECIPSELF], fix_d(es));
    }
    if (!(L.fam & kFamBias)) return;   // the self term alone
    if (d.ncr && t == 0) {
      // charge restraints on the lambda-weighted net charge Q
      // (msld/msld.cu:1400-1530); dE/dQ times q_b goes to every block but
      // the environment, the energy to the bias term
      double q = 0.0;
      for (int b = 0; b < d.nl; b++) q += (b ? d.lambda[b] : 1.0) * d.blockcharge[b];
      const double vol = d.box[0] * d.box[1] * d.box[2];
      double g = 0.0, ec = 0.0;
      if (d.elecmethod == BM_PME) {
        if (d.crk1 != 0.0) {
          // -(2 pi/3) kE k1 / V: the uniform-background shift of a net charge
          const double k = -2.0943951023931955 * kElectricD * d.crk1 / vol;
          g += k;
          ec += k * q;
        }
        if (d.crk2 != 0.0) {
          // -pi kE k2 / (V beta^2), with pi^2/beta^2 held in FP64
          const double k = -kElectricD * d.crk2 * d.pmeexpd / (3.141592653589793 * vol);
          g += k * q;
          ec += 0.5 * k * q * q;
        }
      }
      if (d.crk3 != 0.0) {
        double dq = q - d.crq0;
        if (d.crw > 0.0) dq = fabs(dq) <= d.crw ? 0.0 : dq - copysign(d.crw, dq);
        g += d.crk3 * dq;
        ec += 0.5 * d.crk3 * dq * dq;
      }
      for (int b = 1; b < d.nl; b++) {
        const double f = g * d.blockcharge[b];
        if (f != 0.0) atomic_fix(&d.dudl[b], fix_d(f));
      }
      if (energy) atomic_fix(&d.energy[BM_EBIAS], fix_d(ec));
    }
    if (d.nl > 1)
      for (int b = t; b < d.nl; b += blockDim.x) {
        double bias = d.fixed_bias[b];
        if (bias != 0.0) atomic_fix(&d.dudl[b], fix_d(bias));
        e += bias * d.lambda[b];
      }
    for (int v = t; v < d.nvb; v += blockDim.x) {
      int4 vi = d.vb_i[v];
      double2 vp = d.vb_p[v];
      const int i = vi.x, j = vi.y, ty = vi.z, nn = vi.w;
      const double li = d.lambda[i], lj = d.lambda[j], k = vp.y, l0 = vp.x;
      double ev = 0, fi = 0, fj = 0;
      switch (ty) {
        case 6: ev = k * li * lj; fi = k * lj; fj = k * li; break;
        case 8: ev = k * li * lj / (li + l0); fi = k * l0 * lj / ((li + l0) * (li + l0)); fj = k * li / (li + l0); break;
        case 10: ev = k * lj * (1 - exp(l0 * li)); fi = k * lj * (-l0 * exp(l0 * li)); fj = k * (1 - exp(l0 * li)); break;
        case 1: if (li < l0) { ev = k * pow(li - l0, nn); fi = nn * k * pow(li - l0, nn - 1); } break;
        case 2: if (li > l0) { ev = k * pow(li - l0, nn); fi = nn * k * pow(li - l0, nn - 1); } break;
        case 3: ev = k * pow(li - lj, nn); fi = nn * k * pow(li - lj, nn - 1); fj = -fi; break;
        case 4: ev = (li > l0) ? -k : -k * (1 - (li - l0) * (li - l0) / (l0 * l0)); fi = (li > l0) ? 0 : 2 * k * (li - l0) / (l0 * l0); break;
        case 5: ev = -k * li; fi = -k; break;
        case 7: ev = k * li * (1 - li) / (li + l0); fi = k * (l0 * (1 + l0) / ((li + l0) * (li + l0)) - 1); break;
        case 9: ev = k * lj * (1 - pow((li + l0) / l0, nn)); fi = -k * lj * pow((li + l0) / l0, nn - 1) / l0; fj = k * (1 - pow((li + l0) / l0, nn)); break;
        case 11: ev = k * pow(li, l0) * pow(lj, nn); fi = k * l0 * pow(li, l0 - 1) * pow(lj, nn); fj = k * nn * pow(li, l0) * pow(lj, nn - 1); break;
        case 12: ev = k * li * pow(lj, nn) / (li + l0); fi = k * l0 * pow(lj, nn) / ((li + l0) * (li + l0)); fj = k * nn * li * pow(lj, nn - 1) / (li + l0); break;
        default: break;
      }
      if (fi != 0.0) atomic_fix(&d.dudl[i], fix_d(fi));
      if (fj != 0.0) atomic_fix(&d.dudl[j], fix_d(fj));
      e += ev;
    }
    if (d.nl > 1)
      for (int st = t; st < d.nsites; st += blockDim.x) {
        const double et = theta_bias_site(d, st);
        if (energy && et != 0.0) atomic_fix(&d.energy[BM_ETHETA], fix_d(et));
      }
  }
  if (energy) wave_atomic_fix(&d.energy[slot], fix_d(e));
}

__global__ void __launch_bounds__(256) bonded_kernel(Dev d, BondedLaunch L, bool energy) {
  bonded_body(d, L, blockIdx.x, energy);
}

__global__ void __launch_bounds__(256) bonded_multi(Multi m, bool energy) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  bonded_body(d, bonded_layout(d), b.bid, energy);
}

int grid_bonded(const Dev &d) { return bonded_layout(d).b_bias + 1; }

void launch_bonded(const Dev &d, bool energy, unsigned fam, hipStream_t st) {
  bonded_kernel<<<grid_bonded(d), 256, 0, st>>>(d, bonded_layout(d, fam), energy);
}

void launch_bonded_multi(const Multi &m, bool energy, hipStream_t st) { bonded_multi<<<m.nwg, 256, 0, st>>>(m, energy); }

}  // namespace blade
