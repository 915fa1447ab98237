// oracle.cpp - FP64 host restatement of BLaDE's per-step path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the MI355X kernels and
// the OpenMP CPU baseline timed by bench.py.  The product library never
// links or calls this file.  Built with -ffp-contract=off so that the
// integer list semantics (orc_domdec) round exactly like the GPU path.
//
// Units AKMA (reference main/defines.h:42-61 CHARMM_UNITS branch).

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

const double kElectric = 332.0716;          // main/defines.h:55
const double kBoltz = 0.001987191;          // main/defines.h:53
const double softCoreRadius = 4.0;          // main/defines.h:48
const double twoOverSqrtPi = 1.1283791670955126;

inline double wrapd(double d, double L) { return d - L * std::floor(d / L + 0.5); }

struct Vec3 { double x, y, z; };
inline Vec3 v3(double a, double b, double c) { return Vec3{a, b, c}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator+(Vec3 a, Vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator*(double s, Vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double norm(Vec3 a) { return std::sqrt(dot(a, a)); }

inline Vec3 getp(const double *pos, int n, int i) { return v3(pos[i], pos[n + i], pos[2 * n + i]); }
inline Vec3 minimg(Vec3 d, const double *L) { return v3(wrapd(d.x, L[0]), wrapd(d.y, L[1]), wrapd(d.z, L[2])); }
inline void addf(double *f, int n, int i, Vec3 v) { f[i] += v.x; f[n + i] += v.y; f[2 * n + i] += v.z; }

inline int blockOf(int sb) { return sb & 0xFFFF; }
inline double lam(const double *lambda, int sb) { return (sb & 0xFFFF) ? lambda[sb & 0xFFFF] : 1.0; }

// ---------------------------------------------------------------------
// Pair functions
// ---------------------------------------------------------------------

struct PairOut { double f, e; };

// Electrostatics + vdW of one pair at effective distance r
// (nbdirect/nbdirect.cu:277-398).  f = dE/dr.
PairOut nb_pair(const bm_runparams *rp, double qq, double c12, double c6, double r) {
  const double rc = rp->rcut, rs = rp->rswitch;
  double rinv = 1.0 / r;
  double f = 0, e = 0;
  if (rp->elec_method == BM_FSWITCH) {
    double roff2 = rc * rc, ron2 = rs * rs;
    double ginv = 1 / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
    double A = roff2 * roff2 * (roff2 - 3 * ron2) * ginv, B = 6 * roff2 * ron2 * ginv;
    double C = -(ron2 + roff2) * ginv, D = 2 * ginv / 5;
    double dvc = 8 * (ron2 * roff2 * (rc - rs) - (roff2 * roff2 * rc - ron2 * ron2 * rs) / 5) * ginv;
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    if (r <= rs) {
      f = -kElectric * qq * rinv * rinv;
      e = kElectric * qq * (rinv + dvc);
    } else {
      f = -kElectric * qq * rinv * (A * rinv + B * r + 3 * C * r3 + 5 * D * r5);
      e = kElectric * qq * (A * (rinv - 1 / rc) + B * (rc - r) + C * (roff2 * rc - r3) + D * (roff2 * roff2 * rc - r5));
    }
  } else if (rp->elec_method == BM_PME) {
    double br = rp->beta_ewald * r;
    double erfcrinv = std::erfc(br) * rinv;
    f = -kElectric * qq * (erfcrinv + twoOverSqrtPi * rp->beta_ewald * std::exp(-br * br)) * rinv;
    e = kElectric * qq * erfcrinv;
  } else {
    double roffinv = 1 / rc, roffinv2 = roffinv * roffinv, kqq = kElectric * qq;
    f = -kqq * (rinv * rinv - roffinv2);
    e = kqq * (rinv - 2 * roffinv + r * roffinv2);
  }
  double rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  if (rp->vdw_method == BM_VFSWITCH) {
    double rc3 = rc * rc * rc, rs3 = rs * rs * rs;
    if (r < rs) {
      double dv6 = 1 / (rc3 * rs3);
      f += (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv;
      e += c12 * (rinv6 * rinv6 - dv6 * dv6) - c6 * (rinv6 - dv6);
    } else {
      double k6 = rc3 / (rc3 - rs3), k12 = rc3 * rc3 / (rc3 * rc3 - rs3 * rs3), rcinv3 = 1 / rc3;
      f += (6 * c6 * k6 * (rinv3 - rcinv3) * rinv3 - 12 * c12 * k12 * (rinv6 - rcinv3 * rcinv3) * rinv6) * rinv;
      e += c12 * k12 * (rinv6 - rcinv3 * rcinv3) * (rinv6 - rcinv3 * rcinv3) - c6 * k6 * (rinv3 - rcinv3) * (rinv3 - rcinv3);
    }
  } else if (rp->vdw_method == BM_VSWITCH) {
    if (r < rs) {
      f += (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv;
      e += c12 * rinv6 * rinv6 - c6 * rinv6;
    } else {
      double c2of = rc * rc, c2on = rs * rs;
      double rul3 = (c2of - c2on) * (c2of - c2on) * (c2of - c2on), rul12 = 12 / rul3;
      double rijl = c2on - r * r, riju = c2of - r * r;
      double fsw = riju * riju * (riju - 3 * rijl) / rul3, dfsw = rijl * riju * rul12;
      f += fsw * (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv + dfsw * (c12 * rinv6 - c6) * rinv6;
      e += fsw * (c12 * rinv6 - c6) * rinv6;
    }
  } else {
    double r2 = r * r, r5 = r2 * r2 * r, r6 = r2 * r2 * r2;
    double rinv7 = rinv6 * rinv, rinv12 = rinv6 * rinv6, rinv13 = rinv12 * rinv;
    double rc2 = rc * rc, rc6 = rc2 * rc2 * rc2, rc12 = rc6 * rc6, rc18 = rc12 * rc6;
    f += 6 * c6 * (rinv7 - r5 / rc12) - 12 * c12 * (rinv13 + r5 / rc18);
    e += c12 * (rinv12 + 2 * r6 / rc18 - 3 / rc12) - c6 * (rinv6 + r6 / rc12 - 2 / rc6);
  }
  return PairOut{f, e};
}

// 1-4 pair (bonded/pair.cu:56-175).  Note the VFSWITCH energy below the
// switch distance is shifted by rc^-6 / rc^-12 (pair.cu:92-103), which
// differs from the nbdirect shift; the restatement keeps it.
PairOut nb14_pair(const bm_runparams *rp, double qq, double c12, double c6, double e14fac, double r) {
  const double rc = rp->rcut, rs = rp->rswitch;
  if (!(r < rc)) return PairOut{0, 0};
  double rinv = 1 / r, rinv3 = rinv * rinv * rinv, rinv6 = rinv3 * rinv3;
  double f = 0, e = 0;
  if (rp->vdw_method == BM_VFSWITCH) {
    double rc3 = rc * rc * rc, rc6 = rc3 * rc3, rs3 = rs * rs * rs, rs6 = rs3 * rs3;
    double k6 = rc3 / (rc3 - rs3), k12 = rc6 / (rc6 - rs6);
    bool in = r < rs;
    double A6 = in ? 1 : k6, A12 = in ? 1 : k12, B6 = in ? 0 : 1 / rc3, B12 = in ? 0 : 1 / rc6;
    f = 6 * c6 * A6 * (rinv3 - B6) * rinv3 * rinv - 12 * c12 * A12 * (rinv6 - B12) * rinv6 * rinv;
    double CC6 = in ? -1 / rc6 : 0, CC12 = in ? -1 / (rc6 * rc6) : 0;
    e = c12 * (A12 * (rinv6 - B12) * (rinv6 - B12) + CC12) - c6 * (A6 * (rinv3 - B6) * (rinv3 - B6) + CC6);
  } else if (rp->vdw_method == BM_VSWITCH) {
    double c2of = rc * rc, c2on = rs * rs;
    double rul3 = (c2of - c2on) * (c2of - c2on) * (c2of - c2on), rul12 = 12 / rul3;
    double rijl = c2on - r * r, riju = c2of - r * r;
    double fsw = (r < rs) ? 1 : riju * riju * (riju - 3 * rijl) / rul3;
    double dfsw = (r < rs) ? 0 : rijl * riju * rul12;
    f = fsw * (6 * c6 - 12 * c12 * rinv6) * rinv6 * rinv + dfsw * (c12 * rinv6 - c6) * rinv6;
    e = fsw * (c12 * rinv6 - c6) * rinv6;
  } else {
    double r2 = r * r, r5 = r2 * r2 * r, r6 = r2 * r2 * r2;
    double rinv7 = rinv6 * rinv, rinv12 = rinv6 * rinv6, rinv13 = rinv12 * rinv;
    double rc2 = rc * rc, rc6 = rc2 * rc2 * rc2, rc12 = rc6 * rc6, rc18 = rc12 * rc6;
    f = 6 * c6 * (rinv7 - r5 / rc12) - 12 * c12 * (rinv13 + r5 / rc18);
    e = c12 * (rinv12 + 2 * r6 / rc18 - 3 / rc12) - c6 * (rinv6 + r6 / rc12 - 2 / rc6);
  }
  double kqq = kElectric * qq;
  if (rp->elec_method == BM_FSWITCH) {
    double kqqe14 = kqq * e14fac;
    double roff2 = rc * rc, ron2 = rs * rs;
    double ginv = 1 / ((roff2 - ron2) * (roff2 - ron2) * (roff2 - ron2));
    double A = roff2 * roff2 * (roff2 - 3 * ron2) * ginv, B = 6 * roff2 * ron2 * ginv;
    double C = -(ron2 + roff2) * ginv, D = 2 * ginv / 5;
    double dvc = 8 * (ron2 * roff2 * (rc - rs) - (roff2 * roff2 * rc - ron2 * ron2 * rs) / 5) * ginv;
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    if (r <= rs) {
      f += -kqqe14 * rinv * rinv;
      e += kqqe14 * (rinv + dvc);
    } else {
      f += -kqqe14 * rinv * (A * rinv + B * r + 3 * C * r3 + 5 * D * r5);
      e += kqqe14 * (A * (rinv - 1 / rc) + B * (rc - r) + C * (roff2 * rc - r3) + D * (roff2 * roff2 * rc - r5));
    }
  } else if (rp->elec_method == BM_PME) {
    double br = rp->beta_ewald * r;
    double erfcrinv = (std::erfc(br) + e14fac - 1) * rinv;
    f += kqq * (-erfcrinv - twoOverSqrtPi * rp->beta_ewald * std::exp(-br * br)) * rinv;
    e += kqq * erfcrinv;
  } else {
    double roffinv = 1 / rc, roffinv2 = roffinv * roffinv, kqqe14 = kqq * e14fac;
    f += -kqqe14 * (rinv * rinv - roffinv2);
    e += kqqe14 * (rinv - 2 * roffinv + r * roffinv2);
  }
  return PairOut{f, e};
}

// Soft-core effective distance (nbdirect/nbdirect.cu:257-275).
struct Soft { double reff, dredr, dredll; };
inline Soft soft_core(double r, double ll) {
  Soft s{r, 1, 0};
  double rSoft = softCoreRadius * (1 - ll);
  if (r < rSoft) {
    double rdivrs = r / rSoft;
    double re = 1 - 0.5 * rdivrs;
    re = re * rdivrs * rdivrs * rdivrs + 0.5;
    double dredr = (3 - 2 * rdivrs) * rdivrs * rdivrs;
    double dredll = (re - dredr * rdivrs) * (-softCoreRadius);
    s.reff = re * rSoft;
    s.dredr = dredr;
    s.dredll = dredll;
  }
  return s;
}

// Sorted per-atom exclusion lists (CSR).
struct Excl {
  std::vector<int> start, list;
  Excl(const bm_topology *t) {
    int n = t->n_atoms;
    std::vector<std::vector<int>> tmp(n);
    for (int e = 0; e < t->n_excl; e++) {
      int a = t->excl_idx[2 * e], b = t->excl_idx[2 * e + 1];
      tmp[a].push_back(b);
      tmp[b].push_back(a);
    }
    start.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
      std::sort(tmp[i].begin(), tmp[i].end());
      start[i + 1] = start[i] + (int)tmp[i].size();
    }
    list.reserve(start[n]);
    for (int i = 0; i < n; i++) list.insert(list.end(), tmp[i].begin(), tmp[i].end());
  }
  bool has(int i, int j) const {
    return std::binary_search(list.begin() + start[i], list.begin() + start[i + 1], j);
  }
};

// ---------------------------------------------------------------------
// nbdirect: every non-excluded pair with r<rcut, full shell per atom
// (nbdirect/nbdirect.cu:87-490 pair semantics).
// ---------------------------------------------------------------------
long nbdirect(const bm_topology *t, const bm_runparams *rp, const double *pos,
              const double *lambda, double *force, double *dudl, double *energy) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  const bool alch = t->n_blocks > 1;
  const bool useSoft = alch && rp->soft_core;
  Excl ex(t);
  // cell list with cells >= rcut
  int nc[3];
  for (int d = 0; d < 3; d++) nc[d] = std::max(1, (int)std::floor(L[d] / rp->rcut));
  bool useCells = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
  if (!useCells) nc[0] = nc[1] = nc[2] = 1;
  int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> cellOf(n), cellStart(ncell + 1, 0), cellAtoms(n);
  for (int i = 0; i < n; i++) {
    int c[3];
    for (int d = 0; d < 3; d++) {
      double x = pos[d * n + i];
      double f = x / L[d] - std::floor(x / L[d]);
      c[d] = std::min(nc[d] - 1, (int)(f * nc[d]));
    }
    cellOf[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
    cellStart[cellOf[i] + 1]++;
  }
  for (int c = 0; c < ncell; c++) cellStart[c + 1] += cellStart[c];
  {
    std::vector<int> fill(cellStart.begin(), cellStart.end() - 1);
    for (int i = 0; i < n; i++) cellAtoms[fill[cellOf[i]]++] = i;
  }
  int nthr = omp_get_max_threads();
  int nl = t->n_blocks;
  std::vector<double> dudlT((size_t)nthr * nl, 0.0);
  double eTot = 0;
  long pairs = 0;
#pragma omp parallel reduction(+ : eTot, pairs)
  {
    double *mydudl = &dudlT[(size_t)omp_get_thread_num() * nl];
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < n; i++) {
      Vec3 xi = getp(pos, n, i);
      int ci = cellOf[i];
      int cx = ci / (nc[1] * nc[2]), cy = (ci / nc[2]) % nc[1], cz = ci % nc[2];
      Vec3 fi = v3(0, 0, 0);
      int bi = t->site_block[i];
      double li = lam(lambda, bi);
      double qi = t->charge[i];
      int ti = t->type_idx[i];
      int span = useCells ? 1 : 0;
      for (int ox = -span; ox <= span; ox++)
        for (int oy = -span; oy <= span; oy++)
          for (int oz = -span; oz <= span; oz++) {
            int c = (((cx + ox + nc[0]) % nc[0]) * nc[1] + (cy + oy + nc[1]) % nc[1]) * nc[2] + (cz + oz + nc[2]) % nc[2];
            for (int q = cellStart[c]; q < cellStart[c + 1]; q++) {
              int j = cellAtoms[q];
              if (j == i || ex.has(i, j)) continue;
              Vec3 dr = minimg(xi - getp(pos, n, j), L);
              double r = norm(dr);
              if (!(r < rp->rcut)) continue;
              if (i < j) pairs++;
              int bj = t->site_block[j];
              double lj = lam(lambda, bj);
              double s = 1;
              bool sameSite = (bi & 0xFFFF0000) == (bj & 0xFFFF0000);
              if (alch) s = sameSite ? (bi == bj ? li : 0.0) : li * lj;
              Soft sc{r, 1, 0};
              if (useSoft && (bi || bj)) sc = soft_core(r, s);
              int tj = t->type_idx[j];
              PairOut p = nb_pair(rp, qi * t->charge[j], t->vdw_c12[ti * t->n_types + tj],
                                  t->vdw_c6[ti * t->n_types + tj], sc.reff);
              double fij = p.f * s;
              if (alch && (bi || bj)) {
                double fl = p.e + fij * sc.dredll;
                if (sameSite) {
                  if (bi == bj && blockOf(bi)) mydudl[blockOf(bi)] += 0.5 * fl;
                } else if (blockOf(bi)) {
                  mydudl[blockOf(bi)] += lj * fl;
                }
              }
              fij *= sc.dredr;
              fi = fi + (fij / r) * dr;
              eTot += 0.5 * s * p.e;
            }
          }
      force[i] += fi.x;
      force[n + i] += fi.y;
      force[2 * n + i] += fi.z;
    }
  }
  for (int th = 0; th < nthr; th++)
    for (int b = 0; b < nl; b++) dudl[b] += dudlT[(size_t)th * nl + b];
  energy[BM_ENBDIRECT] += eTot;
  return pairs;
}

// Ewald self term (nbrecip/nbrecip.cu:19-77).
void ewaldself(const bm_topology *t, const bm_runparams *rp, const double *lambda,
               double *dudl, double *energy) {
  double pref = -rp->beta_ewald * kElectric / std::sqrt(M_PI);
  double e = 0;
  for (int i = 0; i < t->n_atoms; i++) {
    int b = blockOf(t->site_block[i]);
    double l = b ? lambda[b] : 1.0;
    double q = t->charge[i];
    double ei = pref * q * q;
    if (b) dudl[b] += 2 * l * ei;
    e += ei * l * l;
  }
  energy[BM_ENBRECIPSELF] += e;
}

// ---------------------------------------------------------------------
// FFT (mixed radix, recursive, FP64)
// ---------------------------------------------------------------------
typedef std::complex<double> cplx;

void fft1d(cplx *a, int n, int stride, int sign, cplx *tmp) {
  if (n == 1) return;
  int p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n;
  int m = n / p;
  // gather p sub-sequences
  for (int r = 0; r < p; r++)
    for (int k = 0; k < m; k++) tmp[r * m + k] = a[(k * p + r) * stride];
  for (int r = 0; r < p; r++) {
    std::vector<cplx> sub(tmp + r * m, tmp + r * m + m), scratch(m);
    fft1d(sub.data(), m, 1, sign, scratch.data());
    std::copy(sub.begin(), sub.end(), tmp + r * m);
  }
  for (int k = 0; k < n; k++) {
    cplx s = 0;
    for (int r = 0; r < p; r++) {
      double ang = sign * 2 * M_PI * (double)((long)r * k % n) / n;
      s += tmp[r * m + (k % m)] * cplx(std::cos(ang), std::sin(ang));
    }
    a[k * stride] = s;
  }
}

void fft3d(int K0, int K1, int K2, cplx *g, int sign) {
  // z lines
#pragma omp parallel
  {
    std::vector<cplx> line(std::max(K0, std::max(K1, K2))), tmp(line.size());
#pragma omp for
    for (int ij = 0; ij < K0 * K1; ij++) {
      for (int k = 0; k < K2; k++) line[k] = g[(size_t)ij * K2 + k];
      fft1d(line.data(), K2, 1, sign, tmp.data());
      for (int k = 0; k < K2; k++) g[(size_t)ij * K2 + k] = line[k];
    }
#pragma omp for
    for (int ik = 0; ik < K0 * K2; ik++) {
      int i = ik / K2, k = ik % K2;
      for (int j = 0; j < K1; j++) line[j] = g[((size_t)i * K1 + j) * K2 + k];
      fft1d(line.data(), K1, 1, sign, tmp.data());
      for (int j = 0; j < K1; j++) g[((size_t)i * K1 + j) * K2 + k] = line[j];
    }
#pragma omp for
    for (int jk = 0; jk < K1 * K2; jk++) {
      int j = jk / K2, k = jk % K2;
      for (int i = 0; i < K0; i++) line[i] = g[((size_t)i * K1 + j) * K2 + k];
      fft1d(line.data(), K0, 1, sign, tmp.data());
      for (int i = 0; i < K0; i++) g[((size_t)i * K1 + j) * K2 + k] = line[i];
    }
  }
}

// Cardinal B-spline M_n(u), support [0,n] (recursion M_n(u) = (u M_{n-1}(u)
// + (n-u) M_{n-1}(u-1))/(n-1)).
double bspline(int n, double u) {
  if (n == 1) return (u >= 0 && u < 1) ? 1.0 : 0.0;  // M_1 on [0,1)
  if (u <= 0 || u >= n) return 0;
  return (u * bspline(n - 1, u) + (n - u) * bspline(n - 1, u - 1)) / (n - 1);
}

// 1/|b(m)|^2 with b(m) built from M_n at integer points
// (system/potential.cxx:1245-1287).
void bsp_moduli(int n, int K, double *out) {
  for (int i = 0; i < K; i++) {
    double re = 0, im = 0;
    for (int l = 1; l < n; l++) {
      double m = bspline(n, l);
      re += m * std::cos(2 * M_PI * i * l / K);
      im += m * std::sin(2 * M_PI * i * l / K);
    }
    out[i] = 1.0 / (re * re + im * im);
  }
}

// Weights and derivatives of the n spline points of one coordinate:
// point (u0+k) mod K gets w[k] = M_n(frac+n-1-k) (nbrecip/nbrecip.cu:108-141).
void spline_weights(int n, double frac, double *w, double *dw) {
  for (int k = 0; k < n; k++) {
    double u = frac + n - 1 - k;
    w[k] = bspline(n, u);
    dw[k] = bspline(n - 1, u) - bspline(n - 1, u - 1);
  }
}

// Smooth PME (nbrecip/nbrecip.cu:83-485): lambda-scaled spread, FFT,
// convolution with B(m) (k_e/2 pi V) exp(-pi^2 m^2/beta^2)/m^2, inverse FFT,
// gather of force 2 l q grad(phi), dU/dl 2 q phi, energy l q phi.
void pme(const bm_topology *t, const bm_runparams *rp, const double *pos,
         const double *lambda, double *force, double *dudl, double *energy) {
  const int n = t->n_atoms, ord = rp->order;
  const int K0 = rp->grid[0], K1 = rp->grid[1], K2 = rp->grid[2];
  const double *L = rp->box;
  std::vector<cplx> grid((size_t)K0 * K1 * K2, 0.0);
  std::vector<int> u0(3 * n);
  std::vector<double> W(3 * n * ord), DW(3 * n * ord);
  const int K[3] = {K0, K1, K2};
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      double u = pos[d * n + i] * K[d] / L[d];
      double fl = std::floor(u);
      int iu = (int)fl;
      u0[3 * i + d] = ((iu % K[d]) + K[d]) % K[d];
      spline_weights(ord, u - fl, &W[(3 * i + d) * ord], &DW[(3 * i + d) * ord]);
    }
  }
  for (int i = 0; i < n; i++) {
    int b = blockOf(t->site_block[i]);
    double lq = (b ? lambda[b] : 1.0) * t->charge[i];
    for (int a = 0; a < ord; a++) {
      int ia = (u0[3 * i] + a) % K0;
      for (int c = 0; c < ord; c++) {
        int ic = (u0[3 * i + 1] + c) % K1;
        double wac = lq * W[(3 * i) * ord + a] * W[(3 * i + 1) * ord + c];
        for (int e = 0; e < ord; e++) {
          int ie = (u0[3 * i + 2] + e) % K2;
          grid[((size_t)ia * K1 + ic) * K2 + ie] += wac * W[(3 * i + 2) * ord + e];
        }
      }
    }
  }
  fft3d(K0, K1, K2, grid.data(), -1);
  std::vector<double> bx(K0), by(K1), bz(K2);
  bsp_moduli(ord, K0, bx.data());
  bsp_moduli(ord, K1, by.data());
  bsp_moduli(ord, K2, bz.data());
  double V = L[0] * L[1] * L[2];
  for (int i = 0; i < K0; i++) {
    int mi = (2 * i <= K0) ? i : i - K0;
    for (int j = 0; j < K1; j++) {
      int mj = (2 * j <= K1) ? j : j - K1;
      for (int k = 0; k < K2; k++) {
        int mk = (2 * k <= K2) ? k : k - K2;
        double m2 = (mi / L[0]) * (mi / L[0]) + (mj / L[1]) * (mj / L[1]) + (mk / L[2]) * (mk / L[2]);
        size_t idx = ((size_t)i * K1 + j) * K2 + k;
        if (idx == 0) { grid[0] = 0; continue; }
        double fac = bx[i] * by[j] * bz[k] * (0.5 * kElectric / M_PI) / V *
                     std::exp(-M_PI * M_PI * m2 / (rp->beta_ewald * rp->beta_ewald)) / m2;
        grid[idx] *= fac;
      }
    }
  }
  fft3d(K0, K1, K2, grid.data(), +1);
  double eTot = 0;
  for (int i = 0; i < n; i++) {
    int b = blockOf(t->site_block[i]);
    double l = b ? lambda[b] : 1.0, q = t->charge[i];
    double phi = 0, gx = 0, gy = 0, gz = 0;
    for (int a = 0; a < ord; a++) {
      int ia = (u0[3 * i] + a) % K0;
      double wa = W[(3 * i) * ord + a], da = DW[(3 * i) * ord + a];
      for (int c = 0; c < ord; c++) {
        int ic = (u0[3 * i + 1] + c) % K1;
        double wc = W[(3 * i + 1) * ord + c], dc = DW[(3 * i + 1) * ord + c];
        for (int e = 0; e < ord; e++) {
          int ie = (u0[3 * i + 2] + e) % K2;
          double we = W[(3 * i + 2) * ord + e], de = DW[(3 * i + 2) * ord + e];
          double P = grid[((size_t)ia * K1 + ic) * K2 + ie].real();
          phi += P * wa * wc * we;
          gx += P * da * wc * we;
          gy += P * wa * dc * we;
          gz += P * wa * wc * de;
        }
      }
    }
    force[i] += 2 * l * q * gx * K0 / L[0];
    force[n + i] += 2 * l * q * gy * K1 / L[1];
    force[2 * n + i] += 2 * l * q * gz * K2 / L[2];
    if (b) dudl[b] += 2 * q * phi;
    eTot += l * q * phi;
  }
  energy[BM_ENBRECIP] += eTot;
}

// ---------------------------------------------------------------------
// Bonded terms (bonded/bonded.cu:44-527), non-soft variants.
// ---------------------------------------------------------------------
struct Scale2 { int b0, b1; double l0, l1; };
inline Scale2 scale2(const int *sb, const double *lambda) {
  Scale2 s{blockOf(sb[0]), blockOf(sb[1]), 1, 1};
  if (s.b0) {
    s.l0 = lambda[s.b0];
    if (s.b1) s.l1 = lambda[s.b1];
  }
  return s;
}
inline void scale_dudl(const Scale2 &s, double e, double *dudl) {
  if (s.b0) {
    dudl[s.b0] += s.l1 * e;
    if (s.b1) dudl[s.b1] += s.l0 * e;
  }
}

// lambda weight of a term's energy E0: E = w E0, dE/dP = dw E0 with
// P = l0 l1; a soft angle, dihedral or improper takes w = P^q
// (bonded/bonded.cu:239-256, 408-425)
struct Weight { double w, dw; };
inline Weight weight(const Scale2 &s, bool soft, double q) {
  const double P = s.l0 * s.l1;
  if (!soft) return {P, 1.0};
  return {std::pow(P, q), q * std::pow(P, q - 1)};
}
inline void weight_dudl(const Scale2 &s, double dEdP, double *dudl) { scale_dudl(s, dEdP, dudl); }

void bonded(const bm_topology *t, const bm_runparams *rp, const double *pos,
            const double *lambda, double *force, double *dudl, double *energy) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  for (int b = 0; b < t->n_bonds; b++) {
    int i = t->bond_idx[2 * b], j = t->bond_idx[2 * b + 1];
    Vec3 dr = minimg(getp(pos, n, i) - getp(pos, n, j), L);
    double r = norm(dr);
    Scale2 s = scale2(&t->bond_sb[2 * b], lambda);
    double db = r - t->bond_b0[b];
    double e0 = 0.5 * t->bond_kb[b] * db * db;
    double f, e;
    if (t->bond_soft && t->bond_soft[b]) {
      // soft bond (bonded/bonded.cu:81-107): with P = l0 l1 and Q = P^p,
      // E = Q E0 g, g = 1 / (1 + (1 - Q) alpha db^2)
      const double P = s.l0 * s.l1, p = rp->soft_bond_power, a = rp->soft_bond_alpha;
      const double Q = std::pow(P, p), g = 1.0 / (1.0 + (1.0 - Q) * a * db * db);
      const double dgdr = -2.0 * g * g * (1.0 - Q) * a * db;
      const double dgdQ = g * g * a * db * db;
      e = Q * e0 * g;
      f = Q * (t->bond_kb[b] * db * g + e0 * dgdr);
      weight_dudl(s, e0 * p * std::pow(P, p - 1) * (g + Q * dgdQ), dudl);
    } else {
      e = e0 * s.l0 * s.l1;
      f = t->bond_kb[b] * db * s.l0 * s.l1;
      scale_dudl(s, e0, dudl);
    }
    addf(force, n, i, (f / r) * dr);
    addf(force, n, j, (-f / r) * dr);
    energy[b < t->n_bonds12 ? BM_EBOND : BM_EUREY] += e;
  }
  for (int a = 0; a < t->n_angles; a++) {
    int i = t->angle_idx[3 * a], j = t->angle_idx[3 * a + 1], k = t->angle_idx[3 * a + 2];
    Vec3 rij = minimg(getp(pos, n, i) - getp(pos, n, j), L);
    Vec3 rkj = minimg(getp(pos, n, k) - getp(pos, n, j), L);
    double dp = dot(rij, rkj);
    Vec3 cr = cross(rij, rkj);
    double mc = norm(cr);
    double th = std::atan2(mc, dp);
    Scale2 s = scale2(&t->angle_sb[2 * a], lambda);
    const Weight wt = weight(s, t->angle_soft && t->angle_soft[a], rp->soft_bond_exp);
    double dth = th - t->angle_t0[a];
    double e = 0.5 * t->angle_k[a] * dth * dth;
    double f = t->angle_k[a] * dth * wt.w;
    weight_dudl(s, wt.dw * e, dudl);
    Vec3 fi = (f / (mc * dot(rij, rij))) * cross(rij, cr);
    Vec3 fk = (-f / (mc * dot(rkj, rkj))) * cross(rkj, cr);
    addf(force, n, i, fi);
    addf(force, n, k, fk);
    addf(force, n, j, -1.0 * (fi + fk));
    energy[BM_EANGLE] += e * wt.w;
  }
  for (int pass = 0; pass < 2; pass++) {
    int cnt = pass ? t->n_imprs : t->n_dihes;
    const int *idx = pass ? t->impr_idx : t->dihe_idx;
    const int *sbv = pass ? t->impr_sb : t->dihe_sb;
    for (int d = 0; d < cnt; d++) {
      int i = idx[4 * d], j = idx[4 * d + 1], k = idx[4 * d + 2], l = idx[4 * d + 3];
      Vec3 rij = minimg(getp(pos, n, i) - getp(pos, n, j), L);
      Vec3 rjk = minimg(getp(pos, n, j) - getp(pos, n, k), L);
      Vec3 rkl = minimg(getp(pos, n, k) - getp(pos, n, l), L);
      Vec3 m = cross(rij, rjk), nv = cross(rjk, rkl);
      double sinp = norm(cross(m, nv)), cosp = dot(m, nv);
      double phi = std::atan2(sinp, cosp);
      if (dot(rij, nv) > 0) phi = -phi;   // bonded.cu:372-374
      double fphi, e;
      if (!pass) {
        double dphi = t->dihe_n[d] * phi - t->dihe_d0[d];
        fphi = -t->dihe_k[d] * t->dihe_n[d] * std::sin(dphi);
        e = t->dihe_k[d] * (std::cos(dphi) + 1);
      } else if (t->impr_n[d] > 0) {
        double dphi = t->impr_n[d] * phi - t->impr_d0[d];
        fphi = -t->impr_k[d] * t->impr_n[d] * std::sin(dphi);
        e = t->impr_k[d] * (std::cos(dphi) + 1);
      } else {
        double dphi = phi - t->impr_d0[d];
        dphi -= 2 * M_PI * std::floor((dphi + M_PI) / (2 * M_PI));
        fphi = t->impr_k[d] * dphi;
        e = 0.5 * t->impr_k[d] * dphi * dphi;
      }
      Scale2 s = scale2(&sbv[2 * d], lambda);
      const int32_t *sf = pass ? t->impr_soft : t->dihe_soft;
      const Weight wt = weight(s, sf && sf[d], rp->soft_bond_exp);
      fphi *= wt.w;
      weight_dudl(s, wt.dw * e, dudl);
      double rjk2 = dot(rjk, rjk), r = std::sqrt(rjk2);
      Vec3 fi = (-fphi * r / dot(m, m)) * m;
      Vec3 fl = (-fphi * r / dot(nv, nv)) * nv;   // fk before negation
      double p = dot(rij, rjk) / rjk2, q = dot(rkl, rjk) / rjk2;
      Vec3 fj = (-p) * fi + (-q) * fl;
      Vec3 fkk = fl - fj;
      Vec3 fjj = fj - fi;
      addf(force, n, i, fi);
      addf(force, n, l, -1.0 * fl);
      addf(force, n, k, fkk);
      addf(force, n, j, fjj);
      energy[pass ? BM_EIMPR : BM_EDIHE] += e * wt.w;
    }
  }
}

// 1-4 pairs and PME exclusion corrections (bonded/pair.cu:197-387).
void nb14(const bm_topology *t, const bm_runparams *rp, const double *pos,
          const double *lambda, double *force, double *dudl, double *energy) {
  const int n = t->n_atoms;
  for (int p = 0; p < t->n_nb14; p++) {
    int i = t->nb14_idx[2 * p], j = t->nb14_idx[2 * p + 1];
    Vec3 dr = minimg(getp(pos, n, i) - getp(pos, n, j), rp->box);
    double r = norm(dr);
    Scale2 s = scale2(&t->nb14_sb[2 * p], lambda);
    Soft sc{r, 1, 0};
    if (rp->soft_core14 && s.b0) sc = soft_core(r, s.l0 * s.l1);
    PairOut o = nb14_pair(rp, t->nb14_qq[p], t->nb14_c12[p], t->nb14_c6[p], t->nb14_e14fac[p], sc.reff);
    double f = o.f * s.l0 * s.l1;
    scale_dudl(s, o.e + f * sc.dredll, dudl);
    f *= sc.dredr;
    addf(force, n, i, (f / r) * dr);
    addf(force, n, j, (-f / r) * dr);
    energy[BM_ENB14] += o.e * s.l0 * s.l1;
  }
}

// One exclusion correction E = -kqq erf(beta r)/r (bonded/pair.cu:181-193)
// and dE/dr / r.  For beta r < 0.01 both come from their Taylor series
// about r = 0 (E -> -kqq 2 beta/sqrt(pi), dE/dr / r finite): with msldEwaldType
// 3 the pairs between alternative substituents of one site are corrected,
// and such atoms may coincide (r = 0), where the closed form is 0/0.
static void nbex_pair(double kqq, double beta, double r, double &fr, double &e) {
  const double x = beta * r;
  if (x < 0.01) {
    const double x2 = x * x;
    e = -kqq * beta * twoOverSqrtPi * (1 - x2 / 3 + x2 * x2 / 10 - x2 * x2 * x2 / 42);
    fr = kqq * beta * beta * beta * twoOverSqrtPi * (2.0 / 3 - 2 * x2 / 5 + x2 * x2 / 7);
    return;
  }
  const double rinv = 1 / r;
  e = -kqq * std::erf(x) * rinv;
  fr = kqq * (std::erf(x) * rinv - twoOverSqrtPi * beta * std::exp(-x * x)) * rinv * rinv;
}

void nbex(const bm_topology *t, const bm_runparams *rp, const double *pos,
          const double *lambda, double *force, double *dudl, double *energy) {
  const int n = t->n_atoms;
  const double beta = rp->beta_ewald;
  for (int p = 0; p < t->n_nbex; p++) {
    int i = t->nbex_idx[2 * p], j = t->nbex_idx[2 * p + 1];
    Vec3 dr = minimg(getp(pos, n, i) - getp(pos, n, j), rp->box);
    double r = norm(dr);
    double kqq = kElectric * t->nbex_qq[p];
    double fr, e;
    nbex_pair(kqq, beta, r, fr, e);
    Scale2 s = scale2(&t->nbex_sb[2 * p], lambda);
    fr *= s.l0 * s.l1;
    scale_dudl(s, e, dudl);
    addf(force, n, i, fr * dr);
    addf(force, n, j, (-fr) * dr);
    energy[BM_ENBRECIPEXCL] += e * s.l0 * s.l1;
  }
}

// CMAP (bonded/bonded.cu:531-720, system/potential.cxx:460-520).  Grid
// slopes per grid unit are those of the periodic cubic spline, found here
// by Gauss-Seidel sweeps on dy[i-1] + 4 dy[i] + dy[i+1] = 3 (y[i+1] - y[i-1])
// (diagonally dominant, each sweep at least halves the error); the cell is
// evaluated as nested one-dimensional cubic Hermite interpolation, first
// along psi, then along phi.  Energy scaled by the product of up to three
// lambdas.
static void spline_slopes_gs(int n, const std::vector<double> &y, std::vector<double> &dy) {
  dy.assign(n, 0.0);
  for (int sweep = 0; sweep < 200; sweep++)
    for (int i = 0; i < n; i++)
      dy[i] = (3 * (y[(i + 1) % n] - y[(i + n - 1) % n]) - dy[(i + n - 1) % n] - dy[(i + 1) % n]) / 4;
}

// cubic Hermite on [0,1]: value and derivative from end values and slopes
static void hermite(double p0, double p1, double m0, double m1, double t, double &v, double &dv) {
  const double t2 = t * t, t3 = t2 * t;
  v = (2 * t3 - 3 * t2 + 1) * p0 + (t3 - 2 * t2 + t) * m0 + (-2 * t3 + 3 * t2) * p1 + (t3 - t2) * m1;
  dv = (6 * t2 - 6 * t) * p0 + (3 * t2 - 4 * t + 1) * m0 + (-6 * t2 + 6 * t) * p1 + (3 * t2 - 2 * t) * m1;
}

void cmap(const bm_topology *t, const bm_runparams *rp, const double *pos, const double *lambda, double *force,
          double *dudl, double *energy) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  // per grid: value, d/dphi, d/dpsi, d2/dphi dpsi at every point
  std::vector<std::vector<double>> F(t->n_cmap_types), FX(F.size()), FY(F.size()), FXY(F.size());
  for (int g = 0; g < t->n_cmap_types; g++) {
    const int ng = t->cmap_ngrid[g];
    const double *v = t->cmap_grid + t->cmap_off[g];
    F[g].assign(v, v + ng * ng);
    FX[g].assign(ng * ng, 0.0);
    FY[g].assign(ng * ng, 0.0);
    FXY[g].assign(ng * ng, 0.0);
    std::vector<double> y(ng), dy;
    for (int j = 0; j < ng; j++) {
      for (int i = 0; i < ng; i++) y[i] = v[ng * i + j];
      spline_slopes_gs(ng, y, dy);
      for (int i = 0; i < ng; i++) FX[g][ng * i + j] = dy[i];
    }
    for (int i = 0; i < ng; i++) {
      for (int j = 0; j < ng; j++) y[j] = v[ng * i + j];
      spline_slopes_gs(ng, y, dy);
      for (int j = 0; j < ng; j++) FY[g][ng * i + j] = dy[j];
      for (int j = 0; j < ng; j++) y[j] = FX[g][ng * i + j];
      spline_slopes_gs(ng, y, dy);
      for (int j = 0; j < ng; j++) FXY[g][ng * i + j] = dy[j];
    }
  }
  for (int c = 0; c < t->n_cmaps; c++) {
    const int *a = t->cmap_idx + 8 * c;
    const int g = t->cmap_type[c], ng = t->cmap_ngrid[g];
    Vec3 rij[2], rjk[2], rkl[2], m[2], nv[2];
    double ang[2];
    for (int h = 0; h < 2; h++) {
      const int *q = a + 4 * h;
      rij[h] = minimg(getp(pos, n, q[0]) - getp(pos, n, q[1]), L);
      rjk[h] = minimg(getp(pos, n, q[1]) - getp(pos, n, q[2]), L);
      rkl[h] = minimg(getp(pos, n, q[2]) - getp(pos, n, q[3]), L);
      m[h] = cross(rij[h], rjk[h]);
      nv[h] = cross(rjk[h], rkl[h]);
      ang[h] = std::atan2(norm(cross(m[h], nv[h])), dot(m[h], nv[h]));
      if (dot(rij[h], nv[h]) > 0) ang[h] = -ang[h];
    }
    // grid coordinates: point k at -180 + k 360/ng degrees
    const double step = 2 * M_PI / ng;
    int cell[2];
    double frac[2];
    for (int h = 0; h < 2; h++) {
      const double x = (ang[h] + M_PI) / step;
      const double fl = std::floor(x);
      frac[h] = x - fl;
      cell[h] = (((long)fl % ng) + ng) % ng;
    }
    // the reference bins with floor(angle/step) + ng/2: for odd ng the
    // cell origin shifts by half a step
    if (ng % 2) {
      for (int h = 0; h < 2; h++) {
        const double x = ang[h] / step;
        const double fl = std::floor(x);
        frac[h] = x - fl;
        cell[h] = ((((long)fl + ng / 2) % ng) + ng) % ng;
      }
    }
    const int i0 = cell[0], i1 = (i0 + 1) % ng, j0 = cell[1], j1 = (j0 + 1) % ng;
    auto at = [&](const std::vector<double> &A, int i, int j) { return A[ng * i + j]; };
    // along psi at phi rows i0 and i1: values and phi-slopes, with psi derivatives
    double v0, dv0, v1, dv1, s0, ds0, s1, ds1;
    hermite(at(F[g], i0, j0), at(F[g], i0, j1), at(FY[g], i0, j0), at(FY[g], i0, j1), frac[1], v0, dv0);
    hermite(at(F[g], i1, j0), at(F[g], i1, j1), at(FY[g], i1, j0), at(FY[g], i1, j1), frac[1], v1, dv1);
    hermite(at(FX[g], i0, j0), at(FX[g], i0, j1), at(FXY[g], i0, j0), at(FXY[g], i0, j1), frac[1], s0, ds0);
    hermite(at(FX[g], i1, j0), at(FX[g], i1, j1), at(FXY[g], i1, j0), at(FXY[g], i1, j1), frac[1], s1, ds1);
    double e, dedt, dpsi, unused;
    hermite(v0, v1, s0, s1, frac[0], e, dedt);
    hermite(dv0, dv1, ds0, ds1, frac[0], dpsi, unused);
    double l[3], ll = 1;
    int b[3];
    for (int k = 0; k < 3; k++) {
      b[k] = t->cmap_sb[3 * c + k] & 0xFFFF;
      l[k] = b[k] ? lambda[b[k]] : 1.0;
      ll *= l[k];
    }
    for (int k = 0; k < 3; k++)
      if (b[k]) dudl[b[k]] += e * ll / l[k];
    energy[BM_ECMAP] += e * ll;
    const double dE[2] = {dedt / step * ll, dpsi / step * ll};
    for (int h = 0; h < 2; h++) {
      const int *q = a + 4 * h;
      const double rjk2 = dot(rjk[h], rjk[h]), r = std::sqrt(rjk2);
      Vec3 fi = (-dE[h] * r / dot(m[h], m[h])) * m[h];
      Vec3 fl = (-dE[h] * r / dot(nv[h], nv[h])) * nv[h];
      const double p = dot(rij[h], rjk[h]) / rjk2, s = dot(rkl[h], rjk[h]) / rjk2;
      Vec3 fj = (-p) * fi + (-s) * fl;
      addf(force, n, q[0], fi);
      addf(force, n, q[3], -1.0 * fl);
      addf(force, n, q[2], fl - fj);
      addf(force, n, q[1], fj - fi);
    }
  }
}

// Root t of sum_i max(theta_i - t, 0)^3 = 1 over one site's blocks by
// Newton's method from t = max(theta) - 1 (msld.cu:785-861: at most 51
// iterations, stop when a step is <= 1e-12).  The sum decreases and is
// convex in t, so the iterates rise monotonically to the root.
double piecewise_root(const double *theta, int nb) {
  double x0 = -1e9;
  for (int i = 0; i < nb; i++) x0 = std::max(x0, theta[i]);
  x0 -= 1.0;
  for (int it = 0; it <= 50; it++) {
    double c = -1.0, dc = 0.0;
    for (int i = 0; i < nb; i++) {
      const double d = theta[i] - x0;
      if (d > 0) {
        c += d * d * d;
        dc += 3 * d * d;
      }
    }
    const double x1 = x0 + c / dc;
    if (std::fabs(x1 - x0) <= 1e-12) return x1;
    x0 = x1;
  }
  return x0;
}

// Fixed and variable lambda biases (msld/msld.cu:1000-1130).
void msldbias(const bm_topology *t, const double *lambda, double *dudl, double *energy) {
  for (int b = 0; b < t->n_blocks; b++) {
    double bias = t->fixed_bias ? t->fixed_bias[b] : 0.0;
    dudl[b] += bias;
    energy[BM_ELAMBDA] += bias * lambda[b];
  }
  for (int v = 0; v < t->n_varbias; v++) {
    int i = t->vb_i[v], j = t->vb_j[v], ty = t->vb_type[v], nn = t->vb_n[v];
    double li = lambda[i], lj = lambda[j], k = t->vb_k[v], l0 = t->vb_l0[v];
    double e = 0, fi = 0, fj = 0;
    switch (ty) {
      case 6: e = k * li * lj; fi = k * lj; fj = k * li; break;
      case 8: e = k * li * lj / (li + l0); fi = k * l0 * lj / ((li + l0) * (li + l0)); fj = k * li / (li + l0); break;
      case 10: e = k * lj * (1 - std::exp(l0 * li)); fi = k * lj * (-l0 * std::exp(l0 * li)); fj = k * (1 - std::exp(l0 * li)); break;
      case 1: if (li < l0) { e = k * std::pow(li - l0, nn); fi = nn * k * std::pow(li - l0, nn - 1); } break;
      case 2: if (li > l0) { e = k * std::pow(li - l0, nn); fi = nn * k * std::pow(li - l0, nn - 1); } break;
      case 3: e = k * std::pow(li - lj, nn); fi = nn * k * std::pow(li - lj, nn - 1); fj = -fi; break;
      case 4: e = (li > l0) ? -k : -k * (1 - (li - l0) * (li - l0) / (l0 * l0)); fi = (li > l0) ? 0 : 2 * k * (li - l0) / (l0 * l0); break;
      case 5: e = -k * li; fi = -k; break;
      case 7: e = k * li * (1 - li) / (li + l0); fi = k * (l0 * (1 + l0) / ((li + l0) * (li + l0)) - 1); break;
      case 9: e = k * lj * (1 - std::pow((li + l0) / l0, nn)); fi = -k * lj * std::pow((li + l0) / l0, nn - 1) / l0; fj = k * (1 - std::pow((li + l0) / l0, nn)); break;
      case 11: e = k * std::pow(li, l0) * std::pow(lj, nn); fi = k * l0 * std::pow(li, l0 - 1) * std::pow(lj, nn); fj = k * nn * std::pow(li, l0) * std::pow(lj, nn - 1); break;
      case 12: e = k * li * std::pow(lj, nn) / (li + l0); fi = k * l0 * std::pow(lj, nn) / ((li + l0) * (li + l0)); fj = k * nn * li * std::pow(lj, nn - 1) / (li + l0); break;
      default: break;
    }
    dudl[i] += fi;
    dudl[j] += fj;
    energy[BM_ELAMBDA] += e;
  }
}

// MSLD charge restraints (msld/msld.cu:1400-1530) on Q = sum_b lambda_b q_b,
// lambda_0 = 1: the first two are PME-only, E1 = k Q with
// k = -(2 pi/3) kE k1 / V and E2 = k Q^2 / 2 with k = -pi kE k2 / (V beta^2);
// the third is harmonic, k3/2 dQ^2, in the distance dQ of Q - q0 from the
// flat band [-w, w].  dU/dlambda_b = dE/dQ q_b for b > 0; energy BM_EBIAS.
// The block charges q_b are summed here from the atoms' charges and block
// codes (Msld::initialize, msld.cu:684-690), not taken from the topology's
// block_charge table.
void chargerestraints(const bm_topology *t, const bm_runparams *rp, const double *lambda, double *dudl,
                      double *energy) {
  std::vector<double> qb(t->n_blocks, 0.0);
  for (int i = 0; i < t->n_atoms; i++) qb[t->site_block[i] & 0xFFFF] += t->charge[i];
  double q = 0;
  for (int b = 0; b < t->n_blocks; b++) q += (b ? lambda[b] : 1.0) * qb[b];
  const double vol = rp->box[0] * rp->box[1] * rp->box[2];
  double dedq[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  if (rp->elec_method == BM_PME && t->cr_k1 != 0) {
    const double k = -(2 * M_PI / 3) * kElectric * t->cr_k1 / vol;
    e[0] = k * q;
    dedq[0] = k;
  }
  if (rp->elec_method == BM_PME && t->cr_k2 != 0) {
    const double k = -M_PI * kElectric * t->cr_k2 / (vol * rp->beta_ewald * rp->beta_ewald);
    e[1] = 0.5 * k * q * q;
    dedq[1] = k * q;
  }
  if (t->cr_k3 != 0) {
    double dq = q - t->cr_q0;
    if (t->cr_w > 0) {
      if (std::fabs(dq) <= t->cr_w) dq = 0;
      else dq -= std::copysign(t->cr_w, dq);
    }
    e[2] = 0.5 * t->cr_k3 * dq * dq;
    dedq[2] = t->cr_k3 * dq;
  }
  for (int v = 0; v < 3; v++) {
    for (int b = 1; b < t->n_blocks; b++) dudl[b] += dedq[v] * qb[b];
    energy[BM_EBIAS] += e[v];
  }
}

// MSLD atom restraints (msld/msld.cu:1328-1364): each group's atoms are
// pulled to the group centroid (minimum-image offsets from its first
// atom), E = k/2 sum |x_i - c|^2 and dU/dx_i = k (x_i - c); not scaled.
void cats(const bm_topology *t, const bm_runparams *rp, const double *pos, double *force, double *energy) {
  const int n = t->n_atoms;
  for (int g = 0; g < t->n_cats; g++) {
    const int m0 = t->cats_bound[g], m1 = t->cats_bound[g + 1];
    const Vec3 x0 = getp(pos, n, t->cats_idx[m0]);
    Vec3 c = v3(0, 0, 0);
    for (int m = m0 + 1; m < m1; m++) c = c + minimg(getp(pos, n, t->cats_idx[m]) - x0, rp->box);
    c = (1.0 / (m1 - m0)) * c;
    for (int m = m0; m < m1; m++) {
      const int a = t->cats_idx[m];
      const Vec3 dx = minimg(getp(pos, n, a) - x0 - c, rp->box);
      addf(force, n, a, t->cats_k * dx);
      energy[BM_ECATS] += 0.5 * t->cats_k * dot(dx, dx);
    }
  }
}

// Restraints (restrain/restrain.cu:15-598).  NOE pairs or a fixed point
// (BM_ENOE), harmonic positions k |x - x0|^n (BM_EHARMONIC), bond, angle and
// dihedral restraints scaled by lambda of their block (BM_EMMFP) and RESD
// two-distance restraints (BM_ERESD); dU/dx into force, the unscaled energy
// of a scaled term into dudl.
void restraints(const bm_topology *t, const bm_runparams *rp, const double *pos, const double *lambda,
                double *force, double *dudl, double *energy) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  for (int k = 0; k < t->n_noe; k++) {
    const double *p = t->noe_par + 10 * k;
    const int i = t->noe_idx[2 * k], j = t->noe_idx[2 * k + 1];
    const Vec3 xj = j >= 0 ? getp(pos, n, j) : v3(p[7], p[8], p[9]);
    const Vec3 dr = minimg(getp(pos, n, i) - xj, L);
    const double r = norm(dr);
    double f = 0, e = 0;
    if (r < p[0]) {
      f = p[1] * (r - p[0]);
      e = 0.5 * f * (r - p[0]);
    } else if (r > p[2]) {
      const double d = r - p[2], rpk = p[4], rsw = p[5], nsw = p[6];
      if (rsw > 0 && d > rsw) {
        const double b = (rpk - rsw) / nsw * std::pow(rsw, nsw + 1);
        const double a = 0.5 * rsw * rsw - rpk * rsw - rsw * (rpk - rsw) / nsw;
        f = p[3] * (rpk - b * std::pow(d, -nsw - 1));
        e = p[3] * (a + b * std::pow(d, -nsw) + rpk * d);
      } else {
        f = p[3] * d;
        e = 0.5 * f * d;
      }
    }
    if (r > 0) {
      addf(force, n, i, (f / r) * dr);
      if (j >= 0) addf(force, n, j, (-f / r) * dr);
    }
    energy[BM_ENOE] += e;
  }
  for (int k = 0; k < t->n_harm; k++) {
    const double *p = t->harm_par + 5 * k;
    const int i = t->harm_idx[k];
    const Vec3 dr = minimg(getp(pos, n, i) - v3(p[1], p[2], p[3]), L);
    const double r2 = dot(dr, dr);
    const double krn2 = r2 > 0 ? p[0] * std::pow(r2, 0.5 * p[4] - 1) : 0.0;
    addf(force, n, i, (p[4] * krn2) * dr);
    energy[BM_EHARMONIC] += krn2 * r2;
  }
  auto scale = [&](int b) { return b ? lambda[b] : 1.0; };
  for (int k = 0; k < t->n_borest; k++) {
    const int i = t->borest_idx[2 * k], j = t->borest_idx[2 * k + 1], b = t->borest_block[k];
    const double kr = t->borest_par[2 * k], r0 = t->borest_par[2 * k + 1], l = scale(b);
    const Vec3 dr = minimg(getp(pos, n, i) - getp(pos, n, j), L);
    const double r = norm(dr), e = 0.5 * kr * (r - r0) * (r - r0), f = kr * (r - r0) * l;
    if (b) dudl[b] += e;
    addf(force, n, i, (f / r) * dr);
    addf(force, n, j, (-f / r) * dr);
    energy[BM_EMMFP] += e * l;
  }
  for (int k = 0; k < t->n_anrest; k++) {
    const int *a = t->anrest_idx + 3 * k, b = t->anrest_block[k];
    const double kt = t->anrest_par[2 * k], t0 = t->anrest_par[2 * k + 1], l = scale(b);
    const Vec3 rij = minimg(getp(pos, n, a[0]) - getp(pos, n, a[1]), L);
    const Vec3 rkj = minimg(getp(pos, n, a[2]) - getp(pos, n, a[1]), L);
    const Vec3 c = cross(rij, rkj);
    const double mc = norm(c), th = std::atan2(mc, dot(rij, rkj));
    const double e = 0.5 * kt * (th - t0) * (th - t0), f = kt * (th - t0) * l;
    if (b) dudl[b] += e;
    const Vec3 fi = (f / (mc * dot(rij, rij))) * cross(rij, c);
    const Vec3 fk = (-f / (mc * dot(rkj, rkj))) * cross(rkj, c);
    addf(force, n, a[0], fi);
    addf(force, n, a[2], fk);
    addf(force, n, a[1], -1.0 * (fi + fk));
    energy[BM_EMMFP] += e * l;
  }
  for (int k = 0; k < t->n_direst; k++) {
    const int *a = t->direst_idx + 4 * k, b = t->direst_block[k], np = t->direst_n[k];
    const double kp = t->direst_par[3 * k], p0 = t->direst_par[3 * k + 1], w = t->direst_par[3 * k + 2];
    const double l = scale(b);
    // phi from m = (j-i) x (k-j), nv = (k-j) x (l-k): atan2(b.(m x nv)/|..|, m.nv)
    const Vec3 rij = minimg(getp(pos, n, a[1]) - getp(pos, n, a[0]), L);
    const Vec3 rjk = minimg(getp(pos, n, a[2]) - getp(pos, n, a[1]), L);
    const Vec3 rkl = minimg(getp(pos, n, a[3]) - getp(pos, n, a[2]), L);
    const Vec3 m = cross(rij, rjk), nv = cross(rjk, rkl);
    const double mm = norm(m), nn = norm(nv), bb = norm(rjk);
    const double phi = std::atan2(dot((1 / bb) * rjk, cross((1 / mm) * m, (1 / nn) * nv)), dot(m, nv) / (mm * nn));
    double fphi, e;
    if (np > 0) {
      double d = np * (phi - p0);
      d -= 2 * M_PI * std::floor((d + M_PI) / (2 * M_PI));
      fphi = kp * np * std::sin(d);
      e = kp * (1 - std::cos(d));
    } else {
      double d = phi - p0;
      d -= 2 * M_PI * std::floor((d + M_PI) / (2 * M_PI));
      if (w > 0) d = std::fabs(d) <= w ? 0.0 : d - std::copysign(w, d);
      fphi = kp * d;
      e = 0.5 * kp * d * d;
    }
    if (b) dudl[b] += e;
    fphi *= l;
    // dphi/dx of this phi convention: -|b|/|m|^2 m on i, +|b|/|n|^2 n on l
    const Vec3 fi = (-fphi * bb / (mm * mm)) * m;
    const Vec3 fk0 = (-fphi * bb / (nn * nn)) * nv;
    const double pp = dot(rij, rjk) / (bb * bb), qq = dot(rkl, rjk) / (bb * bb);
    const Vec3 fj0 = (-pp) * fi + (-qq) * fk0;
    addf(force, n, a[0], fi);
    addf(force, n, a[3], -1.0 * fk0);
    addf(force, n, a[2], fk0 - fj0);
    addf(force, n, a[1], fj0 - fi);
    energy[BM_EMMFP] += e * l;
  }
  for (int k = 0; k < t->n_resd; k++) {
    const int *a = t->resd_idx + 4 * k;
    const double *p = t->resd_par + 4 * k;
    const Vec3 d1 = minimg(getp(pos, n, a[0]) - getp(pos, n, a[1]), L);
    const Vec3 d2 = minimg(getp(pos, n, a[2]) - getp(pos, n, a[3]), L);
    const double r1 = norm(d1), r2 = norm(d2), del = p[0] * r1 + p[1] * r2 - p[2], f = del * p[3];
    addf(force, n, a[0], (f * p[0] / r1) * d1);
    addf(force, n, a[1], (-f * p[0] / r1) * d1);
    addf(force, n, a[2], (f * p[1] / r2) * d2);
    addf(force, n, a[3], (-f * p[1] / r2) * d2);
    energy[BM_ERESD] += 0.5 * f * del;
  }
}

// ---------------------------------------------------------------------
// Holonomic constraints in FP64.
// ---------------------------------------------------------------------

// Newton-Raphson SHAKE for a star or triangle group (holonomic.cu:430-600
// iteration, generalised to any group of <=3 constraints).  Atoms are
// pre-shifted to the minimum image of atom 0.  Returns iterations used.
struct Group {
  int na, nc;
  int a[4];
  int ca[3], cb[3];
  double b0[3];
};

void solve3(int nc, double A[3][3], double *b, double *x) {
  // Gaussian elimination with partial pivoting, nc <= 3
  double M[3][4];
  for (int i = 0; i < nc; i++) {
    for (int j = 0; j < nc; j++) M[i][j] = A[i][j];
    M[i][nc] = b[i];
  }
  for (int c = 0; c < nc; c++) {
    int p = c;
    for (int r = c + 1; r < nc; r++)
      if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
    if (p != c)
      for (int j = 0; j <= nc; j++) std::swap(M[c][j], M[p][j]);
    for (int r = c + 1; r < nc; r++) {
      double f = M[r][c] / M[c][c];
      for (int j = c; j <= nc; j++) M[r][j] -= f * M[c][j];
    }
  }
  for (int r = nc - 1; r >= 0; r--) {
    double s = M[r][nc];
    for (int j = r + 1; j < nc; j++) s -= M[r][j] * x[j];
    x[r] = s / M[r][r];
  }
}

int shake_group(const Group &g, Vec3 *x, const Vec3 *xp, const double *im, double tol, int maxit) {
  Vec3 gp[3];
  for (int c = 0; c < g.nc; c++) gp[c] = xp[g.ca[c]] - xp[g.cb[c]];
  int it = 0;
  for (;;) {
    double res[3];
    bool done = true;
    for (int c = 0; c < g.nc; c++) {
      Vec3 d = x[g.ca[c]] - x[g.cb[c]];
      res[c] = g.b0[c] * g.b0[c] - dot(d, d);
      if (std::fabs(res[c]) >= 2 * g.b0[c] * g.b0[c] * tol) done = false;
    }
    if (done || it >= maxit) break;
    // d(res_c)/d(lambda_e): x_a += lambda_e im_a gp_e (a=ca_e), x_b -= lambda_e im_b gp_e
    double A[3][3];
    for (int c = 0; c < g.nc; c++) {
      Vec3 d = x[g.ca[c]] - x[g.cb[c]];
      for (int e = 0; e < g.nc; e++) {
        double coef = 0;
        if (g.ca[c] == g.ca[e]) coef += im[g.ca[c]];
        if (g.ca[c] == g.cb[e]) coef -= im[g.ca[c]];
        if (g.cb[c] == g.ca[e]) coef -= im[g.cb[c]];
        if (g.cb[c] == g.cb[e]) coef += im[g.cb[c]];
        A[c][e] = 2 * coef * dot(d, gp[e]);
      }
    }
    double lamv[3];
    solve3(g.nc, A, res, lamv);
    for (int e = 0; e < g.nc; e++) {
      x[g.ca[e]] = x[g.ca[e]] + (lamv[e] * im[g.ca[e]]) * gp[e];
      x[g.cb[e]] = x[g.cb[e]] + (-lamv[e] * im[g.cb[e]]) * gp[e];
    }
    it++;
  }
  return it;
}

// Analytic SETTLE for an isosceles rigid triangle (atom 0 apex, equal
// masses and lengths to atoms 1 and 2): rigid-body solution of the SHAKE
// equations (Miyamoto & Kollman 1992), the role of the reference's
// triangle kernel (holonomic.cu:253-377).
void settle(const double *b0, Vec3 *x, const Vec3 *xp, const double *mass) {
  double mA = mass[0], mB = mass[1], M = mA + 2 * mB;
  double dAB = b0[0], dBC = b0[2];
  double rc = 0.5 * dBC;
  double h = std::sqrt(dAB * dAB - rc * rc);
  double ra = 2 * mB * h / M, rb = h - ra;
  Vec3 d = (1.0 / M) * (mA * x[0] + mB * x[1] + mB * x[2]);
  Vec3 a1 = x[0] - d, b1 = x[1] - d, c1 = x[2] - d;
  Vec3 p[3] = {v3(0, 0, 0), xp[1] - xp[0], xp[2] - xp[0]};
  Vec3 ez = cross(p[1], p[2]);
  ez = (1.0 / norm(ez)) * ez;
  Vec3 ex = cross(a1, ez);
  ex = (1.0 / norm(ex)) * ex;
  Vec3 ey = cross(ez, ex);
  double az = dot(a1, ez), bz = dot(b1, ez), cz = dot(c1, ez);
  double sphi = az / ra, cphi = std::sqrt(1 - sphi * sphi);
  double spsi = (bz - cz) / (2 * rc * cphi), cpsi = std::sqrt(1 - spsi * spsi);
  // canonical triangle tilted by psi (about y) then phi (about x)
  double qx[3] = {0, -rc * cpsi, rc * cpsi};
  double qy[3] = {ra * cphi, -rb * cphi - rc * spsi * sphi, -rb * cphi + rc * spsi * sphi};
  double qz[3] = {az, bz, cz};
  double m[3] = {mA, mB, mB};
  Vec3 u[3] = {a1, b1, c1};
  double al = 0, be = 0, ga = 0;
  for (int i = 0; i < 3; i++) {
    double px = dot(p[i], ex), py = dot(p[i], ey);
    double ux = dot(u[i], ex), uy = dot(u[i], ey);
    al += m[i] * (px * qx[i] + py * qy[i]);
    be += m[i] * (px * qy[i] - py * qx[i]);
    ga += m[i] * (px * uy - py * ux);
  }
  double a2b2 = al * al + be * be;
  double sgn = al > 0 ? 1 : -1;
  double sth = (al * ga - sgn * be * std::sqrt(a2b2 - ga * ga)) / a2b2;
  double cth = std::sqrt(1 - sth * sth);
  for (int i = 0; i < 3; i++) {
    double fx = qx[i] * cth - qy[i] * sth, fy = qx[i] * sth + qy[i] * cth;
    x[i] = d + fx * ex + fy * ey + qz[i] * ez;
  }
}

template <class F>
void for_groups(const bm_topology *t, F fn) {
  for (int c = 0; c < t->n_tri; c++) {
    Group g{3, 3, {t->tri_idx[3 * c], t->tri_idx[3 * c + 1], t->tri_idx[3 * c + 2], -1},
            {0, 0, 1}, {1, 2, 2}, {t->tri_b0[3 * c], t->tri_b0[3 * c + 1], t->tri_b0[3 * c + 2]}};
    fn(g, 0);
  }
  for (int c = 0; c < t->n_br1; c++) {
    Group g{2, 1, {t->br1_idx[2 * c], t->br1_idx[2 * c + 1], -1, -1}, {0}, {1}, {t->br1_b0[c]}};
    fn(g, 1);
  }
  for (int c = 0; c < t->n_br2; c++) {
    Group g{3, 2, {t->br2_idx[3 * c], t->br2_idx[3 * c + 1], t->br2_idx[3 * c + 2], -1},
            {0, 0}, {1, 2}, {t->br2_b0[2 * c], t->br2_b0[2 * c + 1]}};
    fn(g, 2);
  }
  for (int c = 0; c < t->n_br3; c++) {
    Group g{4, 3, {t->br3_idx[4 * c], t->br3_idx[4 * c + 1], t->br3_idx[4 * c + 2], t->br3_idx[4 * c + 3]},
            {0, 0, 0}, {1, 2, 3}, {t->br3_b0[3 * c], t->br3_b0[3 * c + 1], t->br3_b0[3 * c + 2]}};
    fn(g, 3);
  }
}

// Virtual sites (holonomic/virtual.cu:11-254).  The site position as a
// function of its hosts' positions, evaluated in forward-mode derivatives:
// each coordinate carries its gradient with respect to the 9 host
// coordinates, so the force on the hosts is the exact J^T F of the site's
// force F.
struct Dn {
  double v, g[9];
};
Dn dn(double v) { Dn r{v, {0}}; return r; }
Dn operator+(const Dn &a, const Dn &b) { Dn r{a.v + b.v, {0}}; for (int k = 0; k < 9; k++) r.g[k] = a.g[k] + b.g[k]; return r; }
Dn operator-(const Dn &a, const Dn &b) { Dn r{a.v - b.v, {0}}; for (int k = 0; k < 9; k++) r.g[k] = a.g[k] - b.g[k]; return r; }
Dn operator*(const Dn &a, const Dn &b) {
  Dn r{a.v * b.v, {0}};
  for (int k = 0; k < 9; k++) r.g[k] = a.g[k] * b.v + a.v * b.g[k];
  return r;
}
Dn operator*(double s, const Dn &a) { Dn r{s * a.v, {0}}; for (int k = 0; k < 9; k++) r.g[k] = s * a.g[k]; return r; }
Dn dsqrt(const Dn &a) {
  Dn r{std::sqrt(a.v), {0}};
  for (int k = 0; k < 9; k++) r.g[k] = a.g[k] / (2 * r.v);
  return r;
}
Dn dinv(const Dn &a) {
  Dn r{1 / a.v, {0}};
  for (int k = 0; k < 9; k++) r.g[k] = -a.g[k] / (a.v * a.v);
  return r;
}
struct V3d { Dn x, y, z; };
V3d vsub(const V3d &a, const V3d &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3d vadd(const V3d &a, const V3d &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3d vscale(const Dn &s, const V3d &a) { return {s * a.x, s * a.y, s * a.z}; }
V3d vscale(double s, const V3d &a) { return {s * a.x, s * a.y, s * a.z}; }
Dn vdot(const V3d &a, const V3d &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3d vcross(const V3d &a, const V3d &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
V3d vunit(const V3d &a) { return vscale(dinv(dsqrt(vdot(a, a))), a); }

// offset of the site from its first host, hosts given relative to the
// first host (h0 = 0) with their derivative seeds
V3d virt_offset(int nh, const V3d &h1, const V3d &h2, double dist, double a, double b) {
  if (nh == 2) {
    const V3d d = vscale(-1.0, h1);   // x_h0 - x_h1
    const Dn s = dist * dinv(dsqrt(vdot(d, d))) + dn(a);
    return vscale(s, d);
  }
  V3d bb = vsub(h2, h1), p1 = h1;
  if (dist < 0) p1 = vadd(h1, vscale(0.5, bb));
  const V3d ax = vunit(p1);
  const V3d c = vunit(vcross(bb, ax));
  const V3d bp = vcross(ax, c);
  const V3d w = vadd(vscale(std::sin(b), c), vscale(std::cos(b), bp));
  return vscale(std::fabs(dist), vadd(vscale(std::cos(a), ax), vscale(std::sin(a), w)));
}

template <class F>
void for_virts(const bm_topology *t, F fn) {
  for (int k = 0; k < t->n_virt2; k++) {
    const int *ix = t->virt2_idx + 3 * k;
    fn(2, ix[0], ix[1], ix[2], -1, t->virt2_par[2 * k], t->virt2_par[2 * k + 1], 0.0);
  }
  for (int k = 0; k < t->n_virt3; k++) {
    const int *ix = t->virt3_idx + 4 * k;
    fn(3, ix[0], ix[1], ix[2], ix[3], t->virt3_par[3 * k], t->virt3_par[3 * k + 1], t->virt3_par[3 * k + 2]);
  }
}

// hosts 1 and 2 relative to host 0 (minimum image), seeded: derivative
// index 3*j+k is coordinate k of host j
void virt_hosts(const double *x, int n, const double *L, int h0, int h1, int h2, V3d &r1, V3d &r2) {
  const Vec3 p0 = getp(x, n, h0);
  const Vec3 d1 = minimg(getp(x, n, h1) - p0, L);
  const Vec3 d2 = h2 >= 0 ? minimg(getp(x, n, h2) - p0, L) : v3(0, 0, 0);
  const double c1[3] = {d1.x, d1.y, d1.z}, c2[3] = {d2.x, d2.y, d2.z};
  Dn *o1[3] = {&r1.x, &r1.y, &r1.z}, *o2[3] = {&r2.x, &r2.y, &r2.z};
  for (int k = 0; k < 3; k++) {
    *o1[k] = dn(c1[k]);
    o1[k]->g[3 + k] = 1.0;
    o1[k]->g[k] = -1.0;   // relative to host 0
    *o2[k] = dn(c2[k]);
    o2[k]->g[6 + k] = 1.0;
    o2[k]->g[k] = -1.0;
  }
}

void virt_place(const bm_topology *t, const bm_runparams *rp, double *x) {
  const int n = t->n_atoms;
  for_virts(t, [&](int nh, int v, int h0, int h1, int h2, double dist, double a, double b) {
    V3d r1, r2;
    virt_hosts(x, n, rp->box, h0, h1, h2, r1, r2);
    const V3d o = virt_offset(nh, r1, r2, dist, a, b);
    // in the periodic image nearest the site's previous position
    // (holonomic/virtual.cu:32-36)
    const Vec3 nw = v3(x[h0] + o.x.v, x[n + h0] + o.y.v, x[2 * n + h0] + o.z.v);
    const Vec3 step = nw - getp(x, n, v);
    const Vec3 p = nw + (minimg(step, rp->box) - step);
    x[v] = p.x;
    x[n + v] = p.y;
    x[2 * n + v] = p.z;
  });
}

// the site's force to its hosts, J^T F; the site's entry is cleared
void virt_spread(const bm_topology *t, const bm_runparams *rp, const double *x, double *f) {
  const int n = t->n_atoms;
  for_virts(t, [&](int nh, int v, int h0, int h1, int h2, double dist, double a, double b) {
    V3d r1, r2;
    virt_hosts(x, n, rp->box, h0, h1, h2, r1, r2);
    const V3d o = virt_offset(nh, r1, r2, dist, a, b);
    const double F[3] = {f[v], f[n + v], f[2 * n + v]};
    const Dn *oc[3] = {&o.x, &o.y, &o.z};
    const int hosts[3] = {h0, h1, h2};
    for (int j = 0; j < nh; j++)
      for (int k = 0; k < 3; k++) {
        // x_v = x_h0 + o: host 0 also carries F itself
        double g = j == 0 ? F[k] : 0.0;
        for (int i = 0; i < 3; i++) g += F[i] * oc[i]->g[3 * j + k];
        f[k * n + hosts[j]] += g;
      }
    f[v] = f[n + v] = f[2 * n + v] = 0.0;
  });
}

}  // namespace

// =====================================================================
// C ABI
// =====================================================================
extern "C" {

int orc_threads(void) { return omp_get_max_threads(); }

void orc_nbex_pair(double kqq, double beta, double r, double *fr, double *e) { nbex_pair(kqq, beta, r, *fr, *e); }

long orc_forces(const bm_topology *t, const bm_runparams *rp, const double *pos,
                const double *lambda, unsigned terms, double *force, double *dudl, double *energy) {
  int n = t->n_atoms;
  std::fill(force, force + 3 * n, 0.0);
  std::fill(dudl, dudl + t->n_blocks, 0.0);
  std::fill(energy, energy + BM_NTERMS, 0.0);
  long pairs = 0;
  if (terms & ORC_NBDIRECT) pairs = nbdirect(t, rp, pos, lambda, force, dudl, energy);
  if (rp->elec_method == BM_PME) {
    if (terms & ORC_EWALDSELF) ewaldself(t, rp, lambda, dudl, energy);
    if (terms & ORC_PME) pme(t, rp, pos, lambda, force, dudl, energy);
    if (terms & ORC_NBEX) nbex(t, rp, pos, lambda, force, dudl, energy);
  }
  if (terms & ORC_BONDED) bonded(t, rp, pos, lambda, force, dudl, energy);
  if ((terms & ORC_BONDED) && t->n_cmaps > 0) cmap(t, rp, pos, lambda, force, dudl, energy);
  if (terms & ORC_NB14) nb14(t, rp, pos, lambda, force, dudl, energy);
  if ((terms & ORC_MSLDBIAS) && t->n_blocks > 1) msldbias(t, lambda, dudl, energy);
  if ((terms & ORC_MSLDBIAS) && t->n_cats > 0) cats(t, rp, pos, force, energy);
  if ((terms & ORC_MSLDBIAS) && t->n_blocks > 1 && (t->cr_k1 != 0 || t->cr_k2 != 0 || t->cr_k3 != 0))
    chargerestraints(t, rp, lambda, dudl, energy);
  if (terms & ORC_RESTRAIN) restraints(t, rp, pos, lambda, force, dudl, energy);
  // forces on virtual sites move to their hosts (system/potential.cxx:1938)
  if (t->n_virt2 + t->n_virt3 > 0) virt_spread(t, rp, pos, force);
  double pot = 0;
  for (int i = 0; i < BM_EPOTENTIAL; i++) pot += energy[i];
  energy[BM_EPOTENTIAL] = pot;
  return pairs;
}

double orc_ewald_recip_direct(const bm_topology *t, const bm_runparams *rp, const double *pos,
                              const double *lambda, int kmax, double *force, double *dudl) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  double V = L[0] * L[1] * L[2], beta = rp->beta_ewald;
  std::vector<double> lq(n);
  for (int i = 0; i < n; i++) {
    int b = blockOf(t->site_block[i]);
    lq[i] = (b ? lambda[b] : 1.0) * t->charge[i];
  }
  if (force) std::fill(force, force + 3 * n, 0.0);
  double etot = 0;
  int nthr = omp_get_max_threads();
  std::vector<double> fT((size_t)nthr * 3 * n, 0.0), dT((size_t)nthr * t->n_blocks, 0.0);
  int span = 2 * kmax + 1;
#pragma omp parallel for reduction(+ : etot) schedule(dynamic, 4)
  for (int idx = 0; idx < span * span * span; idx++) {
    int mx = idx / (span * span) - kmax, my = (idx / span) % span - kmax, mz = idx % span - kmax;
    if (!mx && !my && !mz) continue;
    double kx = mx / L[0], ky = my / L[1], kz = mz / L[2];
    double m2 = kx * kx + ky * ky + kz * kz;
    double g = kElectric / (2 * M_PI * V) * std::exp(-M_PI * M_PI * m2 / (beta * beta)) / m2;
    double C = 0, S = 0;
    for (int i = 0; i < n; i++) {
      double th = 2 * M_PI * (kx * pos[i] + ky * pos[n + i] + kz * pos[2 * n + i]);
      C += lq[i] * std::cos(th);
      S += lq[i] * std::sin(th);
    }
    etot += g * (C * C + S * S);
    double *f = &fT[(size_t)omp_get_thread_num() * 3 * n];
    double *dd = &dT[(size_t)omp_get_thread_num() * t->n_blocks];
    for (int i = 0; i < n; i++) {
      double th = 2 * M_PI * (kx * pos[i] + ky * pos[n + i] + kz * pos[2 * n + i]);
      double cs = std::cos(th), sn = std::sin(th);
      double w = g * 4 * M_PI * lq[i] * (S * cs - C * sn);
      f[i] += w * kx;
      f[n + i] += w * ky;
      f[2 * n + i] += w * kz;
      int b = blockOf(t->site_block[i]);
      if (b) dd[b] += g * 2 * t->charge[i] * (C * cs + S * sn);
    }
  }
  for (int th = 0; th < nthr; th++) {
    if (force)
      for (int i = 0; i < 3 * n; i++) force[i] += fT[(size_t)th * 3 * n + i];
    if (dudl)
      for (int b = 0; b < t->n_blocks; b++) dudl[b] += dT[(size_t)th * t->n_blocks + b];
  }
  return etot;
}

double orc_ewald_real_images(int n, const double *q, const double *pos, const double *box,
                             double beta, double rcut) {
  int nim[3];
  for (int d = 0; d < 3; d++) nim[d] = (int)std::ceil(rcut / box[d]) + 1;
  double e = 0;
#pragma omp parallel for reduction(+ : e)
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++)
      for (int a = -nim[0]; a <= nim[0]; a++)
        for (int b = -nim[1]; b <= nim[1]; b++)
          for (int c = -nim[2]; c <= nim[2]; c++) {
            if (i == j && !a && !b && !c) continue;
            double dx = pos[i] - pos[j] + a * box[0], dy = pos[n + i] - pos[n + j] + b * box[1],
                   dz = pos[2 * n + i] - pos[2 * n + j] + c * box[2];
            double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r < rcut) e += 0.5 * kElectric * q[i] * q[j] * std::erfc(beta * r) / r;
          }
  }
  return e;
}

double orc_bspline(int order, double u) { return bspline(order, u); }
void orc_bspline_moduli(int order, int K, double *out) { bsp_moduli(order, K, out); }

void orc_fft3d(int kx, int ky, int kz, double *data, int sign) {
  fft3d(kx, ky, kz, reinterpret_cast<cplx *>(data), sign);
}

void orc_domdec_columns(const bm_runparams *rp, int n_atoms, int *ncx, int *ncy) {
  // approxBlockBox = (32 V / N)^(1/3); columns = ceil(L / approxBlockBox)
  // (domdec/domdec.cu:110-118)
  double V = rp->box[0] * rp->box[1] * rp->box[2];
  double a = std::cbrt(32.0 * V / n_atoms);
  *ncx = std::max(1, (int)std::ceil(rp->box[0] / a));
  *ncy = std::max(1, (int)std::ceil(rp->box[1] / a));
}

long orc_domdec(const bm_topology *t, const bm_runparams *rp, double *pos, int ncx, int ncy,
                int max_blocks, long max_tiles, int *l2g, int *bcount, int *n_blocks,
                int *tile_i, int *tile_j, unsigned *tile_mask) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  std::vector<uint32_t> u(3 * (size_t)n);
  // wrap into [0,L) and take 32-bit fractional coordinates
  for (int d = 0; d < 3; d++) {
    double invL = 1.0 / L[d];
    for (int i = 0; i < n; i++) {
      double x = pos[d * n + i];
      double k = std::floor(x * invL);
      double kl = k * L[d];
      x = x - kl;
      pos[d * n + i] = x;
      double f = x * invL;
      double s = f * 4294967296.0;
      u[(size_t)d * n + i] = (uint32_t)(int64_t)std::floor(s);
    }
  }
  int ncol = ncx * ncy;
  std::vector<std::vector<int>> col(ncol);
  for (int i = 0; i < n; i++) {
    uint32_t cx = (uint32_t)(((uint64_t)u[i] * (uint64_t)ncx) >> 32);
    uint32_t cy = (uint32_t)(((uint64_t)u[n + i] * (uint64_t)ncy) >> 32);
    col[cx * ncy + cy].push_back(i);
  }
  int nb = 0;
  for (int c = 0; c < ncol; c++) nb += ((int)col[c].size() + 31) / 32;
  if (nb > max_blocks) return -1;
  *n_blocks = nb;
  std::vector<uint32_t> lo(3 * (size_t)nb), hi(3 * (size_t)nb);
  std::vector<int> slotOf(n);
  int b = 0;
  for (int c = 0; c < ncol; c++) {
    auto &v = col[c];
    std::sort(v.begin(), v.end(), [&](int p, int q) {
      uint32_t zp = u[2 * (size_t)n + p], zq = u[2 * (size_t)n + q];
      return zp != zq ? zp < zq : p < q;
    });
    for (size_t s = 0; s < v.size(); s += 32, b++) {
      int cnt = (int)std::min<size_t>(32, v.size() - s);
      bcount[b] = cnt;
      for (int l = 0; l < 32; l++) l2g[32 * b + l] = l < cnt ? v[s + l] : -1;
      for (int d = 0; d < 3; d++) {
        uint32_t mn = 0xFFFFFFFFu, mx = 0;
        for (int l = 0; l < cnt; l++) {
          uint32_t w = u[(size_t)d * n + v[s + l]];
          mn = std::min(mn, w);
          mx = std::max(mx, w);
        }
        lo[3 * b + d] = mn;
        hi[3 * b + d] = mx;
      }
      for (int l = 0; l < cnt; l++) slotOf[v[s + l]] = 32 * b + l;
    }
  }
  // directed tile list: for every block I, every block J (J == I included)
  // whose rebuild box lies within rcut + pad (periodic interval gap), in
  // ascending J; the box test is symmetric, so (I,J) and (J,I) both appear
  const double rl = rp->rcut + rp->cull_pad;
  const double rl2 = rl * rl;
  double scale[3];
  for (int d = 0; d < 3; d++) scale[d] = L[d] * (1.0 / 4294967296.0);
  long nt = 0;
  std::vector<long> firstTile(nb + 1);
  for (int I = 0; I < nb; I++) {
    firstTile[I] = nt;
    for (int J = 0; J < nb; J++) {
      bool keep = (I == J);
      if (!keep) {
        double g[3];
        for (int d = 0; d < 3; d++) {
          uint64_t aL = lo[3 * I + d], aH = hi[3 * I + d], bL = lo[3 * J + d], bH = hi[3 * J + d];
          uint64_t gap = 0;
          const uint64_t P = 1ull << 32;
          if (aH < bL) gap = std::min(bL - aH, aL + P - bH);
          else if (bH < aL) gap = std::min(aL - bH, bL + P - aH);
          g[d] = (double)gap * scale[d];
        }
        double gx = g[0] * g[0], gy = g[1] * g[1], gz = g[2] * g[2];
        double s = gx + gy;
        s = s + gz;
        keep = s <= rl2;
      }
      if (keep) {
        if (nt >= max_tiles) return -1;
        tile_i[nt] = I;
        tile_j[nt] = J;
        for (int l = 0; l < 32; l++) tile_mask[32 * nt + l] = 0;
        nt++;
      }
    }
  }
  firstTile[nb] = nt;
  // exclusion masks: bit j of word i of tile (I,J) marks slots (I,i), (J,j)
  auto tileOf = [&](int I, int J) -> long {
    const int *beg = tile_j + firstTile[I], *end = tile_j + firstTile[I + 1];
    const int *it = std::lower_bound(beg, end, J);
    return (it == end || *it != J) ? -1 : firstTile[I] + (it - beg);
  };
  for (int e = 0; e < t->n_excl; e++) {
    const int sa = slotOf[t->excl_idx[2 * e]], sb = slotOf[t->excl_idx[2 * e + 1]];
    const int I = sa / 32, J = sb / 32;
    const long ta = tileOf(I, J);
    if (ta < 0) continue;  // never within range of each other
    tile_mask[32 * ta + sa % 32] |= 1u << (sb % 32);
    const long tb = tileOf(J, I);
    tile_mask[32 * tb + sb % 32] |= 1u << (sa % 32);
  }
  return nt;
}

void orc_theta_to_lambda(const bm_topology *t, const bm_runparams *rp, const double *theta,
                         double *lambda) {
  if (t->fix_lambda) {
    // the fix flag: lambda is theta as given (msld.cu:924-934)
    for (int b = 0; b < t->n_blocks; b++) lambda[b] = theta[b];
    return;
  }
  for (int s = 0; s < t->n_sites; s++) {
    int b0 = t->site_bound[s], b1 = t->site_bound[s + 1];
    if (t->piecewise) {
      const double t0 = piecewise_root(theta + b0, b1 - b0);
      for (int b = b0; b < b1; b++) {
        const double d = theta[b] - t0;
        lambda[b] = d > 0 ? d * d * d : 0.0;
      }
      continue;
    }
    double fix = 0, norm = 0;
    for (int b = b0; b < b1; b++) {
      bool fixed = t->block_fixed && t->block_fixed[b];
      if (fixed) fix += theta[b];
      else {
        lambda[b] = std::exp(rp->fnex * std::sin(theta[b]));
        norm += lambda[b];
      }
    }
    double r = 1 - fix;
    if (norm > 0) norm = r / norm;
    for (int b = b0; b < b1; b++) {
      bool fixed = t->block_fixed && t->block_fixed[b];
      lambda[b] = fixed ? theta[b] : lambda[b] * norm;
    }
  }
}

void orc_chain_rule(const bm_topology *t, const bm_runparams *rp, const double *theta,
                    const double *lambda, const double *dudl, double *dudtheta) {
  for (int b = 0; b < t->n_blocks; b++) {
    int s = t->block_site[b];
    bool fixedb = t->block_fixed && t->block_fixed[b];
    if (fixedb || t->fix_lambda) { dudtheta[b] = 0; continue; }
    if (t->piecewise) {
      // dlambda_i/dtheta_b = g_i (delta_ib - g_b / G), g = 3 max(theta - t, 0)^2,
      // G = sum g (the root t moves with theta_b by g_b / G; msld.cu:968-990)
      const int c0 = t->site_bound[s], c1 = t->site_bound[s + 1];
      const double t0 = piecewise_root(theta + c0, c1 - c0);
      double G = 0, dot = 0;
      for (int c = c0; c < c1; c++) {
        const double d = theta[c] - t0, g = d > 0 ? 3 * d * d : 0.0;
        G += g;
        dot += g * dudl[c];
      }
      const double d = theta[b] - t0, gb = d > 0 ? 3 * d * d : 0.0;
      dudtheta[b] = gb * (dudl[b] - dot / G);
      continue;
    }
    double fix = 0, dyn = 0;
    for (int c = t->site_bound[s]; c < t->site_bound[s + 1]; c++) {
      bool fixedc = t->block_fixed && t->block_fixed[c];
      if (fixedc) fix += lambda[c];
      else dyn += lambda[c] * dudl[c];
    }
    double r = 1 - fix, f = dudl[b];
    if (r > 0) f -= dyn / r;
    dudtheta[b] = f * lambda[b] * rp->fnex * std::cos(theta[b]);
  }
}

void orc_theta_bias(const bm_topology *t, const bm_runparams *rp, const double *theta, double *dudtheta,
                    double *energy) {
  (void)rp;
  double e = 0;
  for (int s = 0; s < t->n_sites; s++) {
    const int b0 = t->site_bound[s], b1 = t->site_bound[s + 1], nb = b1 - b0;
    if (t->tcoll_k && t->tcoll_k[s] != 0) {
      // collective bias (msld.cu:1155-1181): the reference's force is
      // k n (A h^(n-1) - B l^(n-1)) cos theta, twice the theta derivative of
      // the energy k/2 (A^2 + B^2) it reports; both are kept as they are
      const double k = t->tcoll_k[s], n = t->tcoll_n[s];
      double A = -1, B = -(nb - 1);
      for (int b = b0; b < b1; b++) {
        A += std::pow(0.5 + 0.5 * std::sin(theta[b]), n);
        B += std::pow(0.5 - 0.5 * std::sin(theta[b]), n);
      }
      e += 0.5 * k * (A * A + B * B);
      for (int b = b0; b < b1; b++)
        dudtheta[b] += k * n * (A * std::pow(0.5 + 0.5 * std::sin(theta[b]), n - 1) -
                                B * std::pow(0.5 - 0.5 * std::sin(theta[b]), n - 1)) * std::cos(theta[b]);
    }
    if (t->tinde_k && t->tinde_k[s] != 0) {
      // independent bias (msld.cu:1182-1199), every block of the site
      const double k = t->tinde_k[s];
      for (int b = b0; b < b1; b++) {
        const double sb = 0.5 - 0.5 * std::sin(theta[b]);
        e += k * (1 - sb * sb * sb * sb);
        dudtheta[b] += 2 * k * sb * sb * sb * std::cos(theta[b]);
      }
    }
    if (t->piecewise && s > 0) {
      // flat-bottom harmonic beyond +-nb*width (msld.cu:1201-1226)
      for (int b = b0; b < b1; b++) {
        const double w = nb * t->pw_width;
        double d = 0;
        if (theta[b] > w) d = theta[b] - w;
        else if (theta[b] < -w) d = theta[b] + w;
        e += 0.5 * t->pw_k * d * d;
        dudtheta[b] += t->pw_k * d;
      }
      if (std::fabs(t->edge_k) > 1e-4) {
        // edge bias on the smooth substituent count (msld.cu:1228-1264)
        const double t0 = piecewise_root(theta + b0, nb);
        double G = 0;
        for (int b = b0; b < b1; b++) {
          const double d = theta[b] - t0;
          G += d > 0 ? 3 * d * d : 0.0;
        }
        double N = 0;
        std::vector<double> sg(nb);
        for (int j = 0; j < nb; j++) {
          const double x = std::exp(-t->edge_alpha * (theta[b0 + j] - t0 + t->edge_phi));
          N += 1 / (1 + x);
          sg[j] = t->edge_alpha * x / ((1 + x) * (1 + x));
        }
        const double over = N > t->edge_n ? N - t->edge_n : 0.0;
        e += 0.5 * t->edge_k * over * over;
        double ssum = 0;
        for (int j = 0; j < nb; j++) ssum += sg[j];
        for (int b = b0; b < b1; b++) {
          const double d = theta[b] - t0, gb = d > 0 ? 3 * d * d : 0.0;
          dudtheta[b] += t->edge_k * over * (sg[b - b0] - (gb / G) * ssum);
        }
      }
    }
  }
  energy[BM_ETHETA] += e;
}

void orc_holonomic_position(const bm_topology *t, const bm_runparams *rp, double *x,
                            const double *xprev, double *v, double halfdt) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  for_groups(t, [&](const Group &g, int kind) {
    Vec3 xp[4], xn[4], shift[4];
    double im[4], m[4];
    Vec3 p0 = getp(xprev, n, g.a[0]);
    for (int k = 0; k < g.na; k++) {
      Vec3 pk = getp(xprev, n, g.a[k]);
      shift[k] = k ? (minimg(pk - p0, L) - (pk - p0)) : v3(0, 0, 0);
      xp[k] = pk + shift[k];
      xn[k] = getp(x, n, g.a[k]) + shift[k];
      m[k] = t->mass[g.a[k]];
      im[k] = 1.0 / m[k];
    }
    bool iso = kind == 0 && m[1] == m[2] && g.b0[0] == g.b0[1];
    if (iso) settle(g.b0, xn, xp, m);
    else if (kind == 1) {
      // single bond: closed form of the quadratic (holonomic.cu:397-403)
      Vec3 gp = xp[0] - xp[1], s = xn[0] - xn[1];
      double gg = dot(gp, gp), sg = dot(s, gp), ss = dot(s, s), b02 = g.b0[0] * g.b0[0];
      double lw = (-sg + std::sqrt(sg * sg - gg * (ss - b02))) / (gg * (im[0] + im[1]));
      xn[0] = xn[0] + (im[0] * lw) * gp;
      xn[1] = xn[1] + (-im[1] * lw) * gp;
    } else {
      shake_group(g, xn, xp, im, rp->shake_tol, 10);
    }
    for (int k = 0; k < g.na; k++) {
      int a = g.a[k];
      Vec3 vv = (1.0 / halfdt) * (xn[k] - xp[k]);
      v[a] = vv.x; v[n + a] = vv.y; v[2 * n + a] = vv.z;
      Vec3 xo = xn[k] - shift[k];
      x[a] = xo.x; x[n + a] = xo.y; x[2 * n + a] = xo.z;
    }
  });
}

// holonomic_rectify (holonomic/rectify.cu:13-197), run before the first
// step: in each constraint group the atoms other than the first are put at
// exactly their constrained distances from the first (its minimum image).
// A branch atom keeps its direction from the first atom.  In a triangle the
// second atom keeps its direction and the third goes into the plane of the
// three, on its own side of the 0-1 line, at distances b02 and b12.
void orc_rectify(const bm_topology *t, const bm_runparams *rp, double *x) {
  const int n = t->n_atoms;
  const double *L = rp->box;
  for_groups(t, [&](const Group &g, int kind) {
    const Vec3 p0 = getp(x, n, g.a[0]);
    Vec3 d[3], shift[3], q[3];
    for (int k = 1; k < g.na; k++) {
      const Vec3 pk = getp(x, n, g.a[k]);
      d[k - 1] = minimg(pk - p0, L);
      shift[k - 1] = d[k - 1] - (pk - p0);
    }
    if (kind == 0) {
      const Vec3 e0 = (1.0 / norm(d[0])) * d[0];
      const Vec3 c = cross(d[0], d[1]);
      const Vec3 e2 = (1.0 / norm(c)) * c;
      const Vec3 e1 = cross(e2, e0);
      const double b01 = g.b0[0], b02 = g.b0[1], b12 = g.b0[2];
      const double along = (b01 * b01 + b02 * b02 - b12 * b12) / (2.0 * b01);
      const double across = std::sqrt(b02 * b02 - along * along);
      q[0] = b01 * e0;
      q[1] = along * e0 + across * e1;
    } else {
      for (int k = 0; k < g.na - 1; k++) q[k] = (g.b0[k] / norm(d[k])) * d[k];
    }
    for (int k = 1; k < g.na; k++) {
      const Vec3 xo = (p0 + q[k - 1]) - shift[k - 1];
      const int a = g.a[k];
      x[a] = xo.x; x[n + a] = xo.y; x[2 * n + a] = xo.z;
    }
  });
}

void orc_virtual_place(const bm_topology *t, const bm_runparams *rp, double *x) { virt_place(t, rp, x); }

void orc_holonomic_velocity(const bm_topology *t, const bm_runparams *rp, const double *x, double *v) {
  const int n = t->n_atoms;
  for_groups(t, [&](const Group &g, int) {
    Vec3 xs[4], vs[4];
    double im[4];
    for (int k = 0; k < g.na; k++) {
      xs[k] = getp(x, n, g.a[k]);
      vs[k] = getp(v, n, g.a[k]);
      im[k] = 1.0 / t->mass[g.a[k]];
    }
    Vec3 gc[3];
    for (int c = 0; c < g.nc; c++) gc[c] = minimg(xs[g.ca[c]] - xs[g.cb[c]], rp->box);
    double A[3][3], b[3], lamv[3];
    for (int c = 0; c < g.nc; c++) {
      b[c] = -dot(vs[g.ca[c]] - vs[g.cb[c]], gc[c]);
      for (int e = 0; e < g.nc; e++) {
        double coef = 0;
        if (g.ca[c] == g.ca[e]) coef += im[g.ca[c]];
        if (g.ca[c] == g.cb[e]) coef -= im[g.ca[c]];
        if (g.cb[c] == g.ca[e]) coef -= im[g.cb[c]];
        if (g.cb[c] == g.cb[e]) coef += im[g.cb[c]];
        A[c][e] = coef * dot(gc[c], gc[e]);
      }
    }
    solve3(g.nc, A, b, lamv);
    for (int e = 0; e < g.nc; e++) {
      vs[g.ca[e]] = vs[g.ca[e]] + (lamv[e] * im[g.ca[e]]) * gc[e];
      vs[g.cb[e]] = vs[g.cb[e]] + (-lamv[e] * im[g.cb[e]]) * gc[e];
    }
    for (int k = 0; k < g.na; k++) {
      int a = g.a[k];
      v[a] = vs[k].x; v[n + a] = vs[k].y; v[2 * n + a] = vs[k].z;
    }
  });
}

void orc_philox4x32(const unsigned ctr[4], const unsigned key[2], unsigned out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    c0 = h1 ^ c1 ^ k0;
    c1 = l1;
    c2 = h0 ^ c3 ^ k1;
    c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Normal deviate for (seed, step, dof): Philox block on counter
// (dof/2, step_lo, step_hi, 0x4E524D4C), Box-Muller on two 53-bit uniforms.
double orc_normal(uint64_t seed, uint64_t step, uint64_t dof) {
  unsigned ctr[4] = {(unsigned)(dof >> 1), (unsigned)step, (unsigned)(step >> 32), 0x4E524D4Cu};
  unsigned key[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
  unsigned o[4];
  orc_philox4x32(ctr, key, o);
  double u1 = ((double)((((uint64_t)o[0]) << 21) ^ (o[1] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)((((uint64_t)o[2]) << 21) ^ (o[3] >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double rad = std::sqrt(-2.0 * std::log(u1));
  double ang = 2.0 * M_PI * u2;
  return (dof & 1) ? rad * std::sin(ang) : rad * std::cos(ang);
}

int orc_step(const bm_topology *t, const bm_runparams *rp, long step, int energy_step,
             double *x, double *v, double *lambda, double *energy) {
  const int n = t->n_atoms, nl = t->n_blocks, N1 = 3 * n, N = 3 * n + nl;
  std::vector<double> f(N, 0.0), dudl(nl, 0.0), dudth(nl, 0.0), ism(N), xb(N);
  // virtual sites follow their hosts before the forces (domdec/domdec.cu:215)
  if (t->n_virt2 + t->n_virt3 > 0) virt_place(t, rp, x);
  orc_forces(t, rp, x, lambda, ORC_ALL, f.data(), dudl.data(), energy);
  if (nl > 1) {
    orc_chain_rule(t, rp, x + N1, lambda, dudl.data(), dudth.data());
    orc_theta_bias(t, rp, x + N1, dudth.data(), energy);
    for (int b = 0; b < nl; b++) f[N1 + b] = dudth[b];
    double pot = 0;
    for (int i = 0; i < BM_EPOTENTIAL; i++) pot += energy[i];
    energy[BM_EPOTENTIAL] = pot;
  }
  // massless particles (virtual sites) are not integrated: their inverse
  // square-root mass is infinite in the reference, which skips any
  // coordinate whose value is not finite (update/update.cu:51)
  for (int i = 0; i < n; i++) ism[i] = ism[n + i] = ism[2 * n + i] = t->mass[i] > 0 ? 1 / std::sqrt(t->mass[i]) : 0.0;
  for (int b = 0; b < nl; b++) ism[N1 + b] = t->theta_mass[b] > 0 ? 1 / std::sqrt(t->theta_mass[b]) : 0;
  const double kT = kBoltz * rp->temperature, halfdt = 0.5 * rp->dt;
  double a = std::exp(-rp->gamma * rp->dt), noise = std::sqrt((1 - a * a) * kT);
  double al = std::exp(-rp->msld_gamma * rp->dt), noisel = std::sqrt((1 - al * al) * kT);
  bool cons = t->n_tri + t->n_br1 + t->n_br2 + t->n_br3 > 0;
  if (energy_step) {
    for (int i = 0; i < N; i++)
      if (ism[i] > 0) v[i] -= halfdt * ism[i] * ism[i] * f[i];
    if (cons) orc_holonomic_velocity(t, rp, x, v);
    double ke = 0;
    for (int i = 0; i < N; i++)
      if (ism[i] > 0) ke += 0.5 * (v[i] / ism[i]) * (v[i] / ism[i]);
    energy[BM_EKINETIC] = ke;
    for (int i = 0; i < N; i++)
      if (ism[i] > 0) v[i] -= halfdt * ism[i] * ism[i] * f[i];
  } else {
    for (int i = 0; i < N; i++)
      if (ism[i] > 0) v[i] -= 2 * halfdt * ism[i] * ism[i] * f[i];
  }
  for (int i = 0; i < N; i++) {
    xb[i] = x[i];
    if (ism[i] > 0) x[i] += halfdt * v[i];
  }
  if (cons) orc_holonomic_position(t, rp, x, xb.data(), v, halfdt);
  for (int i = 0; i < N; i++) {
    double fr = a, no = noise;
    if (i >= N1) {
      int b = i - N1;
      if (t->theta_friction) {
        double ab = std::exp(-t->theta_friction[b] * rp->dt);
        fr = ab;
        no = std::sqrt((1 - ab * ab) * kT);
      } else {
        fr = al;
        no = noisel;
      }
    }
    xb[i] = x[i];
    if (ism[i] > 0) {
      v[i] = fr * v[i] + no * ism[i] * orc_normal(rp->seed, (uint64_t)step, (uint64_t)i);
      x[i] += halfdt * v[i];
    }
  }
  if (cons) orc_holonomic_position(t, rp, x, xb.data(), v, halfdt);
  if (nl > 1) {
    if (t->fix_lambda)
      for (int b = 0; b < nl; b++) x[N1 + b] = lambda[b];   // theta reset to the held lambda
    else
      orc_theta_to_lambda(t, rp, x + N1, lambda);
  }
  for (int i = 0; i < N; i++)
    if (!std::isfinite(v[i]) || !std::isfinite(x[i])) return -1;
  return 0;
}

}  // extern "C"
