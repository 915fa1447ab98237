// update.hip - VRORV Langevin integrator for atoms and theta, holonomic
// constraints, MSLD theta<->lambda and the Philox normal generator
// (restates update/update.cu:39-305, holonomic/holonomic.cu:15-758,
// msld/msld.cu:863-998).  FP64 throughout: positions, velocities,
// constraint iterations and the chain rule.
//
// One step (State::update):
//   upd_kick      chain rule dU/dl -> dU/dtheta, V (full or half kick),
//                 backup x, half drift R
//   upd_constrain SETTLE / SHAKE groups, v = dx/halfdt
//   upd_o_drift   O (friction + Philox noise), backup, half drift R
//   upd_constrain + theta -> lambda, step counter
// Energy steps split the kick: V, RATTLE velocity projection, kinetic
// energy, V+R (update.cu:257-279).
#include "kernels.h"

namespace blade {

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 d3(double a, double b, double c) { return D3{a, b, c}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return d3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 crossd(D3 a, D3 b) {
  return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double normd(D3 a) { return sqrt(dotd(a, a)); }
__device__ __forceinline__ D3 getp(const double *p, int n, int i) { return d3(p[i], p[n + i], p[2 * n + i]); }
__device__ __forceinline__ void setp(double *p, int n, int i, D3 v) { p[i] = v.x; p[n + i] = v.y; p[2 * n + i] = v.z; }
__device__ __forceinline__ double wrapd(double x, double L, double invL) { return x - L * floor(x * invL + 0.5); }
__device__ __forceinline__ D3 minimg(const Dev &d, D3 v) {
  return d3(wrapd(v.x, d.box[0], d.invbox[0]), wrapd(v.y, d.box[1], d.invbox[1]), wrapd(v.z, d.box[2], d.invbox[2]));
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -----------------
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
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

// standard normal for (seed, step, dof): the same rule as the oracle
__device__ double normal_dof(unsigned long long seed, unsigned long long step, unsigned long long dof) {
  uint4 c = make_uint4((unsigned)(dof >> 1), (unsigned)step, (unsigned)(step >> 32), 0x4E524D4Cu);
  uint4 o = philox(c, make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  double u1 = ((double)((((unsigned long long)o.x) << 21) ^ (o.y >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)((((unsigned long long)o.z) << 21) ^ (o.w >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  return (dof & 1) ? rad * sn : rad * cs;
}

// dU/dtheta of block b from the fixed-point dU/dlambda (msld.cu:937-975)
__device__ double theta_force(const Dev &d, int b) {
  if (d.block_fixed[b]) return 0.0;
  const int s = d.block_site[b];
  double fix = 0, dyn = 0;
  for (int c = d.site_bound[s]; c < d.site_bound[s + 1]; c++) {
    if (d.block_fixed[c]) fix += d.lambda[c];
    else dyn += d.lambda[c] * ((double)d.dudl[c] * kFixInvD);
  }
  double r = 1 - fix, f = (double)d.dudl[b] * kFixInvD;
  if (r > 0) f -= dyn / r;
  return f * d.lambda[b] * d.fnex * cos(d.pos[3 * d.n + b]);
}

__device__ __forceinline__ void nan_check(const Dev &d, double v, int i) {
  if (!isfinite(v)) atomicCAS(d.nanflag, -1, i);
}

// mode 0: v -= dt f/m, backup, x += dt/2 v; mode 1: v -= dt/2 f/m;
// mode 2: v -= dt/2 f/m, backup, x += dt/2 v
__global__ void upd_kick(Dev d, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int N1 = 3 * d.n, N = N1 + d.nl;
  if (i >= N) return;
  double f;
  if (i < N1) {
    f = (double)d.force[i] * kFixInvD;
  } else {
    f = theta_force(d, i - N1);
    if (mode != 2) d.force[i] = fix_d(f);
  }
  const double ism = d.ism[i];
  double x = d.pos[i];
  if (mode != 1) d.posb[i] = x;
  if (ism > 0) {
    double v = d.vel[i] - (mode == 0 ? 2.0 : 1.0) * d.halfdt * ism * ism * f;
    d.vel[i] = v;
    if (mode != 1) {
      d.pos[i] = x + d.halfdt * v;
      nan_check(d, v, i < N1 ? i % d.n : i);
    }
  }
}

__global__ void upd_o_drift(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int N1 = 3 * d.n, N = N1 + d.nl;
  if (i >= N) return;
  const double ism = d.ism[i];
  const double x = d.pos[i];
  d.posb[i] = x;
  if (ism > 0) {
    double fr = d.fric, no = d.noise;
    if (i >= N1) {
      fr = d.lfric[i - N1];
      no = d.lnoise[i - N1];
    }
    const double xi = normal_dof(d.seed, (unsigned long long)(*d.step), (unsigned long long)i);
    double v = fr * d.vel[i] + no * ism * xi;
    d.vel[i] = v;
    d.pos[i] = x + d.halfdt * v;
    nan_check(d, v, i < N1 ? i % d.n : i);
  }
}

// ---- constraints ------------------------------------------------------
struct Grp {
  int na, nc;
  int a[4];
  int ca[3], cb[3];
  double b0[3];
};

__device__ Grp load_group(const Dev &d, int g, int &kind) {
  Grp G;
  if (g < d.ntri) {
    kind = 0;
    int4 ix = d.tri_i[g];
    double4 b = d.tri_b[g];
    G = Grp{3, 3, {ix.x, ix.y, ix.z, -1}, {0, 0, 1}, {1, 2, 2}, {b.x, b.y, b.z}};
    return G;
  }
  g -= d.ntri;
  if (g < d.nbr1) {
    kind = 1;
    int2 ix = d.br1_i[g];
    G = Grp{2, 1, {ix.x, ix.y, -1, -1}, {0, 0, 0}, {1, 0, 0}, {d.br1_b[g], 0, 0}};
    return G;
  }
  g -= d.nbr1;
  if (g < d.nbr2) {
    kind = 2;
    int4 ix = d.br2_i[g];
    double2 b = d.br2_b[g];
    G = Grp{3, 2, {ix.x, ix.y, ix.z, -1}, {0, 0, 0}, {1, 2, 0}, {b.x, b.y, 0}};
    return G;
  }
  g -= d.nbr2;
  kind = 3;
  int4 ix = d.br3_i[g];
  double4 b = d.br3_b[g];
  G = Grp{4, 3, {ix.x, ix.y, ix.z, ix.w}, {0, 0, 0}, {1, 2, 3}, {b.x, b.y, b.z}};
  return G;
}

__device__ void solve3d(int nc, double A[3][3], const double *b, double *x) {
  double M[3][4];
  for (int i = 0; i < nc; i++) {
    for (int j = 0; j < nc; j++) M[i][j] = A[i][j];
    M[i][nc] = b[i];
  }
  for (int c = 0; c < nc; c++) {
    int p = c;
    for (int r = c + 1; r < nc; r++)
      if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
    if (p != c)
      for (int j = 0; j <= nc; j++) { double t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
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

// Newton-Raphson SHAKE over <=3 constraints (holonomic.cu:430-600)
__device__ void shake_newton(const Grp &g, D3 *x, const D3 *xp, const double *im, double tol) {
  D3 gp[3];
  for (int c = 0; c < g.nc; c++) gp[c] = xp[g.ca[c]] - xp[g.cb[c]];
  for (int it = 0;; it++) {
    double res[3];
    bool done = true;
    for (int c = 0; c < g.nc; c++) {
      D3 dd = x[g.ca[c]] - x[g.cb[c]];
      res[c] = g.b0[c] * g.b0[c] - dotd(dd, dd);
      if (fabs(res[c]) >= 2 * g.b0[c] * g.b0[c] * tol) done = false;
    }
    if (done || it >= 10) break;
    double A[3][3], lam[3];
    for (int c = 0; c < g.nc; c++) {
      D3 dd = x[g.ca[c]] - x[g.cb[c]];
      for (int e = 0; e < g.nc; e++) {
        double coef = 0;
        if (g.ca[c] == g.ca[e]) coef += im[g.ca[c]];
        if (g.ca[c] == g.cb[e]) coef -= im[g.ca[c]];
        if (g.cb[c] == g.ca[e]) coef -= im[g.cb[c]];
        if (g.cb[c] == g.cb[e]) coef += im[g.cb[c]];
        A[c][e] = 2 * coef * dotd(dd, gp[e]);
      }
    }
    solve3d(g.nc, A, res, lam);
    for (int e = 0; e < g.nc; e++) {
      x[g.ca[e]] = x[g.ca[e]] + (lam[e] * im[g.ca[e]]) * gp[e];
      x[g.cb[e]] = x[g.cb[e]] + (-lam[e] * im[g.cb[e]]) * gp[e];
    }
  }
}

// analytic SETTLE for an isosceles rigid triangle (the oracle's settle())
__device__ void settle(const double *b0, D3 *x, const D3 *xp, const double *mass) {
  const double mA = mass[0], mB = mass[1], M = mA + 2 * mB;
  const double rc = 0.5 * b0[2];
  const double h = sqrt(b0[0] * b0[0] - rc * rc);
  const double ra = 2 * mB * h / M, rb = h - ra;
  D3 dcm = (1.0 / M) * (mA * x[0] + mB * x[1] + mB * x[2]);
  D3 a1 = x[0] - dcm, b1 = x[1] - dcm, c1 = x[2] - dcm;
  D3 p[3] = {d3(0, 0, 0), xp[1] - xp[0], xp[2] - xp[0]};
  D3 ez = crossd(p[1], p[2]);
  ez = (1.0 / normd(ez)) * ez;
  D3 ex = crossd(a1, ez);
  ex = (1.0 / normd(ex)) * ex;
  D3 ey = crossd(ez, ex);
  const double az = dotd(a1, ez), bz = dotd(b1, ez), cz = dotd(c1, ez);
  const double sphi = az / ra, cphi = sqrt(1 - sphi * sphi);
  const double spsi = (bz - cz) / (2 * rc * cphi), cpsi = sqrt(1 - spsi * spsi);
  const double qx[3] = {0, -rc * cpsi, rc * cpsi};
  const double qy[3] = {ra * cphi, -rb * cphi - rc * spsi * sphi, -rb * cphi + rc * spsi * sphi};
  const double qz[3] = {az, bz, cz};
  const double m[3] = {mA, mB, mB};
  const D3 u[3] = {a1, b1, c1};
  double al = 0, be = 0, ga = 0;
  for (int i = 0; i < 3; i++) {
    double px = dotd(p[i], ex), py = dotd(p[i], ey);
    double ux = dotd(u[i], ex), uy = dotd(u[i], ey);
    al += m[i] * (px * qx[i] + py * qy[i]);
    be += m[i] * (px * qy[i] - py * qx[i]);
    ga += m[i] * (px * uy - py * ux);
  }
  const double a2b2 = al * al + be * be;
  const double sgn = al > 0 ? 1 : -1;
  const double sth = (al * ga - sgn * be * sqrt(a2b2 - ga * ga)) / a2b2;
  const double cth = sqrt(1 - sth * sth);
  for (int i = 0; i < 3; i++) {
    double fx = qx[i] * cth - qy[i] * sth, fy = qx[i] * sth + qy[i] * cth;
    x[i] = dcm + fx * ex + fy * ey + qz[i] * ez;
  }
}

__device__ void constrain_group(const Dev &d, int gi) {
  int kind;
  Grp g = load_group(d, gi, kind);
  D3 xp[4], xn[4], sh[4];
  double im[4], m[4];
  const D3 p0 = getp(d.posb, d.n, g.a[0]);
  for (int k = 0; k < g.na; k++) {
    D3 pk = getp(d.posb, d.n, g.a[k]);
    sh[k] = k ? (minimg(d, pk - p0) - (pk - p0)) : d3(0, 0, 0);
    xp[k] = pk + sh[k];
    xn[k] = getp(d.pos, d.n, g.a[k]) + sh[k];
    m[k] = d.mass[g.a[k]];
    im[k] = 1.0 / m[k];
  }
  if (kind == 0 && m[1] == m[2] && g.b0[0] == g.b0[1]) {
    settle(g.b0, xn, xp, m);
  } else if (kind == 1) {
    D3 gp = xp[0] - xp[1], s = xn[0] - xn[1];
    double gg = dotd(gp, gp), sg = dotd(s, gp), ss = dotd(s, s), b02 = g.b0[0] * g.b0[0];
    double lw = (-sg + sqrt(sg * sg - gg * (ss - b02))) / (gg * (im[0] + im[1]));
    xn[0] = xn[0] + (im[0] * lw) * gp;
    xn[1] = xn[1] + (-im[1] * lw) * gp;
  } else {
    shake_newton(g, xn, xp, im, d.shaketol);
  }
  const double ih = 1.0 / d.halfdt;
  for (int k = 0; k < g.na; k++) {
    setp(d.vel, d.n, g.a[k], ih * (xn[k] - xp[k]));
    setp(d.pos, d.n, g.a[k], xn[k] - sh[k]);
  }
}

// theta -> lambda for site s (msld.cu:863-898)
__device__ void theta_to_lambda(const Dev &d, int s) {
  const int b0 = d.site_bound[s], b1 = d.site_bound[s + 1];
  double fix = 0, norm = 0;
  for (int b = b0; b < b1; b++) {
    if (d.block_fixed[b]) fix += d.pos[3 * d.n + b];
    else {
      double l = exp(d.fnex * sin(d.pos[3 * d.n + b]));
      d.lambda[b] = l;
      norm += l;
    }
  }
  const double r = 1 - fix;
  if (norm > 0) norm = r / norm;
  for (int b = b0; b < b1; b++) {
    double l = d.block_fixed[b] ? d.pos[3 * d.n + b] : d.lambda[b] * norm;
    d.lambda[b] = l;
    d.lambdaf[b] = (float)l;
  }
}

// groups, then (last call of a step) sites and the step counter
__global__ void upd_constrain(Dev d, int finish) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (t < ng) constrain_group(d, t);
  if (finish) {
    const int s = t - ng;
    if (s >= 0 && s < d.nsites && d.nl > 1) theta_to_lambda(d, s);
    if (t == ng + d.nsites) *d.step += 1;
  }
}

__global__ void upd_theta_to_lambda(Dev d) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < d.nsites) theta_to_lambda(d, s);
}

// RATTLE velocity projection (holonomic.cu:15-211) for every group
__global__ void upd_vconstrain(Dev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.ntri + d.nbr1 + d.nbr2 + d.nbr3) return;
  int kind;
  Grp g = load_group(d, t, kind);
  D3 xs[4], vs[4];
  double im[4];
  for (int k = 0; k < g.na; k++) {
    xs[k] = getp(d.pos, d.n, g.a[k]);
    vs[k] = getp(d.vel, d.n, g.a[k]);
    im[k] = 1.0 / d.mass[g.a[k]];
  }
  D3 gc[3];
  for (int c = 0; c < g.nc; c++) gc[c] = minimg(d, xs[g.ca[c]] - xs[g.cb[c]]);
  double A[3][3], b[3], lam[3];
  for (int c = 0; c < g.nc; c++) {
    b[c] = -dotd(vs[g.ca[c]] - vs[g.cb[c]], gc[c]);
    for (int e = 0; e < g.nc; e++) {
      double coef = 0;
      if (g.ca[c] == g.ca[e]) coef += im[g.ca[c]];
      if (g.ca[c] == g.cb[e]) coef -= im[g.ca[c]];
      if (g.cb[c] == g.ca[e]) coef -= im[g.cb[c]];
      if (g.cb[c] == g.cb[e]) coef += im[g.cb[c]];
      A[c][e] = coef * dotd(gc[c], gc[e]);
    }
  }
  solve3d(g.nc, A, b, lam);
  for (int e = 0; e < g.nc; e++) {
    vs[g.ca[e]] = vs[g.ca[e]] + (lam[e] * im[g.ca[e]]) * gc[e];
    vs[g.cb[e]] = vs[g.cb[e]] + (-lam[e] * im[g.cb[e]]) * gc[e];
  }
  for (int k = 0; k < g.na; k++) setp(d.vel, d.n, g.a[k], vs[k]);
}

__global__ void upd_kinetic(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < 3 * d.n + d.nl) {
    const double ism = d.ism[i];
    if (ism > 0) {
      const double p = d.vel[i] / ism;
      e = 0.5 * p * p;
    }
  }
  wave_atomic_fix(&d.energy[BM_EKINETIC], (long long)__double2ll_rn(e * kFixScaleD));
}

// ---------------------------------------------------------------------
void launch_update(const Dev &d, bool energy_step, hipStream_t st) {
  const int N = 3 * d.n + d.nl;
  const unsigned gN = (N + 255) / 256;
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  const unsigned gG = (ng + 255) / 256 > 0 ? (ng + 255) / 256 : 1;
  const unsigned gF = (ng + d.nsites + 1 + 255) / 256;
  if (energy_step) {
    upd_kick<<<gN, 256, 0, st>>>(d, 1);
    if (ng) upd_vconstrain<<<gG, 256, 0, st>>>(d);
    upd_kinetic<<<gN, 256, 0, st>>>(d);
    upd_kick<<<gN, 256, 0, st>>>(d, 2);
  } else {
    upd_kick<<<gN, 256, 0, st>>>(d, 0);
  }
  if (ng) upd_constrain<<<gG, 256, 0, st>>>(d, 0);
  upd_o_drift<<<gN, 256, 0, st>>>(d);
  upd_constrain<<<gF, 256, 0, st>>>(d, 1);
}

__global__ void upd_set_step(Dev d, long long v) { *d.step = v; }

void launch_set_step(const Dev &d, long long step, hipStream_t st) {
  upd_set_step<<<1, 1, 0, st>>>(d, step);
}

void launch_theta_to_lambda(const Dev &d, hipStream_t st) {
  upd_theta_to_lambda<<<(d.nsites + 63) / 64, 64, 0, st>>>(d);
}

}  // namespace blade
