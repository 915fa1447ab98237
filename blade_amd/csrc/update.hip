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

// No contraction of separate multiplies and adds into fused ones in this
// file: every FP64 expression rounds as written, so the constraint solves
// and updates give the same bits in every kernel that inlines them (the
// fused step tail and the one-term-at-a-time kernels), whatever the
// surrounding code.  The fused forms that are wanted are explicit.
#pragma clang fp contract(off)

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

// standard normals of DOFs 2p and 2p+1 for (seed, step): the same rule as
// the oracle (one Philox block, Box-Muller cosine for the even DOF, sine for
// the odd one)
__device__ double2 normal_pair(unsigned long long seed, unsigned long long step, unsigned long long p) {
  uint4 c = make_uint4((unsigned)p, (unsigned)step, (unsigned)(step >> 32), 0x4E524D4Cu);
  uint4 o = philox(c, make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  double u1 = ((double)((((unsigned long long)o.x) << 21) ^ (o.y >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)((((unsigned long long)o.z) << 21) ^ (o.w >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  return make_double2(rad * cs, rad * sn);
}

__device__ __forceinline__ void upd_normals_body(const Dev &d, int bid) {
  const int p = bid * blockDim.x + threadIdx.x;
  const int N = 3 * d.n + d.nl;
  if (2 * p >= N) return;
  const double2 z = normal_pair(d.seed, (unsigned long long)(*d.step), (unsigned long long)p);
  d.xi[2 * p] = z.x;
  if (2 * p + 1 < N) d.xi[2 * p + 1] = z.y;
}

__global__ void __launch_bounds__(256) upd_normals(Dev d) { upd_normals_body(d, blockIdx.x); }

__global__ void __launch_bounds__(256) upd_normals_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  upd_normals_body(d, b.bid);
}

// dU/dtheta of block b: the theta biases plus the chain rule from the
// fixed-point dU/dlambda (msld.cu:937-998).  Softmax: lambda_b fnex cos
// theta_b (F_b - sum_c lambda_c F_c / (1 - fixed share)); piecewise:
// g_b (F_b - sum_c g_c F_c / G), g = 3 max(theta - t_s, 0)^2, G = sum g.
// Fixed blocks and the fix flag take no chain-rule force.
__device__ double theta_force(const Dev &d, int b) {
  const double fb = (double)d.dtheta[b] * kFixInvD;
  if (d.block_fixed[b] || d.fixlam) return fb;
  const int s = d.block_site[b];
  if (d.piecewise) {
    const double t0 = d.theta0[s];
    double G = 0, dot = 0;
    for (int c = d.site_bound[s]; c < d.site_bound[s + 1]; c++) {
      const double q = d.pos[3 * d.n + c] - t0, g = q > 0 ? 3 * q * q : 0.0;
      G += g;
      dot += g * ((double)d.dudl[c] * kFixInvD);
    }
    const double q = d.pos[3 * d.n + b] - t0, gb = q > 0 ? 3 * q * q : 0.0;
    return fb + gb * ((double)d.dudl[b] * kFixInvD - dot / G);
  }
  double fix = 0, dyn = 0;
  for (int c = d.site_bound[s]; c < d.site_bound[s + 1]; c++) {
    if (d.block_fixed[c]) fix += d.lambda[c];
    else dyn += d.lambda[c] * ((double)d.dudl[c] * kFixInvD);
  }
  double r = 1 - fix, f = (double)d.dudl[b] * kFixInvD;
  if (r > 0) f -= dyn / r;
  return fb + f * d.lambda[b] * d.fnex * cos(d.pos[3 * d.n + b]);
}


// mode 0: v -= dt f/m, backup, x += dt/2 v; mode 1: v -= dt/2 f/m;
// mode 2: v -= dt/2 f/m, backup, x += dt/2 v
__global__ void upd_kick(Dev d, int mode) {
  if (halted_wg(d)) return;
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
    double v = kick_v(d.vel[i], mode == 0 ? 2.0 : 1.0, d.halfdt, ism, f);
    d.vel[i] = v;
    if (mode != 1) {
      d.pos[i] = drift_x(x, d.halfdt, v);
      nan_check(d, v, i < N1 ? i % d.n : i);
    }
  }
}

// O (friction + Philox noise), backup, half drift R of DOF i
__device__ __forceinline__ void o_drift_dof(const Dev &d, int i) {
  const int N1 = 3 * d.n;
  const double ism = d.ism[i];
  const double x = d.pos[i];
  d.posb[i] = x;
  if (ism > 0) {
    double fr = d.fric, no = d.noise;
    if (i >= N1) {
      fr = d.lfric[i - N1];
      no = d.lnoise[i - N1];
    }
    double v = o_v(fr, d.vel[i], no, ism, d.xi[i]);
    d.vel[i] = v;
    d.pos[i] = drift_x(x, d.halfdt, v);
    nan_check(d, v, i < N1 ? i % d.n : i);
  }
}

__global__ void upd_o_drift(Dev d) {
  if (halted_wg(d)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * d.n + d.nl) o_drift_dof(d, i);
}

// ---- constraints ------------------------------------------------------
// Group shapes are compile-time constants so every per-atom array stays in
// registers.  Constraint c of a group joins atoms ca(c) and cb(c):
// triangles (0,1),(0,2),(1,2); branch groups atom 0 to atom c+1.
template <bool TRI> __device__ constexpr int ca_of(int c) { return TRI ? (c < 2 ? 0 : 1) : 0; }
template <bool TRI> __device__ constexpr int cb_of(int c) { return TRI ? (c < 2 ? c + 1 : 2) : c + 1; }

// Coefficient of lambda_e in constraint c for inverse masses im: the
// overlap of the atoms of constraints c and e with signs.
template <bool TRI>
__device__ __forceinline__ double coupling(int c, int e, const double *im) {
  const int a = ca_of<TRI>(c), b = cb_of<TRI>(c), p = ca_of<TRI>(e), q = cb_of<TRI>(e);
  double coef = 0;
  if (a == p) coef += im[a];
  if (a == q) coef -= im[a];
  if (b == p) coef -= im[b];
  if (b == q) coef += im[b];
  return coef;
}

// Gaussian elimination with partial pivoting (the pivot is the first row
// of largest magnitude, as in the oracle), fully unrolled.
template <int NC>
__device__ __forceinline__ void solve_small(double (&A)[NC][NC], const double (&b)[NC], double (&x)[NC]) {
  double M[NC][NC + 1];
#pragma unroll
  for (int i = 0; i < NC; i++) {
#pragma unroll
    for (int j = 0; j < NC; j++) M[i][j] = A[i][j];
    M[i][NC] = b[i];
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
#pragma unroll
    for (int r = c + 1; r < NC; r++) {
      if (fabs(M[r][c]) > fabs(M[c][c])) {
#pragma unroll
        for (int j = 0; j <= NC; j++) { const double t = M[c][j]; M[c][j] = M[r][j]; M[r][j] = t; }
      }
    }
#pragma unroll
    for (int r = c + 1; r < NC; r++) {
      const double f = M[r][c] / M[c][c];
#pragma unroll
      for (int j = c; j <= NC; j++) M[r][j] -= f * M[c][j];
    }
  }
#pragma unroll
  for (int r = NC - 1; r >= 0; r--) {
    double s = M[r][NC];
#pragma unroll
    for (int j = r + 1; j < NC; j++) s -= M[r][j] * x[j];
    x[r] = s / M[r][r];
  }
}

// Newton-Raphson SHAKE over NC <= 3 constraints (holonomic.cu:430-600)
template <int NA, int NC, bool TRI>
__device__ void shake_newton(const double *b0, D3 (&x)[NA], const D3 (&xp)[NA], const double (&im)[NA], double tol) {
  D3 gp[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) gp[c] = xp[ca_of<TRI>(c)] - xp[cb_of<TRI>(c)];
  for (int it = 0;; it++) {
    double res[NC];
    bool done = true;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const D3 dd = x[ca_of<TRI>(c)] - x[cb_of<TRI>(c)];
      res[c] = b0[c] * b0[c] - dotd(dd, dd);
      if (fabs(res[c]) >= 2 * b0[c] * b0[c] * tol) done = false;
    }
    if (done || it >= 10) break;
    double A[NC][NC], lam[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const D3 dd = x[ca_of<TRI>(c)] - x[cb_of<TRI>(c)];
#pragma unroll
      for (int e = 0; e < NC; e++) A[c][e] = 2 * coupling<TRI>(c, e, im) * dotd(dd, gp[e]);
    }
    solve_small<NC>(A, res, lam);
#pragma unroll
    for (int e = 0; e < NC; e++) {
      x[ca_of<TRI>(e)] = x[ca_of<TRI>(e)] + (lam[e] * im[ca_of<TRI>(e)]) * gp[e];
      x[cb_of<TRI>(e)] = x[cb_of<TRI>(e)] + (-lam[e] * im[cb_of<TRI>(e)]) * gp[e];
    }
  }
}

// analytic SETTLE for an isosceles rigid triangle (the oracle's settle())
__device__ void settle(const double *b0, D3 (&x)[3], const D3 (&xp)[3], const double (&mass)[3]) {
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
#pragma unroll
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
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double fx = qx[i] * cth - qy[i] * sth, fy = qx[i] * sth + qy[i] * cth;
    x[i] = dcm + fx * ex + fy * ey + qz[i] * ez;
  }
}

// Positions of one group: old (backup) positions xp and new ones xn, both
// shifted to the periodic image nearest atom 0 (shift sh).
template <int NA>
struct GroupPos {
  D3 xp[NA], xn[NA], sh[NA];
  double m[NA], im[NA];
};

template <int NA>
__device__ __forceinline__ void load_pos(const Dev &d, const int (&a)[NA], GroupPos<NA> &g) {
  const D3 p0 = getp(d.posb, d.n, a[0]);
#pragma unroll
  for (int k = 0; k < NA; k++) {
    const D3 pk = getp(d.posb, d.n, a[k]);
    g.sh[k] = k ? (minimg(d, pk - p0) - (pk - p0)) : d3(0, 0, 0);
    g.xp[k] = pk + g.sh[k];
    g.xn[k] = getp(d.pos, d.n, a[k]) + g.sh[k];
    g.m[k] = d.mass[a[k]];
    g.im[k] = 1.0 / g.m[k];
  }
}

// The constraint step of one group kind on positions g.  Kinds: 0
// triangle (SETTLE when isosceles, else SHAKE), 1 one bond (analytic),
// 2 and 3 two and three bonds to a centre atom (Newton SHAKE).
template <int KIND, int NA>
__device__ __forceinline__ void solve_kind(const Dev &d, const double *b0, GroupPos<NA> &g) {
  if constexpr (KIND == 0) {
    if (g.m[1] == g.m[2] && b0[0] == b0[1]) settle(b0, g.xn, g.xp, g.m);
    else shake_newton<3, 3, true>(b0, g.xn, g.xp, g.im, d.shaketol);
  } else if constexpr (KIND == 1) {
    const double b02 = b0[0] * b0[0];
    const D3 gp = g.xp[0] - g.xp[1], s = g.xn[0] - g.xn[1];
    const double gg = dotd(gp, gp), sg = dotd(s, gp), ss = dotd(s, s);
    const double lw = (-sg + sqrt(sg * sg - gg * (ss - b02))) / (gg * (g.im[0] + g.im[1]));
    g.xn[0] = g.xn[0] + (g.im[0] * lw) * gp;
    g.xn[1] = g.xn[1] + (-g.im[1] * lw) * gp;
  } else if constexpr (KIND == 2) {
    shake_newton<3, 2, false>(b0, g.xn, g.xp, g.im, d.shaketol);
  } else {
    shake_newton<4, 3, false>(b0, g.xn, g.xp, g.im, d.shaketol);
  }
}

__device__ __forceinline__ D3 shfl3(D3 v, int src) {
  return d3(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64));
}

// One lane (k) of a group's four.  Every lane solves the whole group (the
// same inputs give the same result in each) and lane k finishes atom k:
//   TAIL = false: stores the constrained position and v = dx/(dt/2);
//   TAIL = true:  then runs the O step and half drift of the atom's three
//     DOFs in registers, the four lanes exchange the new positions by
//     shuffles, and every lane solves the group's second constraint step
//     with the first step's result as reference positions.
// The arithmetic is that of the separate kernels (constraint, O + drift,
// constraint), operation for operation.
template <int KIND, int NA, bool TAIL>
__device__ void group_lane(const Dev &d, const int (&a)[NA], const double *b0, int k) {
  GroupPos<NA> g;
  load_pos(d, a, g);
  solve_kind<KIND, NA>(d, b0, g);
  int am = -1;
  D3 xn = d3(0, 0, 0), xp = d3(0, 0, 0), sh = d3(0, 0, 0);
#pragma unroll
  for (int m = 0; m < NA; m++)
    if (m == k) {
      am = a[m];
      xn = g.xn[m];
      xp = g.xp[m];
      sh = g.sh[m];
    }
  const double ih = 1.0 / d.halfdt;
  const D3 v1 = ih * (xn - xp), x1 = xn - sh;
  if (!TAIL) {
    if (am < 0) return;
    setp(d.vel, d.n, am, v1);
    setp(d.pos, d.n, am, x1);
    return;
  }
  // O and half drift (o_drift_dof) of the lane's atom
  double x2[3] = {x1.x, x1.y, x1.z}, vv[3] = {v1.x, v1.y, v1.z};
  if (am >= 0) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int i = c * d.n + am;
      const double ism = d.ism[i];
      if (ism > 0) {
        const double v = o_v(d.fric, vv[c], d.noise, ism, d.xi[i]);
        vv[c] = v;
        x2[c] = drift_x(x2[c], d.halfdt, v);
        nan_check(d, v, am);
      }
    }
  }
  const D3 x2v = d3(x2[0], x2[1], x2[2]);
  const int base = (threadIdx.x & 63) & ~3;
  GroupPos<NA> h;
  D3 X1[NA];
#pragma unroll
  for (int m = 0; m < NA; m++) {
    X1[m] = shfl3(x1, base + m);
    h.xn[m] = shfl3(x2v, base + m);
    h.m[m] = g.m[m];
    h.im[m] = g.im[m];
  }
#pragma unroll
  for (int m = 0; m < NA; m++) {
    h.sh[m] = m ? (minimg(d, X1[m] - X1[0]) - (X1[m] - X1[0])) : d3(0, 0, 0);
    h.xp[m] = X1[m] + h.sh[m];
    h.xn[m] = h.xn[m] + h.sh[m];
  }
  solve_kind<KIND, NA>(d, b0, h);
  if (am < 0) return;
  D3 yn = d3(0, 0, 0), yp = d3(0, 0, 0), ys = d3(0, 0, 0);
#pragma unroll
  for (int m = 0; m < NA; m++)
    if (m == k) {
      yn = h.xn[m];
      yp = h.xp[m];
      ys = h.sh[m];
    }
  setp(d.posb, d.n, am, x1);
  setp(d.vel, d.n, am, ih * (yn - yp));
  setp(d.pos, d.n, am, yn - ys);
}

template <bool TAIL>
__device__ void constrain_group(const Dev &d, int gi, int k) {
  if (gi < d.ntri) {
    const int4 ix = d.tri_i[gi];
    const double4 bb = d.tri_b[gi];
    const int a[3] = {ix.x, ix.y, ix.z};
    const double b0[3] = {bb.x, bb.y, bb.z};
    group_lane<0, 3, TAIL>(d, a, b0, k);
    return;
  }
  gi -= d.ntri;
  if (gi < d.nbr1) {
    const int2 ix = d.br1_i[gi];
    const int a[2] = {ix.x, ix.y};
    const double b0[1] = {d.br1_b[gi]};
    group_lane<1, 2, TAIL>(d, a, b0, k);
    return;
  }
  gi -= d.nbr1;
  if (gi < d.nbr2) {
    const int4 ix = d.br2_i[gi];
    const double2 bb = d.br2_b[gi];
    const int a[3] = {ix.x, ix.y, ix.z};
    const double b0[2] = {bb.x, bb.y};
    group_lane<2, 3, TAIL>(d, a, b0, k);
    return;
  }
  gi -= d.nbr2;
  const int4 ix = d.br3_i[gi];
  const double4 bb = d.br3_b[gi];
  const int a[4] = {ix.x, ix.y, ix.z, ix.w};
  const double b0[3] = {bb.x, bb.y, bb.z};
  group_lane<3, 4, TAIL>(d, a, b0, k);
}

// theta -> lambda for site s (msld.cu:863-898).  Piecewise constraint
// (msld.cu:785-861): the root t_s of sum max(theta - t, 0)^3 = 1 by Newton
// from max(theta) - 1 (at most 51 iterations, stop at a step <= 1e-12),
// lambda = max(theta - t_s, 0)^3.  Fix flag (msld.cu:924-934): theta is set
// to the held lambda, or with init the held lambda is taken from theta.
__device__ void theta_to_lambda(const Dev &d, int s, bool init = false) {
  const int b0 = d.site_bound[s], b1 = d.site_bound[s + 1];
  double *th = d.pos + 3 * (size_t)d.n;
  if (d.fixlam) {
    for (int b = b0; b < b1; b++) {
      if (init) {
        d.lambda[b] = th[b];
        d.lambdaf[b] = (float)th[b];
      } else {
        th[b] = d.lambda[b];
      }
    }
    return;
  }
  if (d.piecewise) {
    double x0 = -1e9;
    for (int b = b0; b < b1; b++) x0 = fmax(x0, th[b]);
    x0 -= 1.0;
    for (int it = 0; it <= 50; it++) {
      double c = -1.0, dc = 0.0;
      for (int b = b0; b < b1; b++) {
        const double q = th[b] - x0;
        if (q > 0) {
          c += q * q * q;
          dc += 3 * q * q;
        }
      }
      const double x1 = x0 + c / dc;
      const bool done = fabs(x1 - x0) <= 1e-12;
      x0 = x1;
      if (done) break;
    }
    d.theta0[s] = x0;
    for (int b = b0; b < b1; b++) {
      const double q = th[b] - x0, l = q > 0 ? q * q * q : 0.0;
      d.lambda[b] = l;
      d.lambdaf[b] = (float)l;
    }
    return;
  }
  double fix = 0, norm = 0;
  for (int b = b0; b < b1; b++) {
    if (d.block_fixed[b]) fix += th[b];
    else {
      double l = exp(d.fnex * sin(th[b]));
      d.lambda[b] = l;
      norm += l;
    }
  }
  const double r = 1 - fix;
  if (norm > 0) norm = r / norm;
  for (int b = b0; b < b1; b++) {
    double l = d.block_fixed[b] ? th[b] : d.lambda[b] * norm;
    d.lambda[b] = l;
    d.lambdaf[b] = (float)l;
  }
}

// full kick of theta DOF b (the atoms' kick ran in the force assembly)
__device__ void theta_kick(const Dev &d, int b) {
  const int i = 3 * d.n + b;
  const double f = theta_force(d, b);
  d.force[i] = fix_d(f);
  const double ism = d.ism[i];
  const double x = d.pos[i];
  d.posb[i] = x;
  if (ism > 0) {
    const double v = kick_v(d.vel[i], 2.0, d.halfdt, ism, f);
    d.vel[i] = v;
    d.pos[i] = drift_x(x, d.halfdt, v);
    nan_check(d, v, i);
  }
}

// groups (four lanes each), then (last call of a step) sites and the step
// counter
__global__ void __launch_bounds__(64) upd_constrain(Dev d, int finish) {
  if (halted_wg(d)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (t < 4 * ng) constrain_group<false>(d, t >> 2, t & 3);
  if (finish) {
    const int s = t - 4 * ng;
    if (s >= 0 && s < d.nsites && d.nl > 1) theta_to_lambda(d, s);
    if (s == d.nsites) *d.step += 1;
  }
}

__global__ void upd_theta_to_lambda(Dev d, bool init) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < d.nsites) theta_to_lambda(d, s, init);
}

// RATTLE velocity projection (holonomic.cu:15-211) of one group
template <int NA, int NC, bool TRI>
__device__ void vproject(const Dev &d, const int (&a)[NA]) {
  D3 xs[NA], vs[NA];
  double im[NA];
#pragma unroll
  for (int k = 0; k < NA; k++) {
    xs[k] = getp(d.pos, d.n, a[k]);
    vs[k] = getp(d.vel, d.n, a[k]);
    im[k] = 1.0 / d.mass[a[k]];
  }
  D3 gc[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) gc[c] = minimg(d, xs[ca_of<TRI>(c)] - xs[cb_of<TRI>(c)]);
  double A[NC][NC], b[NC], lam[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    b[c] = -dotd(vs[ca_of<TRI>(c)] - vs[cb_of<TRI>(c)], gc[c]);
#pragma unroll
    for (int e = 0; e < NC; e++) A[c][e] = coupling<TRI>(c, e, im) * dotd(gc[c], gc[e]);
  }
  solve_small<NC>(A, b, lam);
#pragma unroll
  for (int e = 0; e < NC; e++) {
    vs[ca_of<TRI>(e)] = vs[ca_of<TRI>(e)] + (lam[e] * im[ca_of<TRI>(e)]) * gc[e];
    vs[cb_of<TRI>(e)] = vs[cb_of<TRI>(e)] + (-lam[e] * im[cb_of<TRI>(e)]) * gc[e];
  }
#pragma unroll
  for (int k = 0; k < NA; k++) setp(d.vel, d.n, a[k], vs[k]);
}

__global__ void __launch_bounds__(64) upd_vconstrain(Dev d) {
  if (halted_wg(d)) return;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < d.ntri) {
    const int4 ix = d.tri_i[t];
    const int a[3] = {ix.x, ix.y, ix.z};
    vproject<3, 3, true>(d, a);
    return;
  }
  t -= d.ntri;
  if (t < d.nbr1) {
    const int2 ix = d.br1_i[t];
    const int a[2] = {ix.x, ix.y};
    vproject<2, 1, false>(d, a);
    return;
  }
  t -= d.nbr1;
  if (t < d.nbr2) {
    const int4 ix = d.br2_i[t];
    const int a[3] = {ix.x, ix.y, ix.z};
    vproject<3, 2, false>(d, a);
    return;
  }
  t -= d.nbr2;
  if (t < d.nbr3) {
    const int4 ix = d.br3_i[t];
    const int a[4] = {ix.x, ix.y, ix.z, ix.w};
    vproject<4, 3, false>(d, a);
  }
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
  // constraint groups: 64-thread workgroups spread the few thousand groups
  // over many CUs
  const unsigned gV = (ng + 63) / 64 > 0 ? (ng + 63) / 64 : 1;
  const unsigned gG = (4 * ng + 63) / 64 > 0 ? (4 * ng + 63) / 64 : 1;
  const unsigned gF = (4 * ng + d.nsites + 1 + 63) / 64;
  launch_normals(d, st);
  if (energy_step) {
    upd_kick<<<gN, 256, 0, st>>>(d, 1);
    if (ng) upd_vconstrain<<<gV, 64, 0, st>>>(d);
    upd_kinetic<<<gN, 256, 0, st>>>(d);
    upd_kick<<<gN, 256, 0, st>>>(d, 2);
  } else {
    upd_kick<<<gN, 256, 0, st>>>(d, 0);
  }
  if (ng) upd_constrain<<<gG, 64, 0, st>>>(d, 0);
  upd_o_drift<<<gN, 256, 0, st>>>(d);
  upd_constrain<<<gF, 64, 0, st>>>(d, 1);
}

// The rest of a non-energy step after the force assembly kicked the atoms
// (V and half drift): each constraint group on four lanes (constraint, O
// and half drift, constraint); the unconstrained atoms (O and half drift);
// and, in the last workgroup, the theta DOFs (kick, O, half drift), then
// theta -> lambda and the step counter.
__device__ __forceinline__ void upd_tail_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;
  if (bid == nblk - 1) {
    if (d.nl > 1) {
      for (int b = threadIdx.x; b < d.nl; b += blockDim.x) {
        theta_kick(d, b);
        o_drift_dof(d, 3 * d.n + b);
      }
      __syncthreads();
      for (int s = threadIdx.x; s < d.nsites; s += blockDim.x) theta_to_lambda(d, s);
    }
    __syncthreads();
    if (threadIdx.x == 0) *d.step += 1;
    return;
  }
  const int t = bid * blockDim.x + threadIdx.x;
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (t < 4 * ng) {
    constrain_group<true>(d, t >> 2, t & 3);
    return;
  }
  const int f = t - 4 * ng;
  if (f < d.nfree) {
    const int a = d.freeatom[f];
#pragma unroll
    for (int c = 0; c < 3; c++) o_drift_dof(d, c * d.n + a);
  }
}

__global__ void __launch_bounds__(64) upd_tail(Dev d) { upd_tail_body(d, blockIdx.x, gridDim.x); }

__global__ void __launch_bounds__(64) upd_tail_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  upd_tail_body(d, b.bid, b.nblk);
}

// the groups and free atoms, then one workgroup for the theta DOFs
int grid_tail(const Dev &d) {
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  return (4 * ng + d.nfree + 63) / 64 + 1;
}

int grid_normals(const Dev &d) { return ((3 * d.n + d.nl + 1) / 2 + 255) / 256; }

void launch_update_after_kick(const Dev &d, hipStream_t st) { upd_tail<<<grid_tail(d), 64, 0, st>>>(d); }

void launch_normals(const Dev &d, hipStream_t st) { upd_normals<<<grid_normals(d), 256, 0, st>>>(d); }

void launch_tail_multi(const Multi &m, hipStream_t st) { upd_tail_multi<<<m.nwg, 64, 0, st>>>(m); }

void launch_normals_multi(const Multi &m, hipStream_t st) { upd_normals_multi<<<m.nwg, 256, 0, st>>>(m); }

__global__ void upd_set_step(Dev d, long long v) { *d.step = v; }

void launch_set_step(const Dev &d, long long step, hipStream_t st) {
  upd_set_step<<<1, 1, 0, st>>>(d, step);
}

// ---- MC barostat volume move (update/pressure.cu:63-118) ----------------
// backup, then every atom scaled per axis
__global__ void __launch_bounds__(256) upd_scale_positions(Dev d, double sx, double sy, double sz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  const double s[3] = {sx, sy, sz};
  for (int k = 0; k < 3; k++) {
    const double x = d.pos[k * d.n + i];
    d.posb[k * d.n + i] = x;
    d.pos[k * d.n + i] = x * s[k];
  }
}

// each constraint group's other atoms back at the scaled first atom plus
// their old minimum-image offsets (holonomic/rectify.cu:200-320), so the
// constrained lengths do not change
__device__ __forceinline__ void rectify_group(const Dev &d, const int *idx, int na, D3 oldbox) {
  const D3 b0 = getp(d.posb, d.n, idx[0]), x0 = getp(d.pos, d.n, idx[0]);
  for (int j = 1; j < na; j++) {
    D3 r = d3(d.posb[idx[j]] - b0.x, d.posb[d.n + idx[j]] - b0.y, d.posb[2 * d.n + idx[j]] - b0.z);
    r.x -= oldbox.x * floor(r.x / oldbox.x + 0.5);
    r.y -= oldbox.y * floor(r.y / oldbox.y + 0.5);
    r.z -= oldbox.z * floor(r.z / oldbox.z + 0.5);
    setp(d.pos, d.n, idx[j], d3(x0.x + r.x, x0.y + r.y, x0.z + r.z));
  }
}

__global__ void __launch_bounds__(256) upd_rectify_groups(Dev d, double bx, double by, double bz) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  const D3 ob = d3(bx, by, bz);
  int idx[4];
  if (g < d.ntri) {
    const int4 t = d.tri_i[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z;
    rectify_group(d, idx, 3, ob);
    return;
  }
  g -= d.ntri;
  if (g < d.nbr1) {
    const int2 t = d.br1_i[g];
    idx[0] = t.x; idx[1] = t.y;
    rectify_group(d, idx, 2, ob);
    return;
  }
  g -= d.nbr1;
  if (g < d.nbr2) {
    const int4 t = d.br2_i[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z;
    rectify_group(d, idx, 3, ob);
    return;
  }
  g -= d.nbr2;
  if (g < d.nbr3) {
    const int4 t = d.br3_i[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z; idx[3] = t.w;
    rectify_group(d, idx, 4, ob);
  }
}

// ---- initial constraint rectification (holonomic/rectify.cu:13-197) -----
// Before the first step, every constraint group's atoms other than the
// first are put at exactly their constrained distances from the first atom
// (in its minimum image): a branch atom along its present direction; a
// triangle's second atom along its present direction and its third in the
// plane of the three, on its own side of the 0-1 line.  One thread a group.
__global__ void __launch_bounds__(256) upd_rectify_init(Dev d) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  int idx[4], na;
  double b[3];
  bool tri = false;
  if (g < d.ntri) {
    const int4 t = d.tri_i[g];
    const double4 q = d.tri_b[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z; na = 3;
    b[0] = q.x; b[1] = q.y; b[2] = q.z;
    tri = true;
  } else if ((g -= d.ntri) < d.nbr1) {
    const int2 t = d.br1_i[g];
    idx[0] = t.x; idx[1] = t.y; na = 2;
    b[0] = d.br1_b[g];
  } else if ((g -= d.nbr1) < d.nbr2) {
    const int4 t = d.br2_i[g];
    const double2 q = d.br2_b[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z; na = 3;
    b[0] = q.x; b[1] = q.y;
  } else if ((g -= d.nbr2) < d.nbr3) {
    const int4 t = d.br3_i[g];
    const double4 q = d.br3_b[g];
    idx[0] = t.x; idx[1] = t.y; idx[2] = t.z; idx[3] = t.w; na = 4;
    b[0] = q.x; b[1] = q.y; b[2] = q.z;
  } else {
    return;
  }
  const D3 p0 = getp(d.pos, d.n, idx[0]);
  D3 dv[3], sh[3], q[3];
  for (int k = 1; k < na; k++) {
    const D3 pk = getp(d.pos, d.n, idx[k]);
    dv[k - 1] = minimg(d, pk - p0);
    sh[k - 1] = dv[k - 1] - (pk - p0);
  }
  if (tri) {
    const D3 e0 = (1.0 / normd(dv[0])) * dv[0];
    const D3 c = crossd(dv[0], dv[1]);
    const D3 e2 = (1.0 / normd(c)) * c;
    const D3 e1 = crossd(e2, e0);
    const double along = (b[0] * b[0] + b[1] * b[1] - b[2] * b[2]) / (2.0 * b[0]);
    const double across = sqrt(b[1] * b[1] - along * along);
    q[0] = b[0] * e0;
    q[1] = along * e0 + across * e1;
  } else {
    for (int k = 0; k < na - 1; k++) q[k] = (b[k] / normd(dv[k])) * dv[k];
  }
  for (int k = 1; k < na; k++) setp(d.pos, d.n, idx[k], (p0 + q[k - 1]) - sh[k - 1]);
}

void launch_rectify_init(const Dev &d, hipStream_t st) {
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (ng) upd_rectify_init<<<(ng + 255) / 256, 256, 0, st>>>(d);
}

// ---- virtual sites (holonomic/virtual.cu:11-254) -------------------------
// A site sits at its first host plus an offset o fixed by the hosts (hosts 1
// and 2 taken relative to host 0 in the minimum image).  Two hosts:
// o = (dist/|e| + scale) e with e = x_h0 - x_h1.  Three hosts: ax is the unit
// vector from host 0 to host 1 (to the midpoint of hosts 1 and 2 when
// dist < 0, the bisector form), c the unit vector along (x_h2 - x_h1) x ax,
// bp = ax x c, and o = |dist| (cos t ax + sin t (sin p c + cos p bp)).
struct VirtGeom {
  D3 r1, r2;        // hosts 1, 2 - host 0
  D3 e;             // two hosts: x_h0 - x_h1
  double s, inv;    // two hosts: dist/|e| + scale, 1/|e|
  D3 bb, p1, ax, cc, c, bp;   // three hosts
  double ip1, icc;  // 1/|p1|, 1/|cc|
  D3 o;
};

__device__ __forceinline__ void virt_geom(const Dev &d, int k, int4 &ix, VirtGeom &g) {
  if (k < d.nvirt2) {
    ix = d.virt2_i[k];
    const double2 p = d.virt2_p[k];
    const D3 h0 = getp(d.pos, d.n, ix.y);
    g.r1 = minimg(d, getp(d.pos, d.n, ix.z) - h0);
    g.e = -1.0 * g.r1;
    g.inv = 1.0 / sqrt(dotd(g.e, g.e));
    g.s = p.x * g.inv + p.y;
    g.o = g.s * g.e;
    return;
  }
  const int j = k - d.nvirt2;
  ix = d.virt3_i[j];
  const double4 p = d.virt3_p[2 * j], q = d.virt3_p[2 * j + 1];
  const D3 h0 = getp(d.pos, d.n, ix.y);
  g.r1 = minimg(d, getp(d.pos, d.n, ix.z) - h0);
  g.r2 = minimg(d, getp(d.pos, d.n, ix.w) - h0);
  g.bb = g.r2 - g.r1;
  g.p1 = p.x < 0 ? g.r1 + 0.5 * g.bb : g.r1;
  g.ip1 = 1.0 / sqrt(dotd(g.p1, g.p1));
  g.ax = g.ip1 * g.p1;
  g.cc = crossd(g.bb, g.ax);
  g.icc = 1.0 / sqrt(dotd(g.cc, g.cc));
  g.c = g.icc * g.cc;
  g.bp = crossd(g.ax, g.c);
  const D3 w = q.y * g.c + q.x * g.bp;
  g.o = fabs(p.x) * (p.y * g.ax + p.z * w);
}

__global__ void __launch_bounds__(256) upd_virtual_place(Dev d) {
  if (halted_wg(d)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.nvirt2 + d.nvirt3) return;
  int4 ix;
  VirtGeom g;
  virt_geom(d, k, ix, g);
  // in the periodic image nearest the site's previous position, so the
  // site stays where the pair list saw it
  const D3 nw = getp(d.pos, d.n, ix.y) + g.o;
  const D3 step = nw - getp(d.pos, d.n, ix.x);
  setp(d.pos, d.n, ix.x, nw + (minimg(d, step) - step));
}

// The site's force F goes to the hosts as J^T F (J = d x_site / d x_hosts),
// by reverse-mode differentiation of the construction above; the site's own
// entry is cleared.  Hosts are never sites, so the site entries read here
// are not written by other threads; host sums are fixed-point atomics and
// do not depend on their order.
__global__ void __launch_bounds__(256) upd_virtual_spread(Dev d) {
  if (halted_wg(d)) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.nvirt2 + d.nvirt3) return;
  int4 ix;
  VirtGeom g;
  virt_geom(d, k, ix, g);
  const int n = d.n;
  D3 F = d3((double)d.force[ix.x] * kFixInvD, (double)d.force[n + ix.x] * kFixInvD,
            (double)d.force[2 * n + ix.x] * kFixInvD);
  d.force[ix.x] = 0;
  d.force[n + ix.x] = 0;
  d.force[2 * n + ix.x] = 0;
  D3 f[3];
  int nh;
  if (k < d.nvirt2) {
    // o = s e, s = dist/|e| + scale: d o/d e = s I - dist e e^T/|e|^3
    const double dist = d.virt2_p[k].x;
    const D3 ge = g.s * F - (dist * g.inv * g.inv * g.inv * dotd(g.e, F)) * g.e;
    f[0] = F + ge;   // e = x_h0 - x_h1
    f[1] = -1.0 * ge;
    nh = 2;
  } else {
    const int j = k - d.nvirt2;
    const double4 p = d.virt3_p[2 * j], q = d.virt3_p[2 * j + 1];
    const double D = fabs(p.x);
    // adjoints of o = D cos t ax + D sin t sin p c + D sin t cos p bp
    D3 gax = (D * p.y) * F;
    const D3 gbp = (D * p.z * q.x) * F;
    D3 gc = (D * p.z * q.y) * F;
    gax = gax + crossd(g.c, gbp);   // bp = ax x c
    gc = gc + crossd(gbp, g.ax);
    // c = cc/|cc|
    const D3 gcc = g.icc * (gc - dotd(g.c, gc) * g.c);
    // cc = bb x ax
    D3 gbb = crossd(g.ax, gcc);
    gax = gax + crossd(gcc, g.bb);
    // ax = p1/|p1|, p1 = r1 + h bb
    const D3 gp1 = g.ip1 * (gax - dotd(g.ax, gax) * g.ax);
    if (p.x < 0) gbb = gbb + 0.5 * gp1;
    // bb = r2 - r1; r1, r2 relative to host 0
    f[1] = gp1 - gbb;
    f[2] = gbb;
    f[0] = F - f[1] - f[2];
    nh = 3;
  }
  const int h[3] = {ix.y, ix.z, ix.w};
  for (int a = 0; a < nh; a++) {
    atomic_fix(&d.force[h[a]], fix_d(f[a].x));
    atomic_fix(&d.force[n + h[a]], fix_d(f[a].y));
    atomic_fix(&d.force[2 * n + h[a]], fix_d(f[a].z));
  }
}

void launch_virtual_place(const Dev &d, hipStream_t st) {
  const int nv = d.nvirt2 + d.nvirt3;
  if (nv) upd_virtual_place<<<(nv + 255) / 256, 256, 0, st>>>(d);
}

void launch_virtual_spread(const Dev &d, hipStream_t st) {
  const int nv = d.nvirt2 + d.nvirt3;
  if (nv) upd_virtual_spread<<<(nv + 255) / 256, 256, 0, st>>>(d);
}

// restore the backup (a rejected move)
__global__ void __launch_bounds__(256) upd_restore_positions(Dev d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * d.n) d.pos[i] = d.posb[i];
}

void launch_volume_move(const Dev &d, const double *scale, const double *oldbox, hipStream_t st) {
  upd_scale_positions<<<(d.n + 255) / 256, 256, 0, st>>>(d, scale[0], scale[1], scale[2]);
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (ng) upd_rectify_groups<<<(ng + 255) / 256, 256, 0, st>>>(d, oldbox[0], oldbox[1], oldbox[2]);
}

void launch_volume_restore(const Dev &d, hipStream_t st) {
  upd_restore_positions<<<(3 * d.n + 255) / 256, 256, 0, st>>>(d);
}

// holonomic_position (holonomic/holonomic.cu:600-758): every group from the
// backup positions, v = dx/(dt/2)
void launch_constrain_positions(const Dev &d, hipStream_t st) {
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (ng) upd_constrain<<<(4 * ng + 63) / 64, 64, 0, st>>>(d, 0);
}

// holonomic_velocity (holonomic/holonomic.cu:15-211)
void launch_constrain_velocities(const Dev &d, hipStream_t st) {
  const int ng = d.ntri + d.nbr1 + d.nbr2 + d.nbr3;
  if (ng) upd_vconstrain<<<(ng + 63) / 64, 64, 0, st>>>(d);
}

__global__ void upd_chain_rule(Dev d) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < d.nl) d.force[3 * d.n + b] = fix_d(theta_force(d, b));
}

// calc_thetaForce_from_lambdaForce (msld/msld.cu:937-998): dU/dtheta of every
// block into the theta tail of the force array
void launch_chain_rule(const Dev &d, hipStream_t st) {
  upd_chain_rule<<<(d.nl + 63) / 64, 64, 0, st>>>(d);
}

void launch_theta_to_lambda(const Dev &d, hipStream_t st, bool init) {
  upd_theta_to_lambda<<<(d.nsites + 63) / 64, 64, 0, st>>>(d, init);
}

}  // namespace blade
