// pme.hip - smooth particle-mesh Ewald on one GPU
// (restates nbrecip/nbrecip.cu:19-485 and the grid setup of
// system/potential.cxx:1202-1307; cuFFT is replaced by our own FFT).
//
//   pme_spread   lambda*q B-spline spread into a 64-bit fixed-point grid
//                (order-independent sums: bit-reproducible)
//   fft_lines    mixed-radix Stockham FFT of grid lines staged in LDS;
//                pass z reads the fixed-point grid (and clears it for the
//                next step), pass x fuses forward FFT, the convolution
//                B(m) kE/(2 pi V) exp(-pi^2 m^2/beta^2)/m^2 and the inverse
//   pme_gather   potential and gradient at each atom -> force, dU/dlambda,
//                energy; the Ewald self term rides along
#include "kernels.h"

namespace blade {

// B-spline weights of order N: w[t] = M_N(frac+N-1-t) for grid point u0+t,
// dw[t] its derivative (M_{N-1}(x) - M_{N-1}(x-1)).
template <int N, class T>
__device__ __forceinline__ void bspline_w(T frac, T *w, T *dw) {
  T A[N];
  A[0] = T(1);
#pragma unroll
  for (int k = 1; k < N; k++) A[k] = T(0);
#pragma unroll
  for (int m = 2; m <= N; m++) {
    if (m == N) {
#pragma unroll
      for (int k = 0; k < N; k++) dw[N - 1 - k] = A[k] - (k ? A[k - 1] : T(0));
    }
    const T inv = T(1) / T(m - 1);
#pragma unroll
    for (int k = m - 1; k >= 0; k--) {
      T a = (frac + T(k)) * A[k];
      T b = k ? (T(m) - frac - T(k)) * A[k - 1] : T(0);
      A[k] = (a + b) * inv;
    }
  }
#pragma unroll
  for (int k = 0; k < N; k++) w[N - 1 - k] = A[k];
}

// grid index and fraction from the 32-bit fractional coordinate
template <class T>
__device__ __forceinline__ void grid_coord(unsigned u, int K, int &u0, T &frac) {
  unsigned long long p = (unsigned long long)u * (unsigned)K;
  u0 = (int)(p >> 32);
  frac = (T)((unsigned)(p & 0xFFFFFFFFull)) * (T(1) / T(4294967296.0));
}

template <int N>
__global__ void __launch_bounds__(256) pme_spread(Dev d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / (N * N);
  if (i >= d.n) return;
  const int a = (t / N) % N, c = t % N;
  const int b = d.sb[i] & 0xFFFF;
  const double lq = d.qd[i] * (b ? d.lambda[b] : 1.0);
  if (lq == 0.0) return;
  int u0[3];
  double fr[3];
  for (int k = 0; k < 3; k++) grid_coord(frac_u32(d.pos[k * d.n + i], d.invbox[k]), d.K[k], u0[k], fr[k]);
  // FP64 charge and weights: the grid is the one input of the energy sum
  double wx[N], wy[N], wz[N], dw[N];
  bspline_w<N>(fr[0], wx, dw);
  bspline_w<N>(fr[1], wy, dw);
  bspline_w<N>(fr[2], wz, dw);
  int ix = u0[0] + a; if (ix >= d.K[0]) ix -= d.K[0];
  int iy = u0[1] + c; if (iy >= d.K[1]) iy -= d.K[1];
  const double wac = lq * wx[a] * wy[c];
  long long *row = d.qgrid + ((long long)ix * d.K[1] + iy) * d.K[2];
#pragma unroll
  for (int e = 0; e < N; e++) {
    int iz = u0[2] + e; if (iz >= d.K[2]) iz -= d.K[2];
    atomic_fix(&row[iz], fix_d(wac * wz[e]));
  }
}

// ---------------------------------------------------------------------
// FFT
// ---------------------------------------------------------------------
struct FftPlan {
  int K, nstage;
  int radix[12];
  const float2 *tw;   // e^{-2 pi i m / K}, m < K
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R>
__device__ __forceinline__ void small_dft(float2 *v, float sign) {
  if (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
  } else if (R == 4) {
    float2 a = v[0], b = v[1], c = v[2], e = v[3];
    float2 s0 = make_float2(a.x + c.x, a.y + c.y), d0 = make_float2(a.x - c.x, a.y - c.y);
    float2 s1 = make_float2(b.x + e.x, b.y + e.y), d1 = make_float2(b.x - e.x, b.y - e.y);
    // multiply d1 by (sign * i)
    float2 rd1 = make_float2(-sign * d1.y, sign * d1.x);
    v[0] = make_float2(s0.x + s1.x, s0.y + s1.y);
    v[2] = make_float2(s0.x - s1.x, s0.y - s1.y);
    v[1] = make_float2(d0.x + rd1.x, d0.y + rd1.y);
    v[3] = make_float2(d0.x - rd1.x, d0.y - rd1.y);
  } else {
    float2 in[R];
#pragma unroll
    for (int r = 0; r < R; r++) in[r] = v[r];
#pragma unroll
    for (int k = 0; k < R; k++) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int m = (k * r) % R;
        float cs, sn;
        sincospif(2.0f * m / R, &sn, &cs);
        float2 wv = make_float2(cs, sign * sn);
        float2 p = cmul(in[r], wv);
        acc.x += p.x;
        acc.y += p.y;
      }
      v[k] = acc;
    }
  }
}

// one Stockham stage of radix R over `nl` lines held in LDS
template <int R>
__device__ void stockham_stage(const float2 *src, float2 *dst, int K, int nl, int Ns,
                               const float2 *tw, float sign) {
  const int nb = K / R;
  const int stride = K / (Ns * R);
  for (int idx = threadIdx.x; idx < nl * nb; idx += blockDim.x) {
    const int line = idx / nb, j = idx - line * nb;
    const float2 *in = src + line * K;
    float2 *out = dst + line * K;
    float2 v[R];
    const int jm = j % Ns;
#pragma unroll
    for (int r = 0; r < R; r++) {
      v[r] = in[j + r * nb];
      if (r) {
        // table holds e^{-2 pi i m/K}; the inverse transform conjugates it
        float2 wv = tw[jm * r * stride];
        if (sign > 0) wv.y = -wv.y;
        v[r] = cmul(v[r], wv);
      }
    }
    small_dft<R>(v, sign);
    const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; r++) out[idxD + r * Ns] = v[r];
  }
}

// run all stages; returns the buffer holding the result
__device__ float2 *lds_fft(float2 *b0, float2 *b1, const FftPlan &p, int nl, float sign) {
  float2 *src = b0, *dst = b1;
  int Ns = 1;
  for (int s = 0; s < p.nstage; s++) {
    const int R = p.radix[s];
    if (R == 4) stockham_stage<4>(src, dst, p.K, nl, Ns, p.tw, sign);
    else if (R == 2) stockham_stage<2>(src, dst, p.K, nl, Ns, p.tw, sign);
    else if (R == 3) stockham_stage<3>(src, dst, p.K, nl, Ns, p.tw, sign);
    else stockham_stage<5>(src, dst, p.K, nl, Ns, p.tw, sign);
    __syncthreads();
    float2 *t = src; src = dst; dst = t;
    Ns *= R;
  }
  return src;
}

constexpr int kFftLds = 2048;  // complex elements per LDS buffer

struct FftArgs {
  FftPlan plan;
  int axis;        // 0 = x, 1 = y, 2 = z
  int mode;        // 0 fwd from fixed-point grid, 1 fwd, 2 inverse, 3 fwd+conv+inv
  int energy;      // mode 3: accumulate the reciprocal energy
};

// address of element m of line l along `axis`
__device__ __forceinline__ long long line_addr(const Dev &d, int axis, int l, int m) {
  const int K1 = d.K[1], K2 = d.K[2];
  if (axis == 2) return (long long)l * K2 + m;
  if (axis == 1) {
    const int i = l / K2, k = l - i * K2;
    return ((long long)i * K1 + m) * K2 + k;
  }
  const int j = l / K2, k = l - j * K2;
  return ((long long)m * K1 + j) * K2 + k;
}

__global__ void __launch_bounds__(256) fft_lines(Dev d, FftArgs a) {
  __shared__ float2 buf0[kFftLds], buf1[kFftLds];
  const int K = a.plan.K;
  const int nl = kFftLds / K;
  const int nlines = (d.K[0] * d.K[1] * d.K[2]) / K;
  const int l0 = blockIdx.x * nl;
  const int nmine = min(nl, nlines - l0);
  if (nmine <= 0) return;
  // load: consecutive threads take consecutive lines (coalesced along k)
  for (int e = threadIdx.x; e < nmine * K; e += blockDim.x) {
    const int line = e % nmine, m = e / nmine;
    const long long g = line_addr(d, a.axis, l0 + line, m);
    float2 v;
    if (a.mode == 0) {
      long long q = d.qgrid[g];
      d.qgrid[g] = 0;
      v = make_float2((float)((double)q * kFixInvD), 0.f);
    } else {
      v = d.cgrid[g];
    }
    buf0[line * K + m] = v;
  }
  __syncthreads();
  float2 *res;
  double ek = 0.0;
  if (a.mode == 3) {
    res = lds_fft(buf0, buf1, a.plan, nmine, -1.f);
    // convolution: element m of line (j,k) is wavevector (m, j, k)
    const float L0 = (float)d.box[0], L1 = (float)d.box[1], L2 = (float)d.box[2];
    for (int e = threadIdx.x; e < nmine * K; e += blockDim.x) {
      const int line = e / K, m = e - line * K;
      const int l = l0 + line;
      const int j = l / d.K[2], k = l - j * d.K[2];
      const int mi = (2 * m <= d.K[0]) ? m : m - d.K[0];
      const int mj = (2 * j <= d.K[1]) ? j : j - d.K[1];
      const int mk = (2 * k <= d.K[2]) ? k : k - d.K[2];
      const float kx = mi / L0, ky = mj / L1, kz = mk / L2;
      const float m2 = kx * kx + ky * ky + kz * kz;
      float fac = 0.f;
      if (m || j || k)
        fac = d.bmod[0][m] * d.bmod[1][j] * d.bmod[2][k] * d.pmefac * __expf(-d.pmeexp * m2) / m2;
      float2 v = res[e];
      if (a.energy && (m || j || k)) {
        // E = sum_k G(k) |S(k)|^2 with G in FP64 (equal to the gathered
        // sum_i l q phi(x_i) for the same charge grid)
        const double kxd = mi / d.box[0], kyd = mj / d.box[1], kzd = mk / d.box[2];
        const double m2d = kxd * kxd + kyd * kyd + kzd * kzd;
        const double g = d.bmodd[0][m] * d.bmodd[1][j] * d.bmodd[2][k] * d.pmefacd * exp(-d.pmeexpd * m2d) / m2d;
        ek += g * ((double)v.x * v.x + (double)v.y * v.y);
      }
      res[e] = make_float2(v.x * fac, v.y * fac);
    }
    if (a.energy) wave_atomic_fix(&d.energy[BM_ENBRECIP], (long long)__double2ll_rn(ek * kFixScaleD));
    __syncthreads();
    float2 *other = (res == buf0) ? buf1 : buf0;
    res = lds_fft(res, other, a.plan, nmine, 1.f);
  } else {
    res = lds_fft(buf0, buf1, a.plan, nmine, a.mode == 2 ? 1.f : -1.f);
  }
  for (int e = threadIdx.x; e < nmine * K; e += blockDim.x) {
    const int line = e % nmine, m = e / nmine;
    d.cgrid[line_addr(d, a.axis, l0 + line, m)] = res[line * K + m];
  }
}

template <int N>
__global__ void __launch_bounds__(256) pme_gather(Dev d, bool energy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, eself = 0.0;
  if (i < d.n) {
    const int b = d.sb[i] & 0xFFFF;
    const float q = d.q[i];
    const float l = b ? d.lambdaf[b] : 1.f;
    int u0[3];
    float fr[3];
    for (int k = 0; k < 3; k++) grid_coord(frac_u32(d.pos[k * d.n + i], d.invbox[k]), d.K[k], u0[k], fr[k]);
    float wx[N], wy[N], wz[N], dx[N], dy[N], dz[N];
    bspline_w<N>(fr[0], wx, dx);
    bspline_w<N>(fr[1], wy, dy);
    bspline_w<N>(fr[2], wz, dz);
    float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    for (int a = 0; a < N; a++) {
      int ix = u0[0] + a; if (ix >= d.K[0]) ix -= d.K[0];
      for (int c = 0; c < N; c++) {
        int iy = u0[1] + c; if (iy >= d.K[1]) iy -= d.K[1];
        const float2 *row = d.cgrid + ((long long)ix * d.K[1] + iy) * d.K[2];
        float sp = 0.f, sd = 0.f;
#pragma unroll
        for (int t = 0; t < N; t++) {
          int iz = u0[2] + t; if (iz >= d.K[2]) iz -= d.K[2];
          const float P = row[iz].x;
          sp += P * wz[t];
          sd += P * dz[t];
        }
        phi += wx[a] * wy[c] * sp;
        gx += dx[a] * wy[c] * sp;
        gy += wx[a] * dy[c] * sp;
        gz += wx[a] * wy[c] * sd;
      }
    }
    const float s = 2.f * l * q;
    atomic_fix_f(&d.force[i], s * gx * (d.K[0] * (float)d.invbox[0]));
    atomic_fix_f(&d.force[d.n + i], s * gy * (d.K[1] * (float)d.invbox[1]));
    atomic_fix_f(&d.force[2 * d.n + i], s * gz * (d.K[2] * (float)d.invbox[2]));
    // dU/dlambda_b = 2 q phi (nbrecip.cu:445-449); the energy is summed in
    // k-space by the convolution pass
    if (b) atomic_fix(&d.dudl[b], fix_d(2.0 * d.qd[i] * (double)phi));
  }
  (void)e;
  (void)eself;
  (void)energy;
}

// ---------------------------------------------------------------------
static void make_plan(FftPlan &p, int K, const float2 *tw) {
  p.K = K;
  p.tw = tw;
  p.nstage = 0;
  int n = K;
  while (n % 4 == 0) { p.radix[p.nstage++] = 4; n /= 4; }
  while (n % 2 == 0) { p.radix[p.nstage++] = 2; n /= 2; }
  while (n % 3 == 0) { p.radix[p.nstage++] = 3; n /= 3; }
  while (n % 5 == 0) { p.radix[p.nstage++] = 5; n /= 5; }
}

bool pme_grid_supported(int K) {
  if (K < 2 || K > kFftLds) return false;
  int n = K;
  for (int r : {2, 3, 5})
    while (n % r == 0) n /= r;
  return n == 1;
}

void launch_pme(const Dev &d, bool energy, hipStream_t st) {
  const int N = d.order;
  const long long nt = (long long)d.n * N * N;
  const unsigned gs = (unsigned)((nt + 255) / 256);
  if (N == 4) pme_spread<4><<<gs, 256, 0, st>>>(d);
  else if (N == 6) pme_spread<6><<<gs, 256, 0, st>>>(d);
  else pme_spread<8><<<gs, 256, 0, st>>>(d);
  const int G = d.K[0] * d.K[1] * d.K[2];
  const float2 *tw[3] = {d.twiddle, d.twiddle + d.K[0], d.twiddle + d.K[0] + d.K[1]};
  FftArgs a;
  // forward z (from the fixed-point grid), forward y, x with convolution,
  // inverse y, inverse z
  const int seq_axis[5] = {2, 1, 0, 1, 2};
  const int seq_mode[5] = {0, 1, 3, 2, 2};
  for (int s = 0; s < 5; s++) {
    a.axis = seq_axis[s];
    a.mode = seq_mode[s];
    a.energy = energy && a.mode == 3;
    const int K = d.K[a.axis];
    make_plan(a.plan, K, tw[a.axis]);
    const int nl = kFftLds / K;
    const int nlines = G / K;
    fft_lines<<<(nlines + nl - 1) / nl, 256, 0, st>>>(d, a);
  }
  const unsigned ga = (unsigned)((d.n + 255) / 256);
  if (N == 4) pme_gather<4><<<ga, 256, 0, st>>>(d, energy);
  else if (N == 6) pme_gather<6><<<ga, 256, 0, st>>>(d, energy);
  else pme_gather<8><<<ga, 256, 0, st>>>(d, energy);
}

}  // namespace blade
