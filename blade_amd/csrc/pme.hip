// pme.hip - smooth particle-mesh Ewald on one GPU
// (restates nbrecip/nbrecip.cu:19-485 and the grid setup of
// system/potential.cxx:1202-1307; cuFFT is replaced by our own FFT).
//
//   pme_bin      atoms -> spline records (lambda*q, FP64 B-spline weights,
//                first stencil point) in bins of the grid brick holding
//                their first stencil point
//   pme_spread   one workgroup per brick: lambda*q B-spline weights summed
//                as 64-bit fixed point in LDS (order-independent sums:
//                bit-reproducible), brick stored to the real grid
//   fft_lines    mixed-radix Stockham FFT of grid lines staged in LDS;
//                pass z reads the real grid, pass x fuses forward FFT, the
//                convolution
//                B(m) kE/(2 pi V) exp(-pi^2 m^2/beta^2)/m^2 and the inverse
//   force_assemble  potential gradient at each atom -> force, dU/dlambda;
//                together with the nonbonded and bonded contributions
#include <algorithm>

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

// Brick of grid point u on an axis of K points cut into nb bricks; brick b
// covers [b*K/nb, (b+1)*K/nb).
__device__ __forceinline__ int brick_of(int u, int K, int nb) {
  int b = (u * nb) / K;
  if (b + 1 < nb && ((b + 1) * K) / nb <= u) b++;
  return b;
}

// Bin every charged atom by the brick holding its first stencil point, as
// a spread record: lambda*q, the FP64 B-spline weights of each axis and the
// first stencil point (computed once per atom; every brick the stencil
// reaches reads them), and a 4-byte header holding the first stencil point
// alone (10 bits per axis), so that a brick reads the weights only of the
// atoms whose stencils reach it.  Threads walk the list slots (spatially ordered), so
// a wave's atoms share few bricks and one atomic serves each brick.  Bin
// order is arbitrary; the spread sums in fixed point, so the grid does not
// depend on it.
template <int N>
__device__ __forceinline__ void pme_bin_body(const Dev &d, int bid, int nblk) {
  // a halted step bins nothing (its bins would never be cleared)
  if (halted_wg(d)) return;
  constexpr int W = pme_spline_rec_words(N);
  const int s = bid * blockDim.x + threadIdx.x;
  const int i = s < *d.nblocks * kBlockAtoms ? d.l2g[s] : -1;
  bool act = false;
  int br = 0;
  double lq = 0.0, w[3][N];
  int u0[3] = {0, 0, 0};
  if (i >= 0) {
    const int b = d.sb[i] & 0xFFFF;
    lq = d.qd[i] * (b ? d.lambda[b] : 1.0);
    act = lq != 0.0;
    int bb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double fr, dw[N];
      grid_coord(frac_u32(d.pos[k * d.n + i], d.invbox[k]), d.K[k], u0[k], fr);
      bspline_w<N>(fr, w[k], dw);
      bb[k] = brick_of(u0[k], d.K[k], d.pnb[k]);
    }
    br = (bb[0] * d.pnb[1] + bb[1]) * d.pnb[2] + bb[2];
  }
  const int slot = wave_claim(d.pbcount, br, act);
  if (act) {
    if (slot < d.pcap) {
      double *r = d.pbrec + ((long long)br * d.pcap + slot) * W;
      r[0] = lq;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int t = 0; t < N; t++) r[1 + k * N + t] = w[k][t];
#pragma unroll
      for (int k = 0; k < 3; k++) r[1 + 3 * N + k] = (double)u0[k];
      d.pbhdr[(long long)br * d.pcap + slot] = (unsigned)u0[0] | (unsigned)u0[1] << 10 | (unsigned)u0[2] << 20;
    } else {
      atomicOr(d.errflag, 2);
    }
  }
}

template <int N>
__global__ void __launch_bounds__(256) pme_bin(Dev d) { pme_bin_body<N>(d, blockIdx.x, gridDim.x); }

template <int N>
__global__ void __launch_bounds__(256) pme_bin_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  pme_bin_body<N>(d, b.bid, b.nblk);
}

// candidate atoms staged per round.  The spread runs beside the pair
// kernel, so a smaller workgroup footprint (about 30 KB of LDS for one
// system, 15 KB in the batched launch, which has many more bricks) lets
// more bricks share CUs with it: measured over staging sizes 32-256, these
// gave the fastest steps (the spread alone is a little slower)
template <int N> constexpr int spread_stage() { return N <= 6 ? 128 : 64; }
template <int N> constexpr int spread_stage_multi() { return N <= 6 ? 64 : 32; }

// v * 2^32 rounded to the nearest integer (ties to even), as
// __double2ll_rn(v * 2^32) gives, for |v * 2^32| < 2^51: the sum with
// 1.5 * 2^52 is rounded to an integer by the addition itself, and its low
// mantissa bits are that integer plus a constant.
__device__ __forceinline__ long long fix_d_small(double v) {
  const double magic = 6755399441055744.0;   // 1.5 * 2^52
  return __double_as_longlong(fma(v, 4294967296.0, magic)) - __double_as_longlong(magic);
}

// Exclusive prefix sum of v over a 256-thread workgroup; the sum of all
// values goes to total.  Every thread must call it.
__device__ __forceinline__ int block_scan256(int v, int *s_wsum, int &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int t = s_wsum[k];
    if (k < w) pre += t;
    tot += t;
  }
  total = tot;
  return pre + x - v;
}

// The part of a stencil of N points starting at r (relative to the brick,
// in [0, K)) that falls inside a brick of sz points: stencil points
// t0 .. t0+n-1 at brick positions p0 .. p0+n-1 (mod K when the brick is the
// whole axis), packed as t0 | n << 4 | p0 << 8.
template <int N>
__device__ __forceinline__ int stencil_clip(int r, int sz, int K) {
  int t0 = 0, n = N, p0 = r;
  if (sz < K) {
    if (r < sz) {
      n = min(N, sz - r);
    } else {          // starts in the previous brick and wraps into this one
      t0 = K - r;
      if (t0 >= N) return 0;
      n = N - t0;
      p0 = 0;
    }
  }
  return t0 | n << 4 | p0 << 8;
}

// One workgroup (256 threads) per brick.  Every brick spans >= order-1
// points (or a whole axis), so only atoms binned in this brick or the
// previous one along each axis can reach it.  Each round takes up to
// spread_stage<N>() of those atoms, clips each stencil to the brick from
// the record header alone, and stages the spline records (FP64 lambda*q,
// B-spline weights) of the atoms whose stencils reach the brick, with the
// part of each stencil inside it; it lists the (atom, x, y) stencil rows
// that cross the brick and gives each thread one row at a time; its z
// points are added as 64-bit fixed point in LDS.  The brick is
// stored to the real grid with plain stores: no global atomics, and no
// clearing pass.
template <int N, int kStage>
__device__ __forceinline__ void pme_spread_body(const Dev &d, int bid, int nblk) {
  extern __shared__ long long acc[];
  __shared__ double s_w[3][N][kStage];
  __shared__ int s_clip[3][kStage];
  __shared__ double s_lq[kStage];
  __shared__ unsigned short s_row[kStage * N * N];   // atom | x << 8 | y << 11
  __shared__ int s_cnt[9];
  __shared__ int s_bin[8];
  __shared__ int s_wsum[4];
  int bidx[3], lo[3], sz[3];
  {
    int r = bid;
    bidx[2] = r % d.pnb[2]; r /= d.pnb[2];
    bidx[1] = r % d.pnb[1];
    bidx[0] = r / d.pnb[1];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = (bidx[k] * d.K[k]) / d.pnb[k];
    sz[k] = ((bidx[k] + 1) * d.K[k]) / d.pnb[k] - lo[k];
  }
  // padded row of an odd number of 64-bit words: lanes on consecutive rows
  // fall in different LDS banks
  const int rowz = sz[2] + 1 + (sz[2] & 1);
  const int npts = sz[0] * sz[1] * rowz;
  for (int p = threadIdx.x; p < npts; p += blockDim.x) acc[p] = 0;
  // feeding bins (this brick and the previous one along each axis, each
  // once) and their atom counts
  if (threadIdx.x < 8) {
    const int f = threadIdx.x;
    int fb[3];
    bool dup = false;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int off = (f >> k) & 1;
      if (off && d.pnb[k] == 1) dup = true;
      fb[k] = bidx[k] - off < 0 ? d.pnb[k] - 1 : bidx[k] - off;
    }
    const int br = (fb[0] * d.pnb[1] + fb[1]) * d.pnb[2] + fb[2];
    s_bin[f] = br;
    s_cnt[f] = dup ? 0 : min(d.pbcount[br], d.pcap);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int f = 0; f < 8; f++) {
      const int c = s_cnt[f];
      s_cnt[f] = run;
      run += c;
    }
    s_cnt[8] = run;
  }
  __syncthreads();
  const int total = s_cnt[8];
  for (int base = 0; base < total; base += kStage) {
    const int na = min(kStage, total - base);
    int clip[3] = {0, 0, 0};
    long long entry = 0;
    bool reach = false;
    if ((int)threadIdx.x < na) {
      const int g = base + threadIdx.x;
      int f = 0;
      while (f < 7 && s_cnt[f + 1] <= g) f++;
      entry = (long long)s_bin[f] * d.pcap + (g - s_cnt[f]);
      const unsigned h = d.pbhdr[entry];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        // stencil start relative to the brick, in [0, K)
        int r = (int)((h >> (10 * k)) & 1023u) - lo[k];
        if (r < 0) r += d.K[k];
        clip[k] = stencil_clip<N>(r, sz[k], d.K[k]);
      }
      reach = ((clip[0] >> 4) & 15) && ((clip[1] >> 4) & 15) && ((clip[2] >> 4) & 15);
    }
    // the atoms whose stencils reach the brick, compacted
    int nreach;
    const int a = block_scan256(reach ? 1 : 0, s_wsum, nreach);
    int nrow = 0;
    if (reach) {
      const double *rec = d.pbrec + entry * pme_spline_rec_words(N);
      s_lq[a] = rec[0];
#pragma unroll
      for (int k = 0; k < 3; k++) {
#pragma unroll
        for (int t = 0; t < N; t++) s_w[k][t][a] = rec[1 + k * N + t];
        s_clip[k][a] = clip[k];
      }
      nrow = ((clip[0] >> 4) & 15) * ((clip[1] >> 4) & 15);
    }
    __syncthreads();   // s_wsum is reused by the second scan
    int nrows;
    const int off = block_scan256(nrow, s_wsum, nrows);
    if (nrow) {
      const int nx = (clip[0] >> 4) & 15, ny = (clip[1] >> 4) & 15;
      int q = off;
      for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) s_row[q++] = (unsigned short)(a | x << 8 | y << 11);
    }
    __syncthreads();
    for (int it = threadIdx.x; it < nrows; it += blockDim.x) {
      const unsigned e = s_row[it];
      const int a = e & 255, x = (e >> 8) & 7, y = (e >> 11) & 7;
      const int cx = s_clip[0][a], cy = s_clip[1][a], cz = s_clip[2][a];
      int rx = (cx >> 8) + x;
      if (rx >= d.K[0]) rx -= d.K[0];
      int ry = (cy >> 8) + y;
      if (ry >= d.K[1]) ry -= d.K[1];
      const double wac = s_lq[a] * s_w[0][(cx & 15) + x][a] * s_w[1][(cy & 15) + y][a];
      long long *row = acc + (rx * sz[1] + ry) * rowz;
      const int tz = cz & 15, nz = (cz >> 4) & 15, pz0 = cz >> 8;
#pragma unroll
      for (int z = 0; z < N; z++) {
        if (z < nz) {
          int rz = pz0 + z;
          if (rz >= d.K[2]) rz -= d.K[2];
          // |lambda q w w w| <= |q| < 2^19: within fix_d_small's range
          atomicAdd(reinterpret_cast<unsigned long long *>(row + rz),
                    (unsigned long long)fix_d_small(wac * s_w[2][tz + z][a]));
        }
      }
    }
    __syncthreads();
  }
  for (int p = threadIdx.x; p < sz[0] * sz[1] * sz[2]; p += blockDim.x) {
    const int pz = p % sz[2], pxy = p / sz[2];
    const int py = pxy % sz[1], px = pxy / sz[1];
    d.rgrid[((long long)(lo[0] + px) * d.K[1] + lo[1] + py) * d.K[2] + lo[2] + pz] =
        (float)((double)acc[pxy * rowz + pz] * kFixInvD);
  }
}

template <int N>
__global__ void __launch_bounds__(256) pme_spread(Dev d) {
  pme_spread_body<N, spread_stage<N>()>(d, blockIdx.x, gridDim.x);
}

template <int N>
__global__ void __launch_bounds__(256) pme_spread_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  pme_spread_body<N, spread_stage_multi<N>()>(d, b.bid, b.nblk);
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
  int mode;        // 0 fwd from the real grid, 1 fwd, 2 inverse, 3 fwd+conv+inv,
                   // 4 inverse with the real part stored to the real grid
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
      v = make_float2(d.rgrid[g], 0.f);
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
    res = lds_fft(buf0, buf1, a.plan, nmine, a.mode >= 2 ? 1.f : -1.f);
  }
  for (int e = threadIdx.x; e < nmine * K; e += blockDim.x) {
    const int line = e % nmine, m = e / nmine;
    const long long g = line_addr(d, a.axis, l0 + line, m);
    if (a.mode == 4) d.rgrid[g] = res[line * K + m].x;
    else d.cgrid[g] = res[line * K + m];
  }
}

// ---------------------------------------------------------------------
// Compile-time FFT: transform size, radix sequence, line count and LDS
// strides are template constants, so every index is shift/multiply
// arithmetic.  Stockham stages ping-pong between two LDS buffers.
// ---------------------------------------------------------------------
constexpr float kC3 = -0.5f, kS3 = 0.86602540378443865f;                // 2 pi/3
constexpr float kC51 = 0.30901699437494742f, kS51 = 0.95105651629515357f;  // 2 pi/5
constexpr float kC52 = -0.80901699437494742f, kS52 = 0.58778525229247313f; // 4 pi/5
constexpr float kR2 = 0.70710678118654752f;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (s i) for s = +-1
__device__ __forceinline__ float2 cmuli(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// length-R DFT in registers, X_k = sum_r v_r exp(sign 2 pi i r k / R)
template <int R>
__device__ __forceinline__ void ct_dft(float2 (&v)[R], float sign) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const float2 s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    const float2 s1 = cadd(v[1], v[3]), d1 = cmuli(csub(v[1], v[3]), sign);
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
  } else if constexpr (R == 8) {
    float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    ct_dft<4>(e, sign);
    ct_dft<4>(o, sign);
    // o_k * exp(sign 2 pi i k/8)
    const float2 w1 = make_float2(kR2 * (o[1].x - sign * o[1].y), kR2 * (o[1].y + sign * o[1].x));
    const float2 w2 = cmuli(o[2], sign);
    const float2 w3 = make_float2(-kR2 * (o[3].x + sign * o[3].y), kR2 * (sign * o[3].x - o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  } else if constexpr (R == 3) {
    const float2 s = cadd(v[1], v[2]), dd = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x + kC3 * s.x, v[0].y + kC3 * s.y);
    const float2 r = cmuli(dd, sign * kS3);
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else {
    static_assert(R == 5, "radix 2, 3, 4, 5 or 8");
    const float2 s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    const float2 s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    const float2 m1 = make_float2(v[0].x + kC51 * s1.x + kC52 * s2.x, v[0].y + kC51 * s1.y + kC52 * s2.y);
    const float2 m2 = make_float2(v[0].x + kC52 * s1.x + kC51 * s2.x, v[0].y + kC52 * s1.y + kC51 * s2.y);
    const float2 r1 = make_float2(kS51 * d1.x + kS52 * d2.x, kS51 * d1.y + kS52 * d2.y);
    const float2 r2 = make_float2(kS52 * d1.x - kS51 * d2.x, kS52 * d1.y - kS51 * d2.y);
    const float2 q1 = cmuli(r1, sign), q2 = cmuli(r2, sign);
    v[0] = make_float2(v[0].x + s1.x + s2.x, v[0].y + s1.y + s2.y);
    v[1] = cadd(m1, q1);
    v[4] = csub(m1, q1);
    v[2] = cadd(m2, q2);
    v[3] = csub(m2, q2);
  }
}

constexpr int ct_radix(int n) {
  return n % 8 == 0 ? 8 : n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : 5;
}
constexpr bool ct_ok(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// One Stockham stage over NL lines of K points; element e of line l at
// buf[l*LS + e*ES].  LINEFAST: consecutive threads take consecutive lines
// (for ES > 1), otherwise consecutive butterflies of one line.
template <int K, int R, int Ns, int NL, int ES, int LS, bool LINEFAST>
__device__ __forceinline__ void ct_stage(const float2 *src, float2 *dst, const float2 *tw, float sign) {
  constexpr int NB = K / R, TS = K / (Ns * R);
  for (int b = threadIdx.x; b < NL * NB; b += blockDim.x) {
    const int line = LINEFAST ? b % NL : b / NB;
    const int j = LINEFAST ? b / NL : b % NB;
    const float2 *in = src + line * LS;
    float2 *out = dst + line * LS;
    const int jm = j % Ns;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      v[r] = in[(j + r * NB) * ES];
      if (Ns > 1 && r) {
        // table: exp(-2 pi i m/K); the inverse transform conjugates it
        float2 w = tw[jm * r * TS];
        w.y *= -sign;
        v[r] = make_float2(v[r].x * w.x - v[r].y * w.y, v[r].x * w.y + v[r].y * w.x);
      }
    }
    ct_dft<R>(v, sign);
    const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; r++) out[(idxD + r * Ns) * ES] = v[r];
  }
}

// all stages; returns the buffer holding the result
template <int K, int Ns, int Rem, int NL, int ES, int LS, bool LINEFAST>
__device__ __forceinline__ float2 *ct_fft(float2 *src, float2 *dst, const float2 *tw, float sign) {
  if constexpr (Rem == 1) {
    return src;
  } else {
    constexpr int R = ct_radix(Rem);
    ct_stage<K, R, Ns, NL, ES, LS, LINEFAST>(src, dst, tw, sign);
    __syncthreads();
    return ct_fft<K, Ns * R, Rem / R, NL, ES, LS, LINEFAST>(dst, src, tw, sign);
  }
}

// Forward transform along z then y of one x-plane (K1 x K2, in LDS): reads
// the real charge grid, writes the complex grid.  1024 threads: there are
// only K0 planes, so each plane's workgroup carries 16 waves.
template <int K1, int K2>
__device__ __forceinline__ void fft_yz_fwd_body(const Dev &d, int bid, int nblk) {
  __shared__ float2 b0[K1 * K2], b1[K1 * K2];
  __shared__ float2 t1[K1], t2[K2];
  const long long base = (long long)bid * K1 * K2;
  for (int m = threadIdx.x; m < K1; m += blockDim.x) t1[m] = d.twiddle[d.K[0] + m];
  for (int m = threadIdx.x; m < K2; m += blockDim.x) t2[m] = d.twiddle[d.K[0] + K1 + m];
  for (int e = threadIdx.x; e < K1 * K2; e += blockDim.x) b0[e] = make_float2(d.rgrid[base + e], 0.f);
  __syncthreads();
  float2 *r = ct_fft<K2, 1, K2, K1, 1, K2, false>(b0, b1, t2, -1.f);
  float2 *o = r == b0 ? b1 : b0;
  r = ct_fft<K1, 1, K1, K2, K2, 1, true>(r, o, t1, -1.f);
  for (int e = threadIdx.x; e < K1 * K2; e += blockDim.x) d.cgrid[base + e] = r[e];
}

template <int K1, int K2>
__global__ void __launch_bounds__(1024) fft_yz_fwd(Dev d) { fft_yz_fwd_body<K1, K2>(d, blockIdx.x, gridDim.x); }

template <int K1, int K2>
__global__ void __launch_bounds__(1024) fft_yz_fwd_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  fft_yz_fwd_body<K1, K2>(d, b.bid, b.nblk);
}

// Inverse transform along y then z of one x-plane; the real part (the
// potential on the grid) goes to the real grid.
template <int K1, int K2>
__device__ __forceinline__ void fft_yz_inv_body(const Dev &d, int bid, int nblk) {
  __shared__ float2 b0[K1 * K2], b1[K1 * K2];
  __shared__ float2 t1[K1], t2[K2];
  const long long base = (long long)bid * K1 * K2;
  for (int m = threadIdx.x; m < K1; m += blockDim.x) t1[m] = d.twiddle[d.K[0] + m];
  for (int m = threadIdx.x; m < K2; m += blockDim.x) t2[m] = d.twiddle[d.K[0] + K1 + m];
  for (int e = threadIdx.x; e < K1 * K2; e += blockDim.x) b0[e] = d.cgrid[base + e];
  __syncthreads();
  float2 *r = ct_fft<K1, 1, K1, K2, K2, 1, true>(b0, b1, t1, 1.f);
  float2 *o = r == b0 ? b1 : b0;
  r = ct_fft<K2, 1, K2, K1, 1, K2, false>(r, o, t2, 1.f);
  for (int e = threadIdx.x; e < K1 * K2; e += blockDim.x) d.rgrid[base + e] = r[e].x;
}

template <int K1, int K2>
__global__ void __launch_bounds__(1024) fft_yz_inv(Dev d) { fft_yz_inv_body<K1, K2>(d, blockIdx.x, gridDim.x); }

template <int K1, int K2>
__global__ void __launch_bounds__(1024) fft_yz_inv_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  fft_yz_inv_body<K1, K2>(d, b.bid, b.nblk);
}

// Lines along x (NL per workgroup): forward transform, the convolution
// B(m) kE/(2 pi V) exp(-pi^2 m^2/beta^2)/m^2 (with the reciprocal energy
// sum on energy steps), inverse transform.
template <int K0, int NL>
__device__ __forceinline__ void fft_x_conv_body(const Dev &d, int bid, int nblk, int energy) {
  __shared__ float2 b0[K0 * NL], b1[K0 * NL];
  __shared__ float2 t0[K0];
  const int nlines = d.K[1] * d.K[2];
  const int l0 = bid * NL;
  const long long plane = (long long)d.K[1] * d.K[2];
  for (int m = threadIdx.x; m < K0; m += blockDim.x) t0[m] = d.twiddle[m];
  // element m of line l at b[m*NL + l]: consecutive threads, consecutive lines
  for (int e = threadIdx.x; e < K0 * NL; e += blockDim.x) {
    const int l = e % NL, m = e / NL;
    b0[e] = l0 + l < nlines ? d.cgrid[m * plane + l0 + l] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2 *r = ct_fft<K0, 1, K0, NL, NL, 1, true>(b0, b1, t0, -1.f);
  const float L0 = (float)d.box[0], L1 = (float)d.box[1], L2 = (float)d.box[2];
  double ek = 0.0;
  for (int e = threadIdx.x; e < K0 * NL; e += blockDim.x) {
    const int l = e % NL, m = e / NL;
    const int line = l0 + l;
    if (line >= nlines) continue;
    const int j = line / d.K[2], k = line - j * d.K[2];
    const int mi = (2 * m <= K0) ? m : m - K0;
    const int mj = (2 * j <= d.K[1]) ? j : j - d.K[1];
    const int mk = (2 * k <= d.K[2]) ? k : k - d.K[2];
    const float kx = mi / L0, ky = mj / L1, kz = mk / L2;
    const float m2 = kx * kx + ky * ky + kz * kz;
    float fac = 0.f;
    if (m || j || k) fac = d.bmod[0][m] * d.bmod[1][j] * d.bmod[2][k] * d.pmefac * __expf(-d.pmeexp * m2) / m2;
    const float2 v = r[e];
    if (energy && (m || j || k)) {
      const double kxd = mi / d.box[0], kyd = mj / d.box[1], kzd = mk / d.box[2];
      const double m2d = kxd * kxd + kyd * kyd + kzd * kzd;
      const double g = d.bmodd[0][m] * d.bmodd[1][j] * d.bmodd[2][k] * d.pmefacd * exp(-d.pmeexpd * m2d) / m2d;
      ek += g * ((double)v.x * v.x + (double)v.y * v.y);
    }
    r[e] = make_float2(v.x * fac, v.y * fac);
  }
  if (energy) wave_atomic_fix(&d.energy[BM_ENBRECIP], (long long)__double2ll_rn(ek * kFixScaleD));
  __syncthreads();
  float2 *o = r == b0 ? b1 : b0;
  r = ct_fft<K0, 1, K0, NL, NL, 1, true>(r, o, t0, 1.f);
  for (int e = threadIdx.x; e < K0 * NL; e += blockDim.x) {
    const int l = e % NL, m = e / NL;
    if (l0 + l < nlines) d.cgrid[m * plane + l0 + l] = r[e];
  }
}

template <int K0, int NL>
__global__ void __launch_bounds__(256) fft_x_conv(Dev d, int energy) { fft_x_conv_body<K0, NL>(d, blockIdx.x, gridDim.x, energy); }

template <int K0, int NL>
__global__ void __launch_bounds__(256) fft_x_conv_multi(Multi m, int energy) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  fft_x_conv_body<K0, NL>(d, b.bid, b.nblk, energy);
}

// sizes with compiled plane / line kernels
#define BM_FFT_SIZES(X) X(16) X(18) X(20) X(24) X(27) X(32) X(36) X(40) X(48) X(54) X(64)

// Per-atom force assembly, eight lanes per list slot (slots in spatial
// order, so the PME stencils of a wave overlap in cache).  Adds, in fixed
// point, the components selected by mask: the PME force 2 lambda q
// grad(phi) and dU/dlambda = 2 q phi from the potential grid (N = spline
// order; N = 0 compiles no PME part), the nonbonded chunk partials of the
// atom's block, and the atom's bonded slots.  The lanes split the stencil,
// the chunks and the slots; integer sums across them are exact.  One lane
// per force component and stream order make the read-modify-write of the
// force array safe without atomics.
template <int N, bool KICK>
__device__ __forceinline__ void force_assemble_body(const Dev &d, int bid, int nblk, unsigned mask) {
  if (KICK && halted_wg(d)) return;
  // eight lanes per list slot; every lane of a group takes the same branches
  const int g = bid * blockDim.x + threadIdx.x;
  const int s = g >> 3, sub = g & 7;
  const bool live = s < *d.nblocks * kBlockAtoms;
  const int i = live ? d.l2g[s] : -1;
  // loads that depend only on the slot and the atom come first; every
  // store and atomic is left to the end, so none of them orders the loads
  const int j = sub * d.n + i;
  long long F0 = 0;
  double x0 = 0.0, v0 = 0.0, ism = 0.0;
  if (i >= 0 && sub < 3) {
    F0 = d.force[j];
    if (KICK) {
      ism = d.ism[i];
      x0 = d.pos[j];
      v0 = d.vel[j];
    }
  }
  long long fx = 0, fy = 0, fz = 0;
  long long dl = 0;
  int bl = 0;
  if constexpr (N > 0) {
   if ((mask & kAsmPme) && i >= 0) {
    // grid coordinates from the slot's packed fractional position; lane sub
    // takes stencil offset z = u0 + sub of every (x, y) row, so the eight
    // lanes of an atom read one short stretch of a grid row together
    const uint4 ps = d.spos[s];
    const int b = d.spar[s].y & 0xFFFF;
    int u0[3];
    float fr[3];
    grid_coord(ps.x, d.K[0], u0[0], fr[0]);
    grid_coord(ps.y, d.K[1], u0[1], fr[1]);
    grid_coord(ps.z, d.K[2], u0[2], fr[2]);
    float wx[N], wy[N], wz[N], dx[N], dy[N], dz[N];
    bspline_w<N>(fr[0], wx, dx);
    bspline_w<N>(fr[1], wy, dy);
    bspline_w<N>(fr[2], wz, dz);
    float wt = 0.f, dt = 0.f;
#pragma unroll
    for (int t = 0; t < N; t++)
      if (t == sub) {
        wt = wz[t];
        dt = dz[t];
      }
    float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    if (sub < N) {
      // Q = sum_xy wx wy P, Gx = sum_xy dx wy P, Gy = sum_xy wx dy P
      float Q = 0.f, Gx = 0.f, Gy = 0.f;
      int iz = u0[2] + sub;
      if (iz >= d.K[2]) iz -= d.K[2];
#pragma unroll
      for (int a = 0; a < N; a++) {
        int ix = u0[0] + a;
        if (ix >= d.K[0]) ix -= d.K[0];
        const float *plane = d.rgrid + (long long)ix * d.K[1] * d.K[2] + iz;
        float sy = 0.f, syd = 0.f;
#pragma unroll
        for (int c = 0; c < N; c++) {
          int iy = u0[1] + c;
          if (iy >= d.K[1]) iy -= d.K[1];
          const float P = plane[(long long)iy * d.K[2]];
          sy += wy[c] * P;
          syd += dy[c] * P;
        }
        Q += wx[a] * sy;
        Gx += dx[a] * sy;
        Gy += wx[a] * syd;
      }
      phi = wt * Q;
      gx = wt * Gx;
      gy = wt * Gy;
      gz = dt * Q;
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
      phi += __shfl_xor(phi, o, 64);
      gx += __shfl_xor(gx, o, 64);
      gy += __shfl_xor(gy, o, 64);
      gz += __shfl_xor(gz, o, 64);
    }
    if (sub == 0) {
      const float sc = 2.f * (b ? d.lambdaf[b] : 1.f) * __uint_as_float(ps.w);
      fx = fix_f(sc * gx * (d.K[0] * (float)d.invbox[0]));
      fy = fix_f(sc * gy * (d.K[1] * (float)d.invbox[1]));
      fz = fix_f(sc * gz * (d.K[2] * (float)d.invbox[2]));
      // dU/dlambda_b = 2 q phi (nbrecip.cu:445-449); the energy is summed
      // in k-space by the convolution pass
      if (b) {
        bl = b;
        dl = fix_d(2.0 * d.qd[i] * (double)phi);
      }
    }
   }
  }
  if ((mask & kAsmNb) && i >= 0 && *d.ntiles <= d.maxtiles && *d.nchunks <= d.maxchunks) {
    const int I = s / kBlockAtoms, li = s % kBlockAtoms;
    const int c1 = d.cstart[I + 1];
    for (int c = d.cstart[I] + sub; c < c1; c += 8) {
      const float *src = d.ichunk + (long long)c * 3 * kBlockAtoms;
      fx += fix_f(src[li]);
      fy += fix_f(src[kBlockAtoms + li]);
      fz += fix_f(src[2 * kBlockAtoms + li]);
    }
  }
  if ((mask & kAsmBonded) && i >= 0) {
    const int k1 = d.bcsr_start[i + 1];
    for (int k = d.bcsr_start[i] + sub; k < k1; k += 8) {
      fx += d.bslot[k];
      fy += d.bslot[d.nslot + k];
      fz += d.bslot[2 * d.nslot + k];
    }
  }
  // integer sums: the same total in any order
#pragma unroll
  for (int o = 1; o < 8; o <<= 1) {
    fx += __shfl_xor(fx, o, 64);
    fy += __shfl_xor(fy, o, 64);
    fz += __shfl_xor(fz, o, 64);
  }
  // the spread has consumed the PME bins: clear them for the next step
  if (N && (mask & kAsmPme) && g < d.pnb[0] * d.pnb[1] * d.pnb[2]) d.pbcount[g] = 0;
  if (bl) atomic_fix(&d.dudl[bl], dl);
  if (i < 0 || sub >= 3) return;
  // lanes 0, 1, 2 finish the x, y, z components
  const long long F = F0 + (sub == 0 ? fx : sub == 1 ? fy : fz);
  d.force[j] = F;
  if (KICK) {
    // V (full kick), backup, half drift R: as upd_kick mode 0
    const double f = (double)F * kFixInvD;
    d.posb[j] = x0;
    if (ism > 0) {
      const double v = kick_v(v0, 2.0, d.halfdt, ism, f);
      d.vel[j] = v;
      d.pos[j] = drift_x(x0, d.halfdt, v);
      nan_check(d, v, i);
    }
  }
}

template <int N, bool KICK>
__global__ void __launch_bounds__(256) force_assemble(Dev d, unsigned mask) { force_assemble_body<N, KICK>(d, blockIdx.x, gridDim.x, mask); }

template <int N, bool KICK>
__global__ void __launch_bounds__(256) force_assemble_multi(Multi m, unsigned mask) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  force_assemble_body<N, KICK>(d, b.bid, b.nblk, mask);
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

// (at most 1024 points per axis: the spread record header keeps 10 bits of
// the first stencil point per axis)
bool pme_grid_supported(int K) {
  if (K < 2 || K > kFftLds || K > 1024) return false;
  int n = K;
  for (int r : {2, 3, 5})
    while (n % r == 0) n /= r;
  return n == 1;
}

int grid_pme_bin(const Dev &d) { return (int)((32LL * d.maxblocks + 255) / 256); }

int grid_pme_spread(const Dev &d) { return d.pnb[0] * d.pnb[1] * d.pnb[2]; }

size_t lds_pme_spread(const Dev &d) {
  int bmax = 1;
  for (int k = 0; k < 3; k++) bmax *= (d.K[k] + d.pnb[k] - 1) / d.pnb[k] + (k == 2 ? 2 : 0);
  return (size_t)bmax * sizeof(long long);
}

// the three-kernel plane FFT applies (compiled size, square yz-plane in LDS)
bool fft_plane_path(const Dev &d) {
  bool ct0 = false, ct1 = false;
#define BM_IS(SZ) ct0 = ct0 || d.K[0] == SZ; ct1 = ct1 || d.K[1] == SZ;
  BM_FFT_SIZES(BM_IS)
#undef BM_IS
  return ct0 && ct1 && d.K[1] == d.K[2] && d.K[1] * d.K[2] <= 4096;
}

int grid_fft_planes(const Dev &d) { return d.K[0]; }

int grid_fft_x(const Dev &d) { return (d.K[1] * d.K[2] + 15) / 16; }

void launch_pme_fft(const Dev &d, bool energy, hipStream_t st) {
  if (d.fft_mode == 1) {
    launch_pme_fft_mfma(d, energy, st);
    return;
  }
  if (fft_plane_path(d)) {
    // plane path: forward z+y per x-plane, x with convolution, inverse y+z
    const unsigned gx = (unsigned)grid_fft_x(d);
    switch (d.K[1]) {
#define BM_YZF(SZ) case SZ: fft_yz_fwd<SZ, SZ><<<d.K[0], 1024, 0, st>>>(d); break;
      BM_FFT_SIZES(BM_YZF)
#undef BM_YZF
    }
    switch (d.K[0]) {
#define BM_XC(SZ) case SZ: fft_x_conv<SZ, 16><<<gx, 256, 0, st>>>(d, energy ? 1 : 0); break;
      BM_FFT_SIZES(BM_XC)
#undef BM_XC
    }
    switch (d.K[1]) {
#define BM_YZI(SZ) case SZ: fft_yz_inv<SZ, SZ><<<d.K[0], 1024, 0, st>>>(d); break;
      BM_FFT_SIZES(BM_YZI)
#undef BM_YZI
    }
  } else {
    const int G = d.K[0] * d.K[1] * d.K[2];
    const float2 *tw[3] = {d.twiddle, d.twiddle + d.K[0], d.twiddle + d.K[0] + d.K[1]};
    FftArgs a;
    // forward z (from the real grid), forward y, x with convolution,
    // inverse y, inverse z (real part to the real grid)
    const int seq_axis[5] = {2, 1, 0, 1, 2};
    const int seq_mode[5] = {0, 1, 3, 2, 4};
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
  }
}

void launch_pme_grid(const Dev &d, bool energy, hipStream_t st) {
  launch_pme_spread(d, st);
  launch_pme_fft(d, energy, st);
}

void launch_pme_spread(const Dev &d, hipStream_t st) {
  const int N = d.order;
  const unsigned nbr = (unsigned)grid_pme_spread(d);
  const size_t lds = lds_pme_spread(d);
  const unsigned gb = (unsigned)grid_pme_bin(d);
  if (N == 4) {
    pme_bin<4><<<gb, 256, 0, st>>>(d);
    pme_spread<4><<<nbr, 256, lds, st>>>(d);
  } else if (N == 6) {
    pme_bin<6><<<gb, 256, 0, st>>>(d);
    pme_spread<6><<<nbr, 256, lds, st>>>(d);
  } else {
    pme_bin<8><<<gb, 256, 0, st>>>(d);
    pme_spread<8><<<nbr, 256, lds, st>>>(d);
  }
}

void launch_pme_bin_multi(const Multi &m, int order, hipStream_t st) {
  if (order == 4) pme_bin_multi<4><<<m.nwg, 256, 0, st>>>(m);
  else if (order == 6) pme_bin_multi<6><<<m.nwg, 256, 0, st>>>(m);
  else pme_bin_multi<8><<<m.nwg, 256, 0, st>>>(m);
}

void launch_pme_spread_multi(const Multi &m, int order, size_t lds, hipStream_t st) {
  if (order == 4) pme_spread_multi<4><<<m.nwg, 256, lds, st>>>(m);
  else if (order == 6) pme_spread_multi<6><<<m.nwg, 256, lds, st>>>(m);
  else pme_spread_multi<8><<<m.nwg, 256, lds, st>>>(m);
}

// systems of one grid size on the plane path: mp launches K0 workgroups per
// system, mx grid_fft_x() per system.  The plane kernels run 256 threads
// here (the bodies loop over blockDim), so they fit on CUs beside the
// nonbonded kernel of the same step.
void launch_pme_fft_multi(const Multi &mp, const Multi &mx, int K0, int K1, bool energy, hipStream_t st) {
  switch (K1) {
#define BM_YZF(SZ) case SZ: fft_yz_fwd_multi<SZ, SZ><<<mp.nwg, 256, 0, st>>>(mp); break;
    BM_FFT_SIZES(BM_YZF)
#undef BM_YZF
  }
  switch (K0) {
#define BM_XC(SZ) case SZ: fft_x_conv_multi<SZ, 16><<<mx.nwg, 256, 0, st>>>(mx, energy ? 1 : 0); break;
    BM_FFT_SIZES(BM_XC)
#undef BM_XC
  }
  switch (K1) {
#define BM_YZI(SZ) case SZ: fft_yz_inv_multi<SZ, SZ><<<mp.nwg, 256, 0, st>>>(mp); break;
    BM_FFT_SIZES(BM_YZI)
#undef BM_YZI
  }
}

template <bool KICK>
static void assemble_t(const Dev &d, unsigned mask, unsigned g, hipStream_t st) {
  if (!(mask & kAsmPme)) force_assemble<0, KICK><<<g, 256, 0, st>>>(d, mask);
  else if (d.order == 4) force_assemble<4, KICK><<<g, 256, 0, st>>>(d, mask);
  else if (d.order == 6) force_assemble<6, KICK><<<g, 256, 0, st>>>(d, mask);
  else force_assemble<8, KICK><<<g, 256, 0, st>>>(d, mask);
}

int grid_assemble(const Dev &d) {
  const int nbr = d.elecmethod == BM_PME ? d.pnb[0] * d.pnb[1] * d.pnb[2] : 0;
  return (int)((std::max(8LL * 32LL * d.maxblocks, (long long)nbr) + 255) / 256);
}

void launch_force_assemble(const Dev &d, unsigned mask, hipStream_t st, bool kick) {
  const unsigned g = (unsigned)grid_assemble(d);
  if (d.elecmethod != BM_PME) mask &= ~kAsmPme;
  if (kick) assemble_t<true>(d, mask, g, st);
  else assemble_t<false>(d, mask, g, st);
}

// order: the spline order of PME systems, 0 for the others (no gather)
void launch_assemble_multi(const Multi &m, int order, unsigned mask, bool kick, hipStream_t st) {
  if (order == 0) mask &= ~kAsmPme;
#define BM_ASM(NN)                                                       \
  if (kick) force_assemble_multi<NN, true><<<m.nwg, 256, 0, st>>>(m, mask); \
  else force_assemble_multi<NN, false><<<m.nwg, 256, 0, st>>>(m, mask)
  if (order == 0) {
    BM_ASM(0);
  } else if (order == 4) {
    BM_ASM(4);
  } else if (order == 6) {
    BM_ASM(6);
  } else {
    BM_ASM(8);
  }
#undef BM_ASM
}

}  // namespace blade
