// fft_mfma.hip - the PME 3D FFT as batched DFT contractions on the matrix
// cores (the precision form: split-BF16x3 products, FP32 accumulation).
//
// A K-point DFT of a block of lines is one real product
//     [Yr]   [ C   -sS ] [Xr]        C = cos(2 pi j k / K), S = sin(..)
//     [Yi] = [ sS   C  ] [Xi]        s = -1 forward, +1 inverse
// of the 2K x 2K matrix M with the 2K x 32 block of 32 lines.  A wave owns
// one 32-row tile of M (row tile rt) and computes that tile of the result
// for one or more 32-line tiles with v_mfma_f32_32x32x16_bf16.  Operands are
// split into bf16 pairs, x = hi + lo (hi = RNE(x), lo = RNE(x - hi), about
// 2^-17 relative), and every product is hi*hi + hi*lo + lo*hi (lo*lo, about
// 2^-18 relative, is dropped), accumulated in FP32.  The host splits M from
// FP64 in the order the A operand reads it; a wave loads its row tile once
// per kernel and keeps it in registers, and the inverse matrix is the
// forward one with the off-diagonal blocks negated (exact in bf16).
//
// Lines sit in LDS as separate real and imaginary planes, element k of line
// n at n*ls + k*es.  Three kernels make the PME transform:
//   yz forward: per x-plane, z DFT of the real charge grid (the imaginary
//     input is zero, so only the first K columns of M are used), then y DFT;
//   x: forward DFT of 32 x-lines, the convolution (and the k-space energy in
//     FP64 on energy steps), inverse DFT;
//   yz inverse: per x-plane, y DFT, then z DFT keeping only the real rows.
// Grids whose y and z edges differ run the y and z transforms as separate
// 32-line passes.
//
// Lanes of v_mfma_f32_32x32x16_bf16 (checked on gfx950 with integer
// operands): lane l holds A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32],
// j < 8, and D[8(r/4)+4(l/32)+r%4][l%32] in accumulator r < 16.
#include "kernels.h"

namespace blade {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// the grid sizes with compiled MFMA passes (multiples of 16 in the PME size
// rule {16,18,20,24,27,32} * 2^k)
#define BM_MFMA_SIZES(X) X(16) X(32) X(48) X(64) X(80) X(96) X(128)

bool fft_mfma_supported(int K) {
  switch (K) {
#define BM_CASE(SZ) case SZ:
    BM_MFMA_SIZES(BM_CASE)
#undef BM_CASE
    return true;
    default:
      return false;
  }
}

// Launch shape of a DFT over NL lines of K points: RT = 2K/32 row tiles,
// NT 32-line tiles, WL line-tile slots per row tile, P line tiles per wave.
// At most 8 waves (two per SIMD), so a wave has 256 registers for its row
// tile of M (8 KS), its P accumulators (16 P) and the rest.
template <int K, int NL>
struct DftShape {
  static constexpr int RT = 2 * K / 32, KS = 2 * K / 16, NT = (NL + 31) / 32;
  static constexpr int WLMAX = 8 / RT > 0 ? 8 / RT : 1;
  static constexpr int WL = NT < WLMAX ? NT : WLMAX;
  static constexpr int W = RT * WL, P = (NT + WL - 1) / WL;
};

// one wave's row tile of M: fragment ks holds M[32rt + l%32][16ks + 8(l/32) + j]
template <int K>
struct DftRows {
  static constexpr int KS = 2 * K / 16;
  bf16x8 h[KS], l[KS];
  __device__ __forceinline__ void load(const void *mhi, const void *mlo, int rt) {
    const bf16x8 *ph = reinterpret_cast<const bf16x8 *>(mhi) + (long long)rt * KS * 64 + (threadIdx.x & 63);
    const bf16x8 *pl = reinterpret_cast<const bf16x8 *>(mlo) + (long long)rt * KS * 64 + (threadIdx.x & 63);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      h[ks] = ph[ks * 64];
      l[ks] = pl[ks * 64];
    }
  }
  // forward <-> inverse: negate the entries with row < K != column < K
  __device__ __forceinline__ void flip(int rt) {
    const int lane = threadIdx.x & 63;
    const bool rlow = 32 * rt + (lane & 31) < K;
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const bool clow = 16 * ks + 8 * (lane >> 5) < K;
      const unsigned m = rlow != clow ? 0x80008000u : 0u;
      u32x4 a = __builtin_bit_cast(u32x4, h[ks]), b = __builtin_bit_cast(u32x4, l[ks]);
      a ^= m;
      b ^= m;
      h[ks] = __builtin_bit_cast(bf16x8, a);
      l[ks] = __builtin_bit_cast(bf16x8, b);
    }
  }
};

// In-place DFT of NL lines in LDS (real parts re[], imaginary parts im[],
// element k of line n at n*ls + k*es).  REAL_IN: the imaginary input is
// zero (only the first K columns of M).  REAL_OUT: only the real output is
// written (row tiles past K idle).  Ends with the workgroup synchronised.
template <int K, int NL, bool REAL_IN, bool REAL_OUT>
__device__ __forceinline__ void dft_lines(float *re, float *im, int ls, int es, const DftRows<K> &A) {
  using S = DftShape<K, NL>;
  constexpr int KS = REAL_IN ? S::KS / 2 : S::KS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rt = w % S::RT, lt0 = w / S::RT;
  const int n = lane & 31, kh = 8 * (lane >> 5);
  const bool active = w < S::W && !(REAL_OUT && 32 * rt >= K);
  f32x16 acc[S::P];
#pragma unroll
  for (int i = 0; i < S::P; i++) {
#pragma unroll
    for (int r = 0; r < 16; r++) acc[i][r] = 0.f;
    const int lt = lt0 + S::WL * i;
    if (!active || lt >= S::NT) continue;
    const int line = 32 * lt + n < NL ? 32 * lt + n : 0;   // lines past NL read line 0, write nothing
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int k0 = 16 * ks + kh;   // this lane's 8 rows of B: all real or all imaginary
      const float *src = k0 < K ? re + line * ls + k0 * es : im + line * ls + (k0 - K) * es;
      bf16x8 bh, bl;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const float v = src[j * es];
        const __bf16 h = (__bf16)v;
        bh[j] = h;
        bl[j] = (__bf16)(v - (float)h);
      }
      acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.h[ks], bh, acc[i], 0, 0, 0);
      acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.h[ks], bl, acc[i], 0, 0, 0);
      acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.l[ks], bh, acc[i], 0, 0, 0);
    }
  }
  __syncthreads();   // every wave has read its lines
#pragma unroll
  for (int i = 0; i < S::P; i++) {
    const int lt = lt0 + S::WL * i;
    if (!active || lt >= S::NT || 32 * lt + n >= NL) continue;
    const int line = 32 * lt + n;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int m = 32 * rt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (m < K) re[line * ls + m * es] = acc[i][r];
      else if (!REAL_OUT) im[line * ls + (m - K) * es] = acc[i][r];
    }
  }
  __syncthreads();
}

// ---- y-z planes (K1 == K2 == K): one workgroup per x-plane ----------------
// plane element (y, z) at y*(K+4) + z: z lines have ls = K+4, es = 1; y lines
// ls = 1, es = K+4

template <int K, bool INV>
__device__ __forceinline__ void fft_mfma_yz_body(const Dev &d, int bid) {
  constexpr int LD = K + 4;
  __shared__ float re[K * LD], im[K * LD];
  const int rt = (threadIdx.x >> 6) % DftShape<K, K>::RT;
  DftRows<K> A;
  A.load(d.mdft[2][0], d.mdft[2][1], rt);   // the y and z matrices are equal when K1 == K2
  if (INV) A.flip(rt);
  const long long base = (long long)bid * K * K;
  // the plane's loads are all issued before the first LDS store
  constexpr int NTH = 64 * DftShape<K, K>::W, NIT = (K * K + NTH - 1) / NTH;
  if (INV) {
    float2 v[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH;
      v[it] = e < K * K ? d.cgrid[base + e] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH, y = e / K, z = e - y * K;
      if (e < K * K) {
        re[y * LD + z] = v[it].x;
        im[y * LD + z] = v[it].y;
      }
    }
  } else {
    float v[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH;
      v[it] = e < K * K ? d.rgrid[base + e] : 0.f;
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH, y = e / K, z = e - y * K;
      if (e < K * K) re[y * LD + z] = v[it];
    }
  }
  __syncthreads();
  if (!INV) {
    dft_lines<K, K, true, false>(re, im, LD, 1, A);    // z, real input
    dft_lines<K, K, false, false>(re, im, 1, LD, A);   // y
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH, y = e / K, z = e - y * K;
      if (e < K * K) d.cgrid[base + e] = make_float2(re[y * LD + z], im[y * LD + z]);
    }
  } else {
    dft_lines<K, K, false, false>(re, im, 1, LD, A);   // y
    dft_lines<K, K, false, true>(re, im, LD, 1, A);    // z, real output
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH, y = e / K, z = e - y * K;
      if (e < K * K) d.rgrid[base + e] = re[y * LD + z];
    }
  }
}

template <int K, bool INV>
__global__ void __launch_bounds__((64 * DftShape<K, K>::W)) fft_mfma_yz(Dev d) {
  fft_mfma_yz_body<K, INV>(d, blockIdx.x);
}

// one launch over the x-planes of several systems with the same K
template <int K, bool INV>
__global__ void __launch_bounds__((64 * DftShape<K, K>::W)) fft_mfma_yz_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  fft_mfma_yz_body<K, INV>(d, b.bid);
}

// ---- single-axis passes of 32 lines (y and z when K1 != K2) ---------------

// address of element k of line L along axis AXIS of a K0 x K1 x K2 grid
template <int AXIS>
__device__ __forceinline__ long long grid_at(const Dev &d, int L, int k) {
  if (AXIS == 2) return (long long)L * d.K[2] + k;
  if (AXIS == 1) {
    const int x = L / d.K[2], z = L - x * d.K[2];
    return ((long long)x * d.K[1] + k) * d.K[2] + z;
  }
  return (long long)k * d.K[1] * d.K[2] + L;
}

// MODE 0: z forward, real charge grid -> complex grid; 1: y forward;
// 3: y inverse; 4: z inverse, real part -> real potential grid.  Tile
// element k of line n at k*33 + n.
template <int K, int AXIS, int MODE>
__global__ void __launch_bounds__((64 * DftShape<K, 32>::W)) fft_mfma_pass(Dev d) {
  constexpr int LD = 33;
  __shared__ float re[K * LD], im[K * LD];
  const int rt = (threadIdx.x >> 6) % DftShape<K, 32>::RT;
  DftRows<K> A;
  A.load(d.mdft[AXIS][0], d.mdft[AXIS][1], rt);
  if (MODE >= 3) A.flip(rt);
  const int L0 = blockIdx.x * 32;
  // 32K elements over 64 W = 4K threads: 8 per thread, loads first
  constexpr int NTH = 64 * DftShape<K, 32>::W, NIT = 32 * K / NTH;
  static_assert(NIT * NTH == 32 * K, "whole tiles per thread");
  float2 v[NIT];
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int e = threadIdx.x + it * NTH;
    const int n = AXIS == 2 ? e / K : e % 32, k = AXIS == 2 ? e % K : e / 32;
    const long long a = grid_at<AXIS>(d, L0 + n, k);
    v[it] = MODE == 0 ? make_float2(d.rgrid[a], 0.f) : d.cgrid[a];
  }
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int e = threadIdx.x + it * NTH;
    const int n = AXIS == 2 ? e / K : e % 32, k = AXIS == 2 ? e % K : e / 32;
    re[k * LD + n] = v[it].x;
    if (MODE != 0) im[k * LD + n] = v[it].y;
  }
  __syncthreads();
  dft_lines<K, 32, MODE == 0, MODE == 4>(re, im, 1, LD, A);
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int e = threadIdx.x + it * NTH;
    const int n = AXIS == 2 ? e / K : e % 32, k = AXIS == 2 ? e % K : e / 32;
    const long long a = grid_at<AXIS>(d, L0 + n, k);
    if (MODE == 4) d.rgrid[a] = re[k * LD + n];
    else d.cgrid[a] = make_float2(re[k * LD + n], im[k * LD + n]);
  }
}

// ---- x lines: forward DFT, convolution, inverse DFT -----------------------
// the convolution multiplies by B(m) kE/(2 pi V) exp(-pi^2 m^2 / beta^2)/m^2
// and, on energy steps, sums the k-space energy in FP64
template <int K>
__device__ __forceinline__ void fft_mfma_x_conv_body(const Dev &d, int bid, int energy) {
  constexpr int LD = 33;
  __shared__ float re[K * LD], im[K * LD];
  const int rt = (threadIdx.x >> 6) % DftShape<K, 32>::RT;
  DftRows<K> A;
  A.load(d.mdft[0][0], d.mdft[0][1], rt);
  const int L0 = bid * 32;
  constexpr int NTH = 64 * DftShape<K, 32>::W, NIT = 32 * K / NTH;
  static_assert(NIT * NTH == 32 * K, "whole tiles per thread");
  {
    float2 v[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH;
      v[it] = d.cgrid[grid_at<0>(d, L0 + e % 32, e / 32)];
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
      const int e = threadIdx.x + it * NTH, k = e / 32, n = e % 32;
      re[k * LD + n] = v[it].x;
      im[k * LD + n] = v[it].y;
    }
  }
  __syncthreads();
  dft_lines<K, 32, false, false>(re, im, 1, LD, A);
  A.flip(rt);
  const float L0f = (float)d.box[0], L1f = (float)d.box[1], L2f = (float)d.box[2];
  double ek = 0.0;
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int e = threadIdx.x + it * NTH;
    const int m = e / 32, n = e - m * 32;
    const int line = L0 + n;
    const int j = line / d.K[2], k = line - j * d.K[2];
    const int mi = (2 * m <= K) ? m : m - K;
    const int mj = (2 * j <= d.K[1]) ? j : j - d.K[1];
    const int mk = (2 * k <= d.K[2]) ? k : k - d.K[2];
    const float kx = mi / L0f, ky = mj / L1f, kz = mk / L2f;
    const float m2 = kx * kx + ky * ky + kz * kz;
    float fac = 0.f;
    if (m || j || k) fac = d.bmod[0][m] * d.bmod[1][j] * d.bmod[2][k] * d.pmefac * __expf(-d.pmeexp * m2) / m2;
    const float vr = re[m * LD + n], vi = im[m * LD + n];
    if (energy && (m || j || k)) {
      const double kxd = mi / d.box[0], kyd = mj / d.box[1], kzd = mk / d.box[2];
      const double m2d = kxd * kxd + kyd * kyd + kzd * kzd;
      const double g = d.bmodd[0][m] * d.bmodd[1][j] * d.bmodd[2][k] * d.pmefacd * exp(-d.pmeexpd * m2d) / m2d;
      ek += g * ((double)vr * vr + (double)vi * vi);
    }
    re[m * LD + n] = vr * fac;
    im[m * LD + n] = vi * fac;
  }
  if (energy) wave_atomic_fix(&d.energy[BM_ENBRECIP], (long long)__double2ll_rn(ek * kFixScaleD));
  __syncthreads();
  dft_lines<K, 32, false, false>(re, im, 1, LD, A);
#pragma unroll
  for (int it = 0; it < NIT; it++) {
    const int e = threadIdx.x + it * NTH, k = e / 32, n = e % 32;
    d.cgrid[grid_at<0>(d, L0 + n, k)] = make_float2(re[k * LD + n], im[k * LD + n]);
  }
}

template <int K>
__global__ void __launch_bounds__((64 * DftShape<K, 32>::W)) fft_mfma_x_conv(Dev d, int energy) {
  fft_mfma_x_conv_body<K>(d, blockIdx.x, energy);
}

template <int K>
__global__ void __launch_bounds__((64 * DftShape<K, 32>::W)) fft_mfma_x_conv_multi(Multi m, int energy) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  fft_mfma_x_conv_body<K>(d, b.bid, energy);
}

template <int AXIS, int MODE>
static void pass_t(const Dev &d, int nlines, hipStream_t st) {
  const unsigned g = (unsigned)(nlines / 32);
  switch (d.K[AXIS]) {
#define BM_P(SZ) \
  case SZ: fft_mfma_pass<SZ, AXIS, MODE><<<g, 64 * DftShape<SZ, 32>::W, 0, st>>>(d); break;
    BM_MFMA_SIZES(BM_P)
#undef BM_P
  }
}

template <bool INV>
static void yz_t(const Dev &d, hipStream_t st) {
  switch (d.K[1]) {
#define BM_YZ(SZ) \
  case SZ: fft_mfma_yz<SZ, INV><<<(unsigned)d.K[0], 64 * DftShape<SZ, SZ>::W, 0, st>>>(d); break;
    BM_MFMA_SIZES(BM_YZ)
#undef BM_YZ
  }
}

// cubic-plane members (K1 == K2) of one grid size: mp holds
// grid_fft_mfma_planes() workgroups per system, mx grid_fft_mfma_x()
int grid_fft_mfma_planes(const Dev &d) { return d.K[0]; }
int grid_fft_mfma_x(const Dev &d) { return d.K[1] * d.K[2] / 32; }

void launch_pme_fft_mfma_multi(const Multi &mp, const Multi &mx, int K0, int K1, bool energy, hipStream_t st) {
  switch (K1) {
#define BM_YZ(SZ) \
  case SZ: fft_mfma_yz_multi<SZ, false><<<(unsigned)mp.nwg, 64 * DftShape<SZ, SZ>::W, 0, st>>>(mp); break;
    BM_MFMA_SIZES(BM_YZ)
#undef BM_YZ
  }
  switch (K0) {
#define BM_X(SZ) \
  case SZ: fft_mfma_x_conv_multi<SZ><<<(unsigned)mx.nwg, 64 * DftShape<SZ, 32>::W, 0, st>>>(mx, energy ? 1 : 0); break;
    BM_MFMA_SIZES(BM_X)
#undef BM_X
  }
  switch (K1) {
#define BM_YZ(SZ) \
  case SZ: fft_mfma_yz_multi<SZ, true><<<(unsigned)mp.nwg, 64 * DftShape<SZ, SZ>::W, 0, st>>>(mp); break;
    BM_MFMA_SIZES(BM_YZ)
#undef BM_YZ
  }
}

void launch_pme_fft_mfma(const Dev &d, bool energy, hipStream_t st) {
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2];
  const bool planes = K1 == K2;
  if (planes) yz_t<false>(d, st);
  else {
    pass_t<2, 0>(d, K0 * K1, st);
    pass_t<1, 1>(d, K0 * K2, st);
  }
  const unsigned gx = (unsigned)(K1 * K2 / 32);
  switch (K0) {
#define BM_X(SZ) \
  case SZ: fft_mfma_x_conv<SZ><<<gx, 64 * DftShape<SZ, 32>::W, 0, st>>>(d, energy ? 1 : 0); break;
    BM_MFMA_SIZES(BM_X)
#undef BM_X
  }
  if (planes) yz_t<true>(d, st);
  else {
    pass_t<1, 3>(d, K0 * K2, st);
    pass_t<2, 4>(d, K0 * K1, st);
  }
}

}  // namespace blade
