// fft_mfma.hip - the PME 3D FFT as batched DFT contractions on the matrix
// cores (the precision form: split-BF16x3 products, FP32 accumulation).
//
// A pass transforms the grid along one axis.  A workgroup stages 32 lines
// of K points in LDS and computes the K-point DFT of all of them as one
// real product
//     [Yr]   [ C   -sS ] [Xr]        C = cos(2 pi j k / K), S = sin(..)
//     [Yi] = [ sS   C  ] [Xi]        s = -1 forward, +1 inverse
// of a 2K x 2K matrix with the 2K x 32 block of lines.  Each wave computes
// one 32-row tile of the result with v_mfma_f32_32x32x16_bf16.  Operands are
// split into bf16 pairs, x = hi + lo (hi = RNE(x), lo = RNE(x - hi), about
// 2^-17 relative), and every product is hi*hi + hi*lo + lo*hi (lo*lo,
// about 2^-18 relative, is dropped), accumulated in FP32.  The matrix
// halves are precomputed on the host from FP64 in the order the A operand
// reads them.  Five passes make the PME transform: z forward from the real
// charge grid, y forward, x forward + convolution (+ k-space energy in
// FP64) + x inverse, y inverse, z inverse to the real potential grid.
//
// Lanes of v_mfma_f32_32x32x16_bf16 (checked on gfx950 with integer
// operands): lane l holds A[l%32][8(l/32)+j] and B[8(l/32)+j][l%32],
// j < 8, and D[8(r/4)+4(l/32)+r%4][l%32] in accumulator r < 16.
#include "kernels.h"

namespace blade {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaLines = 32;

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

// In-place DFT of the kMfmaLines lines staged in tile (element k of line n
// at tile[n*(K+1)+k]; the pad keeps the lanes of a wave in distinct LDS
// banks).  The workgroup has 2K/32 waves, wave rt computing output rows
// 32rt .. 32rt+31 of the real product.
template <int K>
__device__ __forceinline__ void dft_lines(float2 *tile, const bf16x8 *Mhi, const bf16x8 *Mlo) {
  constexpr int KS = 2 * K / 16;
  const int lane = threadIdx.x & 63, rt = threadIdx.x >> 6;
  const int n = lane & 31, kh = 8 * (lane >> 5);
  const float2 *row = tile + n * (K + 1);
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
  const bf16x8 *ah_p = Mhi + (long long)rt * KS * 64 + lane;
  const bf16x8 *al_p = Mlo + (long long)rt * KS * 64 + lane;
#pragma unroll 2
  for (int ks = 0; ks < KS; ks++) {
    const int k0 = 16 * ks + kh;   // this lane's 8 rows of B: all Xr or all Xi
    bf16x8 bh, bl;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float v = k0 < K ? row[k0 + j].x : row[k0 - K + j].y;
      const __bf16 h = (__bf16)v;
      bh[j] = h;
      bl[j] = (__bf16)(v - (float)h);
    }
    const bf16x8 ah = ah_p[ks * 64], al = al_p[ks * 64];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
  }
  __syncthreads();   // every wave has read the lines
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int m = 32 * rt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m < K) tile[n * (K + 1) + m].x = acc[r];
    else tile[n * (K + 1) + m - K].y = acc[r];
  }
  __syncthreads();
}

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
// 3: y inverse; 4: z inverse, real part -> real potential grid
template <int K, int AXIS, int MODE>
__global__ void __launch_bounds__(64 * (2 * K / 32)) fft_mfma_pass(Dev d) {
  __shared__ float2 tile[kMfmaLines * (K + 1)];
  const int L0 = blockIdx.x * kMfmaLines;
  const int dir = MODE <= 1 ? 0 : 1;
  const bf16x8 *Mhi = reinterpret_cast<const bf16x8 *>(d.mdft[AXIS][dir][0]);
  const bf16x8 *Mlo = reinterpret_cast<const bf16x8 *>(d.mdft[AXIS][dir][1]);
  // stage: consecutive threads take consecutive grid addresses
  for (int e = threadIdx.x; e < kMfmaLines * K; e += blockDim.x) {
    int n, k;
    if (AXIS == 2) { n = e / K; k = e - n * K; }
    else { k = e / kMfmaLines; n = e - k * kMfmaLines; }
    const long long a = grid_at<AXIS>(d, L0 + n, k);
    tile[n * (K + 1) + k] = MODE == 0 ? make_float2(d.rgrid[a], 0.f) : d.cgrid[a];
  }
  __syncthreads();
  dft_lines<K>(tile, Mhi, Mlo);
  for (int e = threadIdx.x; e < kMfmaLines * K; e += blockDim.x) {
    int n, k;
    if (AXIS == 2) { n = e / K; k = e - n * K; }
    else { k = e / kMfmaLines; n = e - k * kMfmaLines; }
    const long long a = grid_at<AXIS>(d, L0 + n, k);
    const float2 v = tile[n * (K + 1) + k];
    if (MODE == 4) d.rgrid[a] = v.x;
    else d.cgrid[a] = v;
  }
}

// x lines: forward DFT, the convolution B(m) kE/(2 pi V) exp(-pi^2 m^2 /
// beta^2)/m^2 (with the k-space energy sum in FP64 on energy steps), and
// the inverse DFT, in LDS
template <int K>
__global__ void __launch_bounds__(64 * (2 * K / 32)) fft_mfma_x_conv(Dev d, int energy) {
  __shared__ float2 tile[kMfmaLines * (K + 1)];
  const int L0 = blockIdx.x * kMfmaLines;
  for (int e = threadIdx.x; e < kMfmaLines * K; e += blockDim.x) {
    const int k = e / kMfmaLines, n = e - k * kMfmaLines;
    tile[n * (K + 1) + k] = d.cgrid[grid_at<0>(d, L0 + n, k)];
  }
  __syncthreads();
  dft_lines<K>(tile, reinterpret_cast<const bf16x8 *>(d.mdft[0][0][0]),
               reinterpret_cast<const bf16x8 *>(d.mdft[0][0][1]));
  const float L0f = (float)d.box[0], L1f = (float)d.box[1], L2f = (float)d.box[2];
  double ek = 0.0;
  for (int e = threadIdx.x; e < kMfmaLines * K; e += blockDim.x) {
    const int m = e / kMfmaLines, n = e - m * kMfmaLines;
    const int line = L0 + n;
    const int j = line / d.K[2], k = line - j * d.K[2];
    const int mi = (2 * m <= K) ? m : m - K;
    const int mj = (2 * j <= d.K[1]) ? j : j - d.K[1];
    const int mk = (2 * k <= d.K[2]) ? k : k - d.K[2];
    const float kx = mi / L0f, ky = mj / L1f, kz = mk / L2f;
    const float m2 = kx * kx + ky * ky + kz * kz;
    float fac = 0.f;
    if (m || j || k) fac = d.bmod[0][m] * d.bmod[1][j] * d.bmod[2][k] * d.pmefac * __expf(-d.pmeexp * m2) / m2;
    float2 &v = tile[n * (K + 1) + m];
    if (energy && (m || j || k)) {
      const double kxd = mi / d.box[0], kyd = mj / d.box[1], kzd = mk / d.box[2];
      const double m2d = kxd * kxd + kyd * kyd + kzd * kzd;
      const double g = d.bmodd[0][m] * d.bmodd[1][j] * d.bmodd[2][k] * d.pmefacd * exp(-d.pmeexpd * m2d) / m2d;
      ek += g * ((double)v.x * v.x + (double)v.y * v.y);
    }
    v = make_float2(v.x * fac, v.y * fac);
  }
  if (energy) wave_atomic_fix(&d.energy[BM_ENBRECIP], (long long)__double2ll_rn(ek * kFixScaleD));
  __syncthreads();
  dft_lines<K>(tile, reinterpret_cast<const bf16x8 *>(d.mdft[0][1][0]),
               reinterpret_cast<const bf16x8 *>(d.mdft[0][1][1]));
  for (int e = threadIdx.x; e < kMfmaLines * K; e += blockDim.x) {
    const int k = e / kMfmaLines, n = e - k * kMfmaLines;
    d.cgrid[grid_at<0>(d, L0 + n, k)] = tile[n * (K + 1) + k];
  }
}

template <int AXIS, int MODE>
static void pass_t(const Dev &d, int nlines, hipStream_t st) {
  const int K = d.K[AXIS];
  const unsigned g = (unsigned)(nlines / kMfmaLines);
  switch (K) {
#define BM_P(SZ) \
  case SZ: fft_mfma_pass<SZ, AXIS, MODE><<<g, 64 * (2 * SZ / 32), 0, st>>>(d); break;
    BM_MFMA_SIZES(BM_P)
#undef BM_P
  }
}

void launch_pme_fft_mfma(const Dev &d, bool energy, hipStream_t st) {
  const int K0 = d.K[0], K1 = d.K[1], K2 = d.K[2];
  pass_t<2, 0>(d, K0 * K1, st);
  pass_t<1, 1>(d, K0 * K2, st);
  const unsigned gx = (unsigned)(K1 * K2 / kMfmaLines);
  switch (K0) {
#define BM_X(SZ) \
  case SZ: fft_mfma_x_conv<SZ><<<gx, 64 * (2 * SZ / 32), 0, st>>>(d, energy ? 1 : 0); break;
    BM_MFMA_SIZES(BM_X)
#undef BM_X
  }
  pass_t<1, 3>(d, K0 * K2, st);
  pass_t<2, 4>(d, K0 * K1, st);
}

}  // namespace blade
