// domdec.hip - cell sort, blocks, tile list and exclusion masks
// (restates domdec/assign_domain.cu:14-45, assign_blocks.cu:24-468,
// cull.cu:44-375, assign_excl.cu:17-371, recull.cu:41-146 on 32-bit
// fractional coordinates; orc_domdec in the oracle is the same rule set).
//
// Rebuild sequence (no host synchronisation, fixed grids, graph-capturable):
//   dd_wrap_count -> dd_scan_columns -> dd_scatter -> dd_sort_columns ->
//   dd_tile_count -> dd_scan_tiles -> dd_tile_write -> dd_excl_mark ->
//   dd_excl_rows -> dd_scan_excl -> dd_excl_assign -> dd_excl_clear ->
//   dd_excl_set
// Every step: dd_pack (slot positions, block boxes, accumulator reset).
//
// The tile list is directed: row I holds every block J (J == I included)
// whose rebuild box is within rcut + pad, in ascending J, so each pair of
// blocks appears as (I,J) and (J,I).  Rows are cut into chunks of at most
// d.tchunk tiles; the nonbonded kernel owns one chunk (one i-block) at a
// time and never writes j-side forces.
//
// Column counts and bucket fills walk the slots of the previous rebuild,
// which are ordered by column, so the lanes of a wave mostly share one
// column and one atomic serves them all (device-scope atomics on one
// address serialise across the chip).
#include "kernels.h"

#pragma clang fp contract(off)

namespace blade {

// One workgroup of 1024 threads: exclusive scan of in[0..n) into out,
// out[n] = total.  Deterministic integer arithmetic.
__device__ void block_scan_1024(const int *in, int *out, int n, int *total) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = tid * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - s;
  for (int i = b; i < e; i++) {
    int v = in[i];
    out[i] = run;
    run += v;
  }
  if (tid == 1023) {
    out[n] = part[1023];
    if (total) *total = part[1023];
  }
}

__device__ __forceinline__ void dd_zero_columns_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  int c = bid * blockDim.x + threadIdx.x;
  if (c < d.ncol) {
    d.colcount[c] = 0;
    d.colfill[c] = 0;
  }
}
__global__ void __launch_bounds__(256) dd_zero_columns(Dev d) { dd_zero_columns_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_zero_columns_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_zero_columns_body(d, b.bid, b.nblk);
}

// wrap every atom into the box, fractional coordinates, column counts
// (one thread per slot of the previous rebuild; every atom has one slot)
__device__ __forceinline__ void dd_wrap_count_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int s = bid * blockDim.x + threadIdx.x;
  const int i = s < 32 * d.maxblocks ? d.l2g[s] : -1;
  int col = 0;
  if (i >= 0) {
    unsigned u[3];
    for (int k = 0; k < 3; k++) {
      double x = d.pos[k * d.n + i];
      nan_check(d, x, i);
      double w = floor(__dmul_rn(x, d.invbox[k]));
      x = __dsub_rn(x, __dmul_rn(w, d.box[k]));
      d.pos[k * d.n + i] = x;
      u[k] = frac_u32(x, d.invbox[k]);
      d.ufrac[k * d.n + i] = u[k];
    }
    unsigned cx = (unsigned)(((unsigned long long)u[0] * (unsigned)d.ncx) >> 32);
    unsigned cy = (unsigned)(((unsigned long long)u[1] * (unsigned)d.ncy) >> 32);
    col = cx * d.ncy + cy;
  }
  wave_claim(d.colcount, col, i >= 0);
}
__global__ void __launch_bounds__(256) dd_wrap_count(Dev d) { dd_wrap_count_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_wrap_count_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_wrap_count_body(d, b.bid, b.nblk);
}

// blocks per column = ceil(count/32); block starts by exclusive scan
__device__ __forceinline__ void dd_scan_columns_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  __shared__ int nb[4096];
  int *tmp = d.bucket;  // not yet used this rebuild: scratch for ncol > 4096
  const bool big = d.ncol > 4096;
  for (int c = threadIdx.x; c < d.ncol; c += 1024) {
    int v = (d.colcount[c] + kBlockAtoms - 1) / kBlockAtoms;
    if (big) tmp[c] = v; else nb[c] = v;
  }
  __syncthreads();
  block_scan_1024(d.colcount, d.colstart, d.ncol, nullptr);
  __syncthreads();
  block_scan_1024(big ? tmp : nb, d.blkstart, d.ncol, d.nblocks);
}
__global__ void __launch_bounds__(1024) dd_scan_columns(Dev d) { dd_scan_columns_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) dd_scan_columns_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_scan_columns_body(d, b.bid, b.nblk);
}

// atoms into column buckets (bucket order is arbitrary; the column sort
// fixes the final order)
__device__ __forceinline__ void dd_scatter_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int s = bid * blockDim.x + threadIdx.x;
  const int i = s < 32 * d.maxblocks ? d.l2g[s] : -1;
  int c = 0;
  if (i >= 0) {
    unsigned cx = (unsigned)(((unsigned long long)d.ufrac[i] * (unsigned)d.ncx) >> 32);
    unsigned cy = (unsigned)(((unsigned long long)d.ufrac[d.n + i] * (unsigned)d.ncy) >> 32);
    c = cx * d.ncy + cy;
  }
  const int f = wave_claim(d.colfill, c, i >= 0);
  if (i >= 0) d.bucket[d.colstart[c] + f] = i;
}
__global__ void __launch_bounds__(256) dd_scatter(Dev d) { dd_scatter_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_scatter_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_scatter_body(d, b.bid, b.nblk);
}

// One workgroup per column: rank atoms by (z fraction, atom index), place
// them into 32-slot blocks, then the exact rebuild box of every block from
// the fractions held in LDS (one half-wave per block).
__device__ __forceinline__ void dd_sort_columns_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  __shared__ unsigned long long key[kMaxColumnAtoms];
  __shared__ unsigned sx[kMaxColumnAtoms], sy[kMaxColumnAtoms];
  __shared__ short sorted[kMaxColumnAtoms];
  const int c = bid;
  const int start = d.colstart[c], cnt = d.colcount[c];
  const int b0 = d.blkstart[c], nbc = d.blkstart[c + 1] - b0;
  if (cnt > kMaxColumnAtoms) {
    if (threadIdx.x == 0) atomicOr(d.errflag, 1);
    return;
  }
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    int a = d.bucket[start + k];
    key[k] = ((unsigned long long)d.ufrac[2 * d.n + a] << 32) | (unsigned)a;
    sx[k] = d.ufrac[a];
    sy[k] = d.ufrac[d.n + a];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    unsigned long long kk = key[k];
    int rank = 0;
    for (int m = 0; m < cnt; m++) rank += key[m] < kk;
    int a = (int)(kk & 0xFFFFFFFFull);
    int slot = (b0 + rank / kBlockAtoms) * kBlockAtoms + rank % kBlockAtoms;
    d.l2g[slot] = a;
    d.g2s[a] = slot;
    sorted[rank] = (short)k;
  }
  for (int s = threadIdx.x; s < nbc * kBlockAtoms; s += blockDim.x)
    if (s >= cnt) d.l2g[b0 * kBlockAtoms + s] = -1;
  // slots past the last block may hold atoms of an earlier, longer list:
  // the next rebuild walks every slot, so clear them
  if (c == 0)
    for (int s = *d.nblocks * kBlockAtoms + threadIdx.x; s < d.maxblocks * kBlockAtoms; s += blockDim.x)
      d.l2g[s] = -1;
  __syncthreads();
  // block boxes: half-wave h of the workgroup takes blocks h, h+8, ...
  const int h = threadIdx.x >> 5, l = threadIdx.x & 31;
  for (int bb = h; bb < ((nbc + 7) / 8) * 8; bb += 8) {
    const int r = bb * kBlockAtoms + l;
    const bool in = bb < nbc && r < cnt;
    unsigned lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    if (in) {
      const int k = sorted[r];
      const unsigned u[3] = {sx[k], sy[k], (unsigned)(key[k] >> 32)};
      for (int q = 0; q < 3; q++) lo[q] = hi[q] = u[q];
    }
    for (int q = 0; q < 3; q++)
      for (int o = 16; o > 0; o >>= 1) {
        lo[q] = min(lo[q], (unsigned)__shfl_xor((int)lo[q], o, 32));
        hi[q] = max(hi[q], (unsigned)__shfl_xor((int)hi[q], o, 32));
      }
    if (bb < nbc && l < 3) {
      d.bblo[3 * (b0 + bb) + l] = l == 0 ? lo[0] : l == 1 ? lo[1] : lo[2];
      d.bbhi[3 * (b0 + bb) + l] = l == 0 ? hi[0] : l == 1 ? hi[1] : hi[2];
    }
  }
}
__global__ void __launch_bounds__(256) dd_sort_columns(Dev d) { dd_sort_columns_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_sort_columns_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_sort_columns_body(d, b.bid, b.nblk);
}

// Periodic gap between two rebuild boxes (fractional counts), squared, in
// A^2, against (rcut+pad)^2; rounding as in the oracle (no contraction).
__device__ __forceinline__ bool dd_close(const Dev &d, const unsigned (&aLo)[3], const unsigned (&aHi)[3],
                                         const unsigned (&bLo)[3], const unsigned (&bHi)[3]) {
  const unsigned long long P = 1ull << 32;
  double g[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    unsigned long long aL = aLo[k], aH = aHi[k], bL = bLo[k], bH = bHi[k];
    unsigned long long gap = 0;
    if (aH < bL) gap = min(bL - aH, aL + P - bH);
    else if (bH < aL) gap = min(aL - bH, bL + P - aH);
    g[k] = __dmul_rn((double)gap, d.ddscale[k]);
  }
  double s = __dadd_rn(__dmul_rn(g[0], g[0]), __dmul_rn(g[1], g[1]));
  s = __dadd_rn(s, __dmul_rn(g[2], g[2]));
  return s <= d.rlist2;
}

constexpr int kBoxLds = 2048;   // block boxes staged in LDS by the tile kernels

// stage the first kBoxLds block boxes in LDS (every thread of the block
// must call it)
__device__ __forceinline__ void dd_stage_boxes(const Dev &d, int nb, unsigned *lo, unsigned *hi) {
  const int m = 3 * min(nb, kBoxLds);
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    lo[k] = d.bblo[k];
    hi[k] = d.bbhi[k];
  }
  __syncthreads();
}

__device__ __forceinline__ bool dd_tile_keep(const Dev &d, const unsigned *lo, const unsigned *hi, int I,
                                             const unsigned (&aLo)[3], const unsigned (&aHi)[3], int J) {
  if (J == I) return true;
  unsigned bLo[3], bHi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    bLo[k] = J < kBoxLds ? lo[3 * J + k] : d.bblo[3 * J + k];
    bHi[k] = J < kBoxLds ? hi[3 * J + k] : d.bbhi[3 * J + k];
  }
  return dd_close(d, aLo, aHi, bLo, bHi);
}

// one wave per i-block: number of blocks J within the list radius (all J)
__device__ __forceinline__ void dd_tile_count_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  __shared__ unsigned s_lo[3 * kBoxLds], s_hi[3 * kBoxLds];
  const int nb = *d.nblocks;
  dd_stage_boxes(d, nb, s_lo, s_hi);
  const int I = (bid * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (I >= nb) return;
  unsigned aLo[3], aHi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    aLo[k] = d.bblo[3 * I + k];
    aHi[k] = d.bbhi[3 * I + k];
  }
  int cnt = 0;
  for (int J0 = 0; J0 < nb; J0 += 64) {
    const int J = J0 + lane;
    const bool keep = J < nb && dd_tile_keep(d, s_lo, s_hi, I, aLo, aHi, J);
    cnt += __popcll(__ballot(keep));
  }
  if (lane == 0) {
    d.tcount[I] = cnt;
    d.ccount[I] = (cnt + d.tchunk - 1) / d.tchunk;
  }
}
__global__ void __launch_bounds__(256) dd_tile_count(Dev d) { dd_tile_count_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_tile_count_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_tile_count_body(d, b.bid, b.nblk);
}

// row starts of tiles and chunks; flags a list that outgrows its buffers
__device__ __forceinline__ void dd_scan_tiles_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int nb = *d.nblocks;
  block_scan_1024(d.tcount, d.tstart, nb, d.ntiles);
  __syncthreads();
  block_scan_1024(d.ccount, d.cstart, nb, d.nchunks);
  __syncthreads();
  if (threadIdx.x == 0 && (*d.ntiles > d.maxtiles || *d.nchunks > d.maxchunks)) atomicOr(d.errflag, 4);
}
__global__ void __launch_bounds__(1024) dd_scan_tiles(Dev d) { dd_scan_tiles_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) dd_scan_tiles_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_scan_tiles_body(d, b.bid, b.nblk);
}

__device__ __forceinline__ void dd_tile_write_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  __shared__ unsigned s_lo[3 * kBoxLds], s_hi[3 * kBoxLds];
  const int nb = *d.nblocks;
  if (*d.ntiles > d.maxtiles || *d.nchunks > d.maxchunks) return;
  dd_stage_boxes(d, nb, s_lo, s_hi);
  const int I = (bid * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (I >= nb) return;
  unsigned aLo[3], aHi[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    aLo[k] = d.bblo[3 * I + k];
    aHi[k] = d.bbhi[3 * I + k];
  }
  int base = d.tstart[I];
  const unsigned long long below = (1ull << lane) - 1;
  for (int J0 = 0; J0 < nb; J0 += 64) {
    const int J = J0 + lane;
    const bool keep = J < nb && dd_tile_keep(d, s_lo, s_hi, I, aLo, aHi, J);
    const unsigned long long m = __ballot(keep);
    if (keep) {
      int t = base + __popcll(m & below);
      d.tile_i[t] = I;
      d.tile_j[t] = J;
      d.tile_flag[t] = 0;
    }
    base += __popcll(m);
  }
  const int t0 = d.tstart[I], cnt = d.tcount[I], c0 = d.cstart[I];
  const int tc = d.tchunk;
  for (int k = lane; k * tc < cnt; k += 64) d.chunk[c0 + k] = make_int4(I, t0 + k * tc, min(tc, cnt - k * tc), 0);
}
__global__ void __launch_bounds__(256) dd_tile_write(Dev d) { dd_tile_write_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_tile_write_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_tile_write_body(d, b.bid, b.nblk);
}

// tile (I,J) by binary search in row I's ascending J list, -1 if absent
__device__ __forceinline__ int dd_find_tile(const Dev &d, int I, int J) {
  int lo = d.tstart[I], hi = d.tstart[I + 1];
  const int end = hi;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (d.tile_j[mid] < J) lo = mid + 1; else hi = mid;
  }
  return (lo < end && d.tile_j[lo] == J) ? lo : -1;
}

// exclusion pair -> the tiles (I,J) and (J,I) holding it
__device__ __forceinline__ void dd_excl_mark_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int e = bid * blockDim.x + threadIdx.x;
  if (e >= d.nexcl) return;
  if (*d.ntiles > d.maxtiles) return;
  const int sa = d.g2s[d.excl[2 * e]], sb = d.g2s[d.excl[2 * e + 1]];
  const int I = sa / kBlockAtoms, J = sb / kBlockAtoms;
  const int ta = dd_find_tile(d, I, J);
  const int tb = ta < 0 ? -1 : (I == J ? ta : dd_find_tile(d, J, I));
  d.excl_tile[2 * e] = ta;
  d.excl_tile[2 * e + 1] = tb;
  if (ta >= 0) d.tile_flag[ta] = 1;
  if (tb >= 0) d.tile_flag[tb] = 1;
}
__global__ void __launch_bounds__(256) dd_excl_mark(Dev d) { dd_excl_mark_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_excl_mark_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_excl_mark_body(d, b.bid, b.nblk);
}

// one wave per row: tiles of the row holding exclusions
__device__ __forceinline__ void dd_excl_rows_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int I = (bid * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (I >= *d.nblocks) return;
  int cnt = 0;
  if (*d.ntiles <= d.maxtiles)
    for (int t = d.tstart[I] + lane; t - lane < d.tstart[I + 1]; t += 64)
      cnt += __popcll(__ballot(t < d.tstart[I + 1] && d.tile_flag[t]));
  if (lane == 0) d.ecount[I] = cnt;
}
__global__ void __launch_bounds__(256) dd_excl_rows(Dev d) { dd_excl_rows_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_excl_rows_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_excl_rows_body(d, b.bid, b.nblk);
}

__device__ __forceinline__ void dd_scan_excl_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  block_scan_1024(d.ecount, d.estart, *d.nblocks, d.nexcltiles);
}
__global__ void __launch_bounds__(1024) dd_scan_excl(Dev d) { dd_scan_excl_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(1024) dd_scan_excl_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_scan_excl_body(d, b.bid, b.nblk);
}

// one wave per row: compact mask index of every flagged tile, -1 otherwise
__device__ __forceinline__ void dd_excl_assign_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int I = (bid * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (I >= *d.nblocks || *d.ntiles > d.maxtiles) return;
  int base = d.estart[I];
  const unsigned long long below = (1ull << lane) - 1;
  const int t1 = d.tstart[I + 1];
  for (int t0 = d.tstart[I]; t0 < t1; t0 += 64) {
    const int t = t0 + lane;
    const bool f = t < t1 && d.tile_flag[t];
    const unsigned long long m = __ballot(f);
    if (t < t1) d.tile_ex[t] = f ? base + __popcll(m & below) : -1;
    base += __popcll(m);
  }
}
__global__ void __launch_bounds__(256) dd_excl_assign(Dev d) { dd_excl_assign_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_excl_assign_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_excl_assign_body(d, b.bid, b.nblk);
}

__device__ __forceinline__ void dd_excl_clear_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const long long n = (long long)(*d.nexcltiles) * kBlockAtoms;
  for (long long i = bid * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)nblk * blockDim.x)
    d.emask[i] = 0u;
}
__global__ void __launch_bounds__(256) dd_excl_clear(Dev d) { dd_excl_clear_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_excl_clear_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_excl_clear_body(d, b.bid, b.nblk);
}

// bit j of word i of tile (I,J): slots (I,i) and (J,j) are excluded
__device__ __forceinline__ void dd_excl_set_body(const Dev &d, int bid, int nblk) {
  if (halted_wg(d)) return;   // a halted step leaves the list as it is
  const int e = bid * blockDim.x + threadIdx.x;
  if (e >= d.nexcl) return;
  if (*d.ntiles > d.maxtiles) return;
  const int ta = d.excl_tile[2 * e], tb = d.excl_tile[2 * e + 1];
  if (ta < 0) return;
  const int sa = d.g2s[d.excl[2 * e]], sb = d.g2s[d.excl[2 * e + 1]];
  const int la = sa % kBlockAtoms, lb = sb % kBlockAtoms;
  atomicOr(&d.emask[d.tile_ex[ta] * kBlockAtoms + la], 1u << lb);
  if (tb >= 0) atomicOr(&d.emask[d.tile_ex[tb] * kBlockAtoms + lb], 1u << la);
}
__global__ void __launch_bounds__(256) dd_excl_set(Dev d) { dd_excl_set_body(d, blockIdx.x, gridDim.x); }
__global__ void __launch_bounds__(256) dd_excl_set_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_excl_set_body(d, b.bid, b.nblk);
}

// Every step: slot coordinates, per-block box centre/half-extent for the
// per-step cull, and reset of the force/energy accumulators.
__device__ __forceinline__ void dd_pack_body(const Dev &d, int bid) {
  const int s = bid * blockDim.x + threadIdx.x;
  const int nb = *d.nblocks;
  const int nf = 3 * d.n + d.nl;
  if (s < nf) d.force[s] = 0;
  if (s < d.nl) {
    d.dudl[s] = 0;
    d.dtheta[s] = 0;
  }
  if (s < BM_NTERMS) d.energy[s] = 0;
  if (s == 0) {
    d.workctr[0] = 0;
    d.workctr[1] = 0;
  }
  if (bid * (int)blockDim.x >= nb * kBlockAtoms) return;  // whole wave idle
  const bool in = s < nb * kBlockAtoms;
  const int a = in ? d.l2g[s] : -1;
  unsigned u[3] = {0, 0, 0};
  if (a >= 0) {
    for (int k = 0; k < 3; k++) {
      const double x = d.pos[k * d.n + a];
      nan_check(d, x, a);
      u[k] = frac_u32(x, d.invbox[k]);
    }
    d.spos[s] = make_uint4(u[0], u[1], u[2], __float_as_uint(d.q[a]));
    d.spar[s] = make_int2(d.type[a], d.sb[a]);
    d.slo[s] = (float)(d.qd[a] - (double)d.q[a]);
  } else if (in) {
    d.spos[s] = make_uint4(0, 0, 0, 0);
    d.spar[s] = make_int2(-1, 0);
    d.slo[s] = 0.f;
  }
  // box of the 32-slot block (a half wave); slot 0 of a block is occupied
  unsigned c[3], h[3];
  for (int k = 0; k < 3; k++) {
    unsigned ref = __shfl(u[k], 0, 32);
    int off = (int)(u[k] - ref);
    int mn = a >= 0 ? off : 0x7FFFFFFF, mx = a >= 0 ? off : (int)0x80000000;
    for (int o = 16; o > 0; o >>= 1) {
      mn = min(mn, __shfl_xor(mn, o, 32));
      mx = max(mx, __shfl_xor(mx, o, 32));
    }
    c[k] = ref + (unsigned)((mn + mx) >> 1);
    h[k] = (unsigned)(mx - mn + 1) >> 1;
  }
  int alch = (a >= 0 && d.sb[a] != 0) ? 1 : 0;
  for (int o = 16; o > 0; o >>= 1) alch |= __shfl_xor(alch, o, 32);
  if (in && (s & (kBlockAtoms - 1)) == 0) {
    const int b = s / kBlockAtoms;
    d.bbc[b] = make_uint4(c[0], c[1], c[2], 0);
    d.bbh[b] = make_uint4(h[0], h[1], h[2], (unsigned)alch);
  }
}

__global__ void __launch_bounds__(256) dd_pack(Dev d) { dd_pack_body(d, blockIdx.x); }

__global__ void __launch_bounds__(256) dd_pack_multi(Multi m) {
  const MultiBlk b = multi_block(m);
  const Dev d = m.devs[b.sys];
  dd_pack_body(d, b.bid);
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
void launch_domdec_rebuild(const Dev &d, int maxBlocks, hipStream_t st) {
  const int n = d.n;
  const unsigned slots = (unsigned)((32LL * maxBlocks + 255) / 256);
  const unsigned rows = (unsigned)((maxBlocks + 3) / 4);   // a wave per block
  (void)n;
  dd_zero_columns<<<(d.ncol + 255) / 256, 256, 0, st>>>(d);
  dd_wrap_count<<<slots, 256, 0, st>>>(d);
  dd_scan_columns<<<1, 1024, 0, st>>>(d);
  dd_scatter<<<slots, 256, 0, st>>>(d);
  dd_sort_columns<<<d.ncol, 256, 0, st>>>(d);
  dd_tile_count<<<rows, 256, 0, st>>>(d);
  dd_scan_tiles<<<1, 1024, 0, st>>>(d);
  dd_tile_write<<<rows, 256, 0, st>>>(d);
  if (d.nexcl) dd_excl_mark<<<(d.nexcl + 255) / 256, 256, 0, st>>>(d);
  dd_excl_rows<<<rows, 256, 0, st>>>(d);
  dd_scan_excl<<<1, 1024, 0, st>>>(d);
  dd_excl_assign<<<rows, 256, 0, st>>>(d);
  if (d.nexcl) {
    dd_excl_clear<<<256, 256, 0, st>>>(d);
    dd_excl_set<<<(d.nexcl + 255) / 256, 256, 0, st>>>(d);
  }
}

// The rebuild's kernels in order, for one launch each over several
// systems: workgroups of kernel k for system d (0 leaves d out of that
// launch), and the launches over the per-kernel plans m[kRebuildKernels].
int grid_rebuild(const Dev &d, int k) {
  const int slots = (int)((32LL * d.maxblocks + 255) / 256), rows = (d.maxblocks + 3) / 4;
  switch (k) {
    case 0: return (d.ncol + 255) / 256;                // dd_zero_columns
    case 1: return slots;                               // dd_wrap_count
    case 2: return 1;                                   // dd_scan_columns
    case 3: return slots;                               // dd_scatter
    case 4: return d.ncol;                              // dd_sort_columns
    case 5: return rows;                                // dd_tile_count
    case 6: return 1;                                   // dd_scan_tiles
    case 7: return rows;                                // dd_tile_write
    case 8: return d.nexcl ? (d.nexcl + 255) / 256 : 0; // dd_excl_mark
    case 9: return rows;                                // dd_excl_rows
    case 10: return 1;                                  // dd_scan_excl
    case 11: return rows;                               // dd_excl_assign
    case 12: return d.nexcl ? 256 : 0;                  // dd_excl_clear
    default: return d.nexcl ? (d.nexcl + 255) / 256 : 0;  // dd_excl_set
  }
}

void launch_domdec_rebuild_multi(const Multi *m, hipStream_t st) {
#define BM_RB(K, KER, TH) \
  if (m[K].nwg) KER##_multi<<<m[K].nwg, TH, 0, st>>>(m[K]);
  BM_RB(0, dd_zero_columns, 256)
  BM_RB(1, dd_wrap_count, 256)
  BM_RB(2, dd_scan_columns, 1024)
  BM_RB(3, dd_scatter, 256)
  BM_RB(4, dd_sort_columns, 256)
  BM_RB(5, dd_tile_count, 256)
  BM_RB(6, dd_scan_tiles, 1024)
  BM_RB(7, dd_tile_write, 256)
  BM_RB(8, dd_excl_mark, 256)
  BM_RB(9, dd_excl_rows, 256)
  BM_RB(10, dd_scan_excl, 1024)
  BM_RB(11, dd_excl_assign, 256)
  BM_RB(12, dd_excl_clear, 256)
  BM_RB(13, dd_excl_set, 256)
#undef BM_RB
}

int grid_pack(const Dev &d) {
  long long work = (long long)d.maxblocks * kBlockAtoms;
  long long nf = 3LL * d.n + d.nl;
  if (nf > work) work = nf;
  if (work < BM_NTERMS) work = BM_NTERMS;
  return (int)((work + 255) / 256);
}

void launch_domdec_pack(const Dev &d, int maxBlocks, hipStream_t st) {
  (void)maxBlocks;
  dd_pack<<<grid_pack(d), 256, 0, st>>>(d);
}

void launch_pack_multi(const Multi &m, hipStream_t st) { dd_pack_multi<<<m.nwg, 256, 0, st>>>(m); }

}  // namespace blade
