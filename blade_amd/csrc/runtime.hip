// runtime.hip - device buffers of one system and the bm_* kernel-layer ABI.
//
// bm_system owns every device buffer of one simulation system (the roles of
// the reference's State, Potential, Domdec and the device half of Msld) and
// drives the per-step launch sequence:
//   [rebuild]  domdec kernels               (every freq_domdec steps)
//   pack       slot coordinates, block boxes, zero accumulators
//   nbdirect   tile kernel
//   pme        spread, 5 FFT passes (convolution fused), gather + self
//   bonded     bonds..nbex, lambda biases
//   update     kick/drift, constraints, O/drift, constraints, theta->lambda
// Each (rebuild, energy) step variant can be captured once into a hipGraph
// and replayed; all counts the kernels need live in device memory, so the
// captured graphs need no host input between replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <tuple>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "kernels.h"

// bf16 of a float, rounded to nearest even (finite values), and back
static uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static double bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Slopes dy of the periodic cubic spline through y[0..n) at unit spacing:
// dy[i-1] + 4 dy[i] + dy[i+1] = 3 (y[i+1] - y[i-1]), indices mod n.  The
// cyclic tridiagonal system is solved by the Thomas algorithm on the
// matrix with its corners folded into the diagonal, and the
// Sherman-Morrison correction for the corners.
static void periodic_spline_slopes(int n, const double *y, double *dy) {
  std::vector<double> r(n), u(n, 0.0), cp(n), x(n), z(n);
  for (int i = 0; i < n; i++) r[i] = 3.0 * (y[(i + 1) % n] - y[(i + n - 1) % n]);
  const double gam = -4.0;                       // -diagonal
  auto thomas = [&](const std::vector<double> &rhs, std::vector<double> &out) {
    // diagonal 4, with 4 - gam at 0 and 4 - 1/gam at n-1; off-diagonals 1
    const double b0 = 4.0 - gam;
    out[0] = rhs[0] / b0;
    cp[0] = 1.0 / b0;
    for (int i = 1; i < n; i++) {
      const double bi = (i == n - 1 ? 4.0 - 1.0 / gam : 4.0) - cp[i - 1];
      cp[i] = 1.0 / bi;
      out[i] = (rhs[i] - out[i - 1]) / bi;
    }
    for (int i = n - 2; i >= 0; i--) out[i] -= cp[i] * out[i + 1];
  };
  thomas(r, x);
  u[0] = gam;
  u[n - 1] = 1.0;
  thomas(u, z);
  const double fac = (x[0] + x[n - 1] / gam) / (1.0 + z[0] + z[n - 1] / gam);
  for (int i = 0; i < n; i++) dy[i] = x[i] - fac * z[i];
}

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &m) {
  g_err = m;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return set_err(BM_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));           \
  } while (0)

double bspline_host(int n, double u) {
  if (n == 1) return (u >= 0 && u < 1) ? 1.0 : 0.0;  // M_1 on [0,1)
  if (u <= 0 || u >= n) return 0;
  return (u * bspline_host(n - 1, u) + (n - u) * bspline_host(n - 1, u - 1)) / (n - 1);
}

}  // namespace

struct bm_system {
  int device = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;         // PME branch of the force DAG (default priority)
  hipStream_t st3 = nullptr;         // bonded / dU/dlambda branch (default priority)
  hipEvent_t evFork = nullptr, evJoin = nullptr, evJoin3 = nullptr;
  blade::Dev d{};
  int maxBlocks = 0;
  long long maxTiles = 0;
  bool alch = false, soft = false, bias = false;
  int nwgNb = 0;
  long long maxTiles0 = 0;           // list capacities at creation (growth is bounded by 12x)
  int pcap0 = 0;
  int capacityGen = 0;               // advanced whenever a capacity changes (batches check it)
  int freqDomdec = 10;
  double kT = 0.0;                   // kB T of the thermostat (replica exchange)
  std::vector<void *> allocs;
  hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraph_t graphSrc[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipEvent_t ev[8] = {};
  double msNb = 0, msPme = 0, msTotal = 0;
  long nMeas = 0;

  template <class T>
  int alloc(T *&p, size_t count) {
    if (count == 0) count = 1;
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, count * sizeof(T)));
    HIPCHK(hipMemset(q, 0, count * sizeof(T)));
    allocs.push_back(q);
    p = static_cast<T *>(q);
    return BM_OK;
  }
  // free p (an allocation of this system) and allocate count zeroed items
  template <class T>
  int realloc_dev(T *&p, size_t count) {
    if (p) {
      auto it = std::find(allocs.begin(), allocs.end(), static_cast<void *>(p));
      if (it != allocs.end()) allocs.erase(it);
      (void)hipFree(p);
      p = nullptr;
    }
    return alloc(p, count);
  }
  void drop_graphs() {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        if (graph[a][b]) (void)hipGraphExecDestroy(graph[a][b]);
        if (graphSrc[a][b]) (void)hipGraphDestroy(graphSrc[a][b]);
        graph[a][b] = nullptr;
        graphSrc[a][b] = nullptr;
      }
  }
  template <class T>
  int upload(T *&p, const std::vector<T> &v) {
    int rc = alloc(p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return BM_OK;
  }

  // all force terms in three branches that only read positions: the pair
  // kernel's chunk partials on the main stream, the PME chain (bins,
  // spread, FFT and convolution) on st2, and the bonded slots with the
  // FP64 dU/dlambda of the alchemical atoms on st3 (with the step's
  // normals); then one per-atom assembly into the force array after all
  // three (kick: the assembly also kicks the atoms of a non-energy step).
  // Under stream capture the fork/join becomes three parallel graph
  // branches.  The PME chain is a row of five dependent kernels, so the
  // bonded work is kept off its stream.
  void force_launches(bool energy, bool kick = false) {
    (void)hipEventRecord(evFork, st);
    (void)hipStreamWaitEvent(st2, evFork, 0);
    (void)hipStreamWaitEvent(st3, evFork, 0);
    blade::launch_nbdirect(d, energy, alch, soft, nwgNb, st);
    if (d.elecmethod == BM_PME) blade::launch_pme_grid(d, energy, st2);
    // a non-energy step's O step runs after the assembly: its normals here
    if (kick) blade::launch_normals(d, st3);
    if (alch) blade::launch_nb_dudl(d, soft, st3);
    blade::launch_bonded(d, energy, blade::kFamAll, st3);
    (void)hipEventRecord(evJoin, st2);
    (void)hipEventRecord(evJoin3, st3);
    (void)hipStreamWaitEvent(st, evJoin, 0);
    (void)hipStreamWaitEvent(st, evJoin3, 0);
    blade::launch_force_assemble(d, blade::kAsmNb | blade::kAsmBonded | blade::kAsmPme, st, kick);
  }

  // a step: virtual sites placed (domdec/domdec.cu:215), lists, forces,
  // update.  With virtual sites the sites' forces move to their hosts
  // between the assembly and the kick, so the assembly does not kick.
  void step_launches(bool rebuild, bool energy) {
    const bool virt = d.nvirt2 + d.nvirt3 > 0;
    blade::launch_virtual_place(d, st);
    if (rebuild) blade::launch_domdec_rebuild(d, maxBlocks, st);
    blade::launch_domdec_pack(d, maxBlocks, st);
    force_launches(energy, !energy && !virt);
    blade::launch_virtual_spread(d, st);
    if (energy || virt) blade::launch_update(d, energy, st);
    else blade::launch_update_after_kick(d, st);
  }

  ~bm_system() {
    if (st3) (void)hipStreamSynchronize(st3);
    if (st2) (void)hipStreamSynchronize(st2);
    if (st) (void)hipStreamSynchronize(st);   // no graph may still run
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        if (graph[a][b]) (void)hipGraphExecDestroy(graph[a][b]);
        if (graphSrc[a][b]) (void)hipGraphDestroy(graphSrc[a][b]);
      }
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (evFork) (void)hipEventDestroy(evFork);
    if (evJoin) (void)hipEventDestroy(evJoin);
    if (evJoin3) (void)hipEventDestroy(evJoin3);
    for (void *p : allocs) (void)hipFree(p);
    if (st3) (void)hipStreamDestroy(st3);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

const char *bm_version(void) { return "blade-mi355x 0.1 (gfx950)"; }
const char *bm_last_error(void) { return g_err.c_str(); }

int bm_system_create(int device, const bm_topology *t, const bm_runparams *rp, const double *pos,
                     const double *vel, const double *theta, const double *theta_vel,
                     bm_system **out) {
  using namespace blade;
  if (!t || !rp || !pos || !out) return set_err(BM_EINVAL, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(BM_ENOGPU, "no HIP device");
  if (rp->tile_chunk < 0 || rp->tile_chunk > kTileChunk) return set_err(BM_EINVAL, "tile_chunk 0..16");
  if (device < 0 || device >= ndev) return set_err(BM_EINVAL, "device out of range");
  // pairs interact through their nearest image; the list search compares
  // periodic gaps between block intervals over every block pair, so it
  // holds below twice the cutoff too (the reference runs its 18.856 A
  // water box at its default 10 A cutoff)
  for (int k = 0; k < 3; k++)
    if (!(rp->box[k] > rp->rcut)) return set_err(BM_EINVAL, "box edge must exceed rcut");
  if (rp->elec_method == BM_PME) {
    for (int k = 0; k < 3; k++)
      if (!pme_grid_supported(rp->grid[k]))
        return set_err(BM_ENOTSUP, "PME grid needs factors 2,3,5 and at most 1024 points per axis");
    if (rp->order != 4 && rp->order != 6 && rp->order != 8) return set_err(BM_ENOTSUP, "PME order 4/6/8");
    if (rp->fft_mode != 0 && rp->fft_mode != 1) return set_err(BM_EINVAL, "fft_mode 0 or 1");
    if (rp->fft_mode == 1)
      for (int k = 0; k < 3; k++)
        if (!fft_mfma_supported(rp->grid[k]))
          return set_err(BM_ENOTSUP, "the MFMA FFT needs grid edges that are multiples of 16 from 16 to 128");
  }
  HIPCHK(hipSetDevice(device));
  bm_system *s = new bm_system();
  s->device = device;
  auto fail = [&](int rc) {
    delete s;
    return rc;
  };
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->st2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->st3, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->evFork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->evJoin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->evJoin3, hipEventDisableTiming) != hipSuccess)
    return fail(set_err(BM_EHIP, "stream"));
  for (auto &e : s->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(set_err(BM_EHIP, "event"));
  Dev &d = s->d;
  const int n = t->n_atoms, nl = std::max(1, t->n_blocks);
  d.n = n;
  d.nl = nl;
  d.nsites = std::max(1, t->n_sites);
  for (int k = 0; k < 3; k++) {
    d.box[k] = rp->box[k];
    d.invbox[k] = 1.0 / rp->box[k];
    d.boxf[k] = (float)rp->box[k];
    d.fracscale[k] = (float)(rp->box[k] / 4294967296.0);
    d.ddscale[k] = rp->box[k] * (1.0 / 4294967296.0);
  }
  const double rl = rp->rcut + rp->cull_pad;
  d.rlist2 = rl * rl;
  s->alch = t->n_blocks > 1;
  s->soft = s->alch && rp->soft_core;
  s->bias = s->alch;
  s->freqDomdec = rp->freq_domdec > 0 ? rp->freq_domdec : 10;
  s->kT = kBoltzD * rp->temperature;
  // columns of about 32 atoms' volume (domdec/domdec.cu:110-118)
  {
    double V = rp->box[0] * rp->box[1] * rp->box[2];
    double a = std::cbrt(32.0 * V / n);
    d.ncx = std::max(1, (int)std::ceil(rp->box[0] / a));
    d.ncy = std::max(1, (int)std::ceil(rp->box[1] / a));
    d.ncol = d.ncx * d.ncy;
  }
  s->maxBlocks = n / kBlockAtoms + 1 + d.ncol;
  d.maxblocks = s->maxBlocks;
  // directed tiles: blocks within rcut+pad of a block (its box edge about
  // the column width a) fill a sphere of radius rcut+pad+sqrt(3)a; 1.5x
  // margin, at most every pair of blocks
  {
    const double V = rp->box[0] * rp->box[1] * rp->box[2];
    const double a = std::cbrt(32.0 * V / n);
    const double R = rl + std::sqrt(3.0) * a;
    const double perRow = 1.5 * (4.0 / 3.0) * M_PI * R * R * R / (32.0 * V / n) + 16;
    const long long MBl = s->maxBlocks;
    s->maxTiles = std::min(MBl * MBl, MBl * (long long)std::ceil(perRow));
  }
  d.maxtiles = s->maxTiles;
  s->maxTiles0 = s->maxTiles;
  d.tchunk = rp->tile_chunk > 0 ? rp->tile_chunk : kTileChunk;
  d.maxchunks = (int)(s->maxTiles / d.tchunk + s->maxBlocks + 1);
  const int MB = s->maxBlocks;
  int rc = 0;
#define ALLOC(p, c) \
  if ((rc = s->alloc(p, (c)))) return fail(rc)
#define UPLOAD(p, v) \
  if ((rc = s->upload(p, v))) return fail(rc)
  const size_t N = 3 * (size_t)n + nl;
  // state
  {
    std::vector<double> x(N, 0.0), v(N, 0.0), ism(N, 0.0);
    for (size_t i = 0; i < 3 * (size_t)n; i++) {
      x[i] = pos[i];
      v[i] = vel ? vel[i] : 0.0;
    }
    for (int i = 0; i < n; i++) {
      double m = t->mass[i];
      double w = m > 0 ? 1.0 / std::sqrt(m) : 0.0;
      ism[i] = ism[n + i] = ism[2 * (size_t)n + i] = w;
    }
    for (int b = 0; b < t->n_blocks; b++) {
      x[3 * (size_t)n + b] = theta ? theta[b] : 0.0;
      v[3 * (size_t)n + b] = theta_vel ? theta_vel[b] : 0.0;
      double m = t->theta_mass ? t->theta_mass[b] : 0.0;
      ism[3 * (size_t)n + b] = m > 0 ? 1.0 / std::sqrt(m) : 0.0;
    }
    UPLOAD(d.pos, x);
    UPLOAD(d.vel, v);
    UPLOAD(d.ism, ism);
    ALLOC(d.posb, N);
    ALLOC(d.xi, N);
  }
  ALLOC(d.force, N);
  ALLOC(d.dudl, nl);
  ALLOC(d.energy, BM_NTERMS);
  {
    // with the fix flag lambda is the initial theta (msld.cu:924-934)
    std::vector<double> lam(nl, 1.0);
    if (t->fix_lambda && theta)
      for (int b = 0; b < t->n_blocks; b++) lam[b] = theta[b];
    UPLOAD(d.lambda, lam);
    std::vector<float> lamf(nl, 1.0f);
    for (int b = 0; b < nl; b++) lamf[b] = (float)lam[b];
    UPLOAD(d.lambdaf, lamf);
    std::vector<int> nf(1, -1);
    UPLOAD(d.nanflag, nf);
    std::vector<long long> st0(1, 0);
    UPLOAD(d.step, st0);
  }
  // per-atom nonbonded data
  {
    std::vector<float> q(n);
    std::vector<int> ty(n), sb(n);
    std::vector<double> ms(n);
    for (int i = 0; i < n; i++) {
      q[i] = (float)t->charge[i];
      ty[i] = t->type_idx[i];
      sb[i] = t->site_block ? t->site_block[i] : 0;
      ms[i] = t->mass[i];
    }
    UPLOAD(d.q, q);
    std::vector<double> qd(t->charge, t->charge + n);
    UPLOAD(d.qd, qd);
    std::vector<double> bq2(nl, 0.0);
    for (int i = 0; i < n; i++) bq2[(t->site_block ? t->site_block[i] : 0) & 0xFFFF] += qd[i] * qd[i];
    UPLOAD(d.blockq2, bq2);
    UPLOAD(d.type, ty);
    UPLOAD(d.sb, sb);
    std::vector<int> al;
    for (int i = 0; i < n; i++)
      if (sb[i]) al.push_back(i);
    d.nalchat = (int)al.size();
    UPLOAD(d.alchat, al);
    UPLOAD(d.mass, ms);
    ALLOC(d.qlam, n);
    std::vector<float2> vdw((size_t)t->n_types * t->n_types);
    for (size_t k = 0; k < vdw.size(); k++) vdw[k] = make_float2((float)t->vdw_c12[k], (float)t->vdw_c6[k]);
    UPLOAD(d.vdw, vdw);
    std::vector<double2> vdwd(vdw.size());
    for (size_t k = 0; k < vdw.size(); k++) vdwd[k] = make_double2(t->vdw_c12[k], t->vdw_c6[k]);
    UPLOAD(d.vdwd, vdwd);
    d.ntypes = t->n_types;
  }
  d.rc = (float)rp->rcut;
  d.rs = (float)rp->rswitch;
  d.rc2 = d.rc * d.rc;
  d.beta = (float)rp->beta_ewald;
  d.rcd = rp->rcut;
  d.rsd = rp->rswitch;
  d.betad = rp->beta_ewald;
  d.vdwmethod = rp->vdw_method;
  d.elecmethod = rp->elec_method;
  // domdec buffers
  ALLOC(d.ufrac, 3 * (size_t)n);
  ALLOC(d.colcount, d.ncol);
  ALLOC(d.colstart, d.ncol + 1);
  ALLOC(d.colfill, d.ncol);
  ALLOC(d.bucket, std::max(n, d.ncol));
  ALLOC(d.blkstart, d.ncol + 1);
  ALLOC(d.nblocks, 1);
  {
    // the first rebuild walks atoms in index order; later ones walk the
    // previous rebuild's slots
    std::vector<int> id((size_t)MB * kBlockAtoms, -1);
    for (int i = 0; i < n; i++) id[i] = i;
    UPLOAD(d.l2g, id);
  }
  ALLOC(d.g2s, n);
  ALLOC(d.bblo, 3 * (size_t)MB);
  ALLOC(d.bbhi, 3 * (size_t)MB);
  ALLOC(d.bbc, MB);
  ALLOC(d.bbh, MB);
  ALLOC(d.spos, (size_t)MB * kBlockAtoms);
  ALLOC(d.slo, (size_t)MB * kBlockAtoms);
  ALLOC(d.spar, (size_t)MB * kBlockAtoms);
  ALLOC(d.tcount, MB);
  ALLOC(d.tstart, MB + 1);
  ALLOC(d.ntiles, 1);
  // +1: the exclusive scans also write the total at index n
  ALLOC(d.tile_i, s->maxTiles + 1);
  ALLOC(d.tile_j, s->maxTiles + 1);
  ALLOC(d.tile_ex, s->maxTiles + 1);
  ALLOC(d.tile_flag, s->maxTiles + 1);
  ALLOC(d.ccount, MB);
  ALLOC(d.cstart, MB + 1);
  ALLOC(d.nchunks, 1);
  ALLOC(d.chunk, d.maxchunks);
  ALLOC(d.ichunk, (size_t)d.maxchunks * 3 * kBlockAtoms);
  ALLOC(d.ecount, MB);
  ALLOC(d.estart, MB + 1);
  ALLOC(d.nexcltiles, 1);
  d.nexcl = t->n_excl;
  ALLOC(d.emask, 2 * (size_t)std::max(1, t->n_excl) * kBlockAtoms);
  ALLOC(d.excl_tile, 2 * (size_t)t->n_excl);
  {
    std::vector<int> ex(t->excl_idx, t->excl_idx + 2 * (size_t)t->n_excl);
    UPLOAD(d.excl, ex);
  }
  ALLOC(d.workctr, 4);
  ALLOC(d.errflag, 1);
  // PME
  if (rp->elec_method == BM_PME) {
    for (int k = 0; k < 3; k++) d.K[k] = rp->grid[k];
    d.order = rp->order;
    const size_t G = (size_t)d.K[0] * d.K[1] * d.K[2];
    ALLOC(d.rgrid, G);
    ALLOC(d.cgrid, G);
    // spread bricks: floor(K/8) per axis (at least one), so each spans >= 8
    // points (>= order-1 for orders up to 8, or the whole axis); a 64^3 grid
    // gives 512 bricks, one workgroup each, all resident at once
    int nbr = 1;
    for (int k = 0; k < 3; k++) {
      d.pnb[k] = std::max(1, d.K[k] / 8);
      nbr *= d.pnb[k];
    }
    d.pcap = 64 + 4 * ((n + nbr - 1) / nbr);
    s->pcap0 = d.pcap;
    ALLOC(d.pbcount, nbr);
    ALLOC(d.pbrec, (size_t)nbr * d.pcap * pme_spline_rec_words(d.order));
    ALLOC(d.pbhdr, (size_t)nbr * d.pcap);
    std::vector<float2> tw;
    for (int k = 0; k < 3; k++)
      for (int m = 0; m < d.K[k]; m++) {
        double a = -2.0 * M_PI * m / d.K[k];
        tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
      }
    UPLOAD(d.twiddle, tw);
    d.fft_mode = rp->fft_mode;
    if (d.fft_mode == 1) {
      // the real-form forward DFT matrix [[C, S], [-S, C]] per axis, split
      // into bf16 hi + lo from FP64, in the order the MFMA A operand reads
      // it: fragment (row tile rt, k-step ks), lane l, element j holds
      // M[32 rt + l%32][16 ks + 8(l/32) + j]; the kernels negate the
      // off-diagonal blocks for the inverse
      for (int ax = 0; ax < 3; ax++) {
        const int K = d.K[ax], KS = 2 * K / 16, RT = 2 * K / 32;
        std::vector<uint16_t> hi((size_t)4 * K * K), lo((size_t)4 * K * K);
        for (int rt = 0; rt < RT; rt++)
          for (int ks = 0; ks < KS; ks++)
            for (int l = 0; l < 64; l++)
              for (int j = 0; j < 8; j++) {
                const int m = 32 * rt + (l & 31), kk = 16 * ks + 8 * (l >> 5) + j;
                const double th = 2.0 * M_PI * (double)((long long)(m % K) * (kk % K) % K) / K;
                const double c = std::cos(th), sn = std::sin(th);
                const double v = (m < K) == (kk < K) ? c : (m < K ? sn : -sn);
                const uint16_t h = bf16_rne((float)v);
                const size_t at = (((size_t)rt * KS + ks) * 64 + l) * 8 + j;
                hi[at] = h;
                lo[at] = bf16_rne((float)(v - bf16_value(h)));
              }
        uint16_t *ph = nullptr, *pl = nullptr;
        UPLOAD(ph, hi);
        UPLOAD(pl, lo);
        d.mdft[ax][0] = ph;
        d.mdft[ax][1] = pl;
      }
    }
    for (int k = 0; k < 3; k++) {
      // 1/|b(m)|^2 from the integer-point B-spline values (potential.cxx:1245-1287)
      std::vector<float> bm(d.K[k]);
      std::vector<double> bmd(d.K[k]);
      for (int i = 0; i < d.K[k]; i++) {
        double re = 0, im = 0;
        for (int l = 1; l < d.order; l++) {
          double mv = bspline_host(d.order, l);
          re += mv * std::cos(2 * M_PI * i * l / d.K[k]);
          im += mv * std::sin(2 * M_PI * i * l / d.K[k]);
        }
        bmd[i] = 1.0 / (re * re + im * im);
        bm[i] = (float)bmd[i];
      }
      UPLOAD(d.bmod[k], bm);
      UPLOAD(d.bmodd[k], bmd);
    }
    const double V = rp->box[0] * rp->box[1] * rp->box[2];
    d.pmefacd = 0.5 * kElectricD / M_PI / V;
    d.pmeexpd = M_PI * M_PI / (rp->beta_ewald * rp->beta_ewald);
    d.pmefac = (float)d.pmefacd;
    d.pmeexp = (float)d.pmeexpd;
  }
  // bonded terms
  {
    d.nbond = t->n_bonds;
    d.nbond12 = t->n_bonds12;
    std::vector<int2> bi(d.nbond), bsb(d.nbond);
    std::vector<double2> bp(d.nbond);
    for (int b = 0; b < d.nbond; b++) {
      bi[b] = make_int2(t->bond_idx[2 * b], t->bond_idx[2 * b + 1]);
      bsb[b] = make_int2(t->bond_sb[2 * b], t->bond_sb[2 * b + 1]);
      bp[b] = make_double2(t->bond_kb[b], t->bond_b0[b]);
    }
    UPLOAD(d.bond_i, bi); UPLOAD(d.bond_sb, bsb); UPLOAD(d.bond_p, bp);
    d.sb_alpha = rp->soft_bond_alpha;
    d.sb_pow = rp->soft_bond_power;
    d.snb_exp = rp->soft_bond_exp;
    d.bond_soft = d.angle_soft = nullptr;
    if (t->bond_soft) {
      std::vector<int> f(t->bond_soft, t->bond_soft + d.nbond);
      UPLOAD(d.bond_soft, f);
    }
    if (t->angle_soft) {
      std::vector<int> f(t->angle_soft, t->angle_soft + t->n_angles);
      UPLOAD(d.angle_soft, f);
    }
    d.nangle = t->n_angles;
    std::vector<int4> ai(d.nangle);
    std::vector<int2> asb(d.nangle);
    std::vector<double2> ap(d.nangle);
    for (int a = 0; a < d.nangle; a++) {
      ai[a] = make_int4(t->angle_idx[3 * a], t->angle_idx[3 * a + 1], t->angle_idx[3 * a + 2], 0);
      asb[a] = make_int2(t->angle_sb[2 * a], t->angle_sb[2 * a + 1]);
      ap[a] = make_double2(t->angle_k[a], t->angle_t0[a]);
    }
    UPLOAD(d.angle_i, ai); UPLOAD(d.angle_sb, asb); UPLOAD(d.angle_p, ap);
    for (int pass = 0; pass < 2; pass++) {
      int cnt = pass ? t->n_imprs : t->n_dihes;
      const int *idx = pass ? t->impr_idx : t->dihe_idx;
      const int *sbv = pass ? t->impr_sb : t->dihe_sb;
      const double *kk = pass ? t->impr_k : t->dihe_k;
      const int *nn = pass ? t->impr_n : t->dihe_n;
      const double *d0 = pass ? t->impr_d0 : t->dihe_d0;
      const int *sf = pass ? t->impr_soft : t->dihe_soft;
      std::vector<int4> di(cnt);
      std::vector<int2> dsb(cnt);
      std::vector<double4> dp(cnt);
      for (int q = 0; q < cnt; q++) {
        di[q] = make_int4(idx[4 * q], idx[4 * q + 1], idx[4 * q + 2], idx[4 * q + 3]);
        dsb[q] = make_int2(sbv[2 * q], sbv[2 * q + 1]);
        dp[q] = make_double4(kk[q], (double)nn[q], d0[q], sf && sf[q] ? 1.0 : 0.0);
      }
      if (pass) { d.nimpr = cnt; UPLOAD(d.impr_i, di); UPLOAD(d.impr_sb, dsb); UPLOAD(d.impr_p, dp); }
      else { d.ndihe = cnt; UPLOAD(d.dihe_i, di); UPLOAD(d.dihe_sb, dsb); UPLOAD(d.dihe_p, dp); }
    }
    d.nnb14 = t->n_nb14;
    std::vector<int2> pi(d.nnb14), psb(d.nnb14);
    std::vector<double4> pp(d.nnb14);
    for (int q = 0; q < d.nnb14; q++) {
      pi[q] = make_int2(t->nb14_idx[2 * q], t->nb14_idx[2 * q + 1]);
      psb[q] = make_int2(t->nb14_sb[2 * q], t->nb14_sb[2 * q + 1]);
      pp[q] = make_double4(t->nb14_qq[q], t->nb14_c12[q], t->nb14_c6[q], t->nb14_e14fac[q]);
    }
    UPLOAD(d.nb14_i, pi); UPLOAD(d.nb14_sb, psb); UPLOAD(d.nb14_p, pp);
    d.nnbex = rp->elec_method == BM_PME ? t->n_nbex : 0;
    std::vector<int2> xi(d.nnbex), xsb(d.nnbex);
    std::vector<double> xp(d.nnbex);
    for (int q = 0; q < d.nnbex; q++) {
      xi[q] = make_int2(t->nbex_idx[2 * q], t->nbex_idx[2 * q + 1]);
      xsb[q] = make_int2(t->nbex_sb[2 * q], t->nbex_sb[2 * q + 1]);
      xp[q] = t->nbex_qq[q];
    }
    UPLOAD(d.nbex_i, xi); UPLOAD(d.nbex_sb, xsb); UPLOAD(d.nbex_p, xp);
    // atom restraint groups
    d.ncats = t->n_cats;
    const int ncm = d.ncats ? t->cats_bound[d.ncats] : 0;
    {
      std::vector<int> cb(t->cats_bound, t->cats_bound + (d.ncats ? d.ncats + 1 : 0));
      std::vector<int> ci(t->cats_idx, t->cats_idx + ncm);
      if (d.ncats) { UPLOAD(d.cats_bound, cb); UPLOAD(d.cats_idx, ci); }
      d.cats_k = t->cats_k;
    }
    // CMAP grids: the slopes of periodic cubic splines through the grid
    // values (potential.cxx:460-520) give each point its value, both first
    // derivatives and the cross derivative, for bicubic Hermite cells
    d.ncmap = t->n_cmaps;
    if (d.ncmap) {
      std::vector<int4> ci(2 * (size_t)d.ncmap), cs(d.ncmap);
      for (int q = 0; q < d.ncmap; q++) {
        const int32_t *a = t->cmap_idx + 8 * q;
        ci[2 * q] = make_int4(a[0], a[1], a[2], a[3]);
        ci[2 * q + 1] = make_int4(a[4], a[5], a[6], a[7]);
        cs[q] = make_int4(t->cmap_sb[3 * q], t->cmap_sb[3 * q + 1], t->cmap_sb[3 * q + 2], t->cmap_type[q]);
      }
      std::vector<int2> cg(t->n_cmap_types);
      std::vector<double4> tab;
      for (int g = 0; g < t->n_cmap_types; g++) {
        const int ng = t->cmap_ngrid[g];
        const double *v = t->cmap_grid + t->cmap_off[g];
        cg[g] = make_int2(ng, (int)tab.size());
        std::vector<double> fx((size_t)ng * ng), fy((size_t)ng * ng), fxy((size_t)ng * ng), y(ng), dy(ng);
        for (int j = 0; j < ng; j++) {
          for (int i = 0; i < ng; i++) y[i] = v[(size_t)ng * i + j];
          periodic_spline_slopes(ng, y.data(), dy.data());
          for (int i = 0; i < ng; i++) fx[(size_t)ng * i + j] = dy[i];
        }
        for (int i = 0; i < ng; i++) {
          periodic_spline_slopes(ng, v + (size_t)ng * i, fy.data() + (size_t)ng * i);
          periodic_spline_slopes(ng, fx.data() + (size_t)ng * i, fxy.data() + (size_t)ng * i);
        }
        for (size_t p = 0; p < (size_t)ng * ng; p++) tab.push_back(make_double4(v[p], fx[p], fy[p], fxy[p]));
      }
      UPLOAD(d.cmap_i, ci); UPLOAD(d.cmap_s, cs); UPLOAD(d.cmap_g, cg); UPLOAD(d.cmap_tab, tab);
    }
    // restraints
    {
      d.nnoe = t->n_noe; d.nharm = t->n_harm; d.nbor = t->n_borest;
      d.nanr = t->n_anrest; d.ndir = t->n_direst; d.nresd = t->n_resd;
      std::vector<int2> ni(d.nnoe);
      for (int q = 0; q < d.nnoe; q++) ni[q] = make_int2(t->noe_idx[2 * q], t->noe_idx[2 * q + 1]);
      std::vector<double> np(t->noe_par, t->noe_par + 10 * (size_t)d.nnoe);
      std::vector<int> hi(t->harm_idx, t->harm_idx + d.nharm);
      std::vector<double> hp(t->harm_par, t->harm_par + 5 * (size_t)d.nharm);
      std::vector<int4> bi(d.nbor), ai(d.nanr), di(d.ndir), ri(d.nresd);
      std::vector<double2> bp(d.nbor), ap(d.nanr);
      std::vector<int> ab(d.nanr);
      std::vector<int2> dbn(d.ndir);
      std::vector<double4> dp(d.ndir), rpar(d.nresd);
      for (int q = 0; q < d.nbor; q++) {
        bi[q] = make_int4(t->borest_idx[2 * q], t->borest_idx[2 * q + 1], t->borest_block[q], 0);
        bp[q] = make_double2(t->borest_par[2 * q], t->borest_par[2 * q + 1]);
      }
      for (int q = 0; q < d.nanr; q++) {
        ai[q] = make_int4(t->anrest_idx[3 * q], t->anrest_idx[3 * q + 1], t->anrest_idx[3 * q + 2], 0);
        ab[q] = t->anrest_block[q];
        ap[q] = make_double2(t->anrest_par[2 * q], t->anrest_par[2 * q + 1]);
      }
      for (int q = 0; q < d.ndir; q++) {
        const int32_t *a = t->direst_idx + 4 * q;
        di[q] = make_int4(a[0], a[1], a[2], a[3]);
        dbn[q] = make_int2(t->direst_block[q], t->direst_n[q]);
        dp[q] = make_double4(t->direst_par[3 * q], t->direst_par[3 * q + 1], t->direst_par[3 * q + 2], 0.0);
      }
      for (int q = 0; q < d.nresd; q++) {
        const int32_t *a = t->resd_idx + 4 * q;
        const double *p = t->resd_par + 4 * q;
        ri[q] = make_int4(a[0], a[1], a[2], a[3]);
        rpar[q] = make_double4(p[0], p[1], p[2], p[3]);
      }
      UPLOAD(d.noe_i, ni); UPLOAD(d.noe_p, np); UPLOAD(d.harm_i, hi); UPLOAD(d.harm_p, hp);
      UPLOAD(d.bor_i, bi); UPLOAD(d.bor_p, bp); UPLOAD(d.anr_i, ai); UPLOAD(d.anr_b, ab); UPLOAD(d.anr_p, ap);
      UPLOAD(d.dir_i, di); UPLOAD(d.dir_bn, dbn); UPLOAD(d.dir_p, dp); UPLOAD(d.resd_i, ri); UPLOAD(d.resd_p, rpar);
    }
    // bonded force slots and the atom -> slot CSR (the topology is static)
    d.sl_bond = 0;
    d.sl_angle = d.sl_bond + 2 * d.nbond;
    d.sl_dihe = d.sl_angle + 3 * d.nangle;
    d.sl_impr = d.sl_dihe + 4 * d.ndihe;
    d.sl_nb14 = d.sl_impr + 4 * d.nimpr;
    d.sl_nbex = d.sl_nb14 + 2 * d.nnb14;
    d.sl_cats = d.sl_nbex + 2 * d.nnbex;
    d.sl_cmap = d.sl_cats + ncm;
    d.sl_noe = d.sl_cmap + 8 * d.ncmap;
    d.sl_harm = d.sl_noe + 2 * d.nnoe;
    d.sl_bor = d.sl_harm + d.nharm;
    d.sl_anr = d.sl_bor + 2 * d.nbor;
    d.sl_dir = d.sl_anr + 3 * d.nanr;
    d.sl_resd = d.sl_dir + 4 * d.ndir;
    d.nslot = d.sl_resd + 4 * d.nresd;
    std::vector<int> owner(d.nslot);
    for (int q = 0; q < d.nbond; q++)
      for (int k = 0; k < 2; k++) owner[d.sl_bond + 2 * q + k] = t->bond_idx[2 * q + k];
    for (int q = 0; q < d.nangle; q++)
      for (int k = 0; k < 3; k++) owner[d.sl_angle + 3 * q + k] = t->angle_idx[3 * q + k];
    for (int q = 0; q < d.ndihe; q++)
      for (int k = 0; k < 4; k++) owner[d.sl_dihe + 4 * q + k] = t->dihe_idx[4 * q + k];
    for (int q = 0; q < d.nimpr; q++)
      for (int k = 0; k < 4; k++) owner[d.sl_impr + 4 * q + k] = t->impr_idx[4 * q + k];
    for (int q = 0; q < d.nnb14; q++)
      for (int k = 0; k < 2; k++) owner[d.sl_nb14 + 2 * q + k] = t->nb14_idx[2 * q + k];
    for (int q = 0; q < d.nnbex; q++)
      for (int k = 0; k < 2; k++) owner[d.sl_nbex + 2 * q + k] = t->nbex_idx[2 * q + k];
    for (int q = 0; q < ncm; q++) owner[d.sl_cats + q] = t->cats_idx[q];
    for (int q = 0; q < 8 * d.ncmap; q++) owner[d.sl_cmap + q] = t->cmap_idx[q];
    for (int q = 0; q < d.nnoe; q++) {   // a point NOE's second slot stays zero, owned by atom i
      owner[d.sl_noe + 2 * q] = t->noe_idx[2 * q];
      owner[d.sl_noe + 2 * q + 1] = t->noe_idx[2 * q + 1] >= 0 ? t->noe_idx[2 * q + 1] : t->noe_idx[2 * q];
    }
    for (int q = 0; q < d.nharm; q++) owner[d.sl_harm + q] = t->harm_idx[q];
    for (int q = 0; q < 2 * d.nbor; q++) owner[d.sl_bor + q] = t->borest_idx[q];
    for (int q = 0; q < 3 * d.nanr; q++) owner[d.sl_anr + q] = t->anrest_idx[q];
    for (int q = 0; q < 4 * d.ndir; q++) owner[d.sl_dir + q] = t->direst_idx[q];
    for (int q = 0; q < 4 * d.nresd; q++) owner[d.sl_resd + q] = t->resd_idx[q];
    std::vector<int> cstartv(n + 1, 0), pos(d.nslot);
    for (int q = 0; q < d.nslot; q++) cstartv[owner[q] + 1]++;
    for (int a = 0; a < n; a++) cstartv[a + 1] += cstartv[a];
    {
      std::vector<int> fill(cstartv.begin(), cstartv.end() - 1);
      for (int q = 0; q < d.nslot; q++) pos[q] = fill[owner[q]]++;
    }
    ALLOC(d.bslot, 3 * (size_t)d.nslot);
    UPLOAD(d.bcsr_start, cstartv);
    UPLOAD(d.bspos, pos);
  }
  // constraints
  {
    d.ntri = t->n_tri; d.nbr1 = t->n_br1; d.nbr2 = t->n_br2; d.nbr3 = t->n_br3;
    std::vector<int4> ti(d.ntri), b2i(d.nbr2), b3i(d.nbr3);
    std::vector<double4> tb(d.ntri), b3b(d.nbr3);
    std::vector<int2> b1i(d.nbr1);
    std::vector<double> b1b(d.nbr1);
    std::vector<double2> b2b(d.nbr2);
    for (int c = 0; c < d.ntri; c++) {
      ti[c] = make_int4(t->tri_idx[3 * c], t->tri_idx[3 * c + 1], t->tri_idx[3 * c + 2], 0);
      tb[c] = make_double4(t->tri_b0[3 * c], t->tri_b0[3 * c + 1], t->tri_b0[3 * c + 2], 0);
    }
    for (int c = 0; c < d.nbr1; c++) {
      b1i[c] = make_int2(t->br1_idx[2 * c], t->br1_idx[2 * c + 1]);
      b1b[c] = t->br1_b0[c];
    }
    for (int c = 0; c < d.nbr2; c++) {
      b2i[c] = make_int4(t->br2_idx[3 * c], t->br2_idx[3 * c + 1], t->br2_idx[3 * c + 2], 0);
      b2b[c] = make_double2(t->br2_b0[2 * c], t->br2_b0[2 * c + 1]);
    }
    for (int c = 0; c < d.nbr3; c++) {
      b3i[c] = make_int4(t->br3_idx[4 * c], t->br3_idx[4 * c + 1], t->br3_idx[4 * c + 2], t->br3_idx[4 * c + 3]);
      b3b[c] = make_double4(t->br3_b0[3 * c], t->br3_b0[3 * c + 1], t->br3_b0[3 * c + 2], 0);
    }
    UPLOAD(d.tri_i, ti); UPLOAD(d.tri_b, tb);
    UPLOAD(d.br1_i, b1i); UPLOAD(d.br1_b, b1b);
    UPLOAD(d.br2_i, b2i); UPLOAD(d.br2_b, b2b);
    UPLOAD(d.br3_i, b3i); UPLOAD(d.br3_b, b3b);
    // atoms in no group; the step kernels rely on every atom being in at
    // most one group (they move a group's atoms in one thread)
    std::vector<int> owner(n, 0);
    auto own = [&](int a) { return ++owner[a] == 1; };
    bool disjoint = true;
    for (int c = 0; c < d.ntri; c++)
      for (int k = 0; k < 3; k++) disjoint &= own(t->tri_idx[3 * c + k]);
    for (int c = 0; c < d.nbr1; c++)
      for (int k = 0; k < 2; k++) disjoint &= own(t->br1_idx[2 * c + k]);
    for (int c = 0; c < d.nbr2; c++)
      for (int k = 0; k < 3; k++) disjoint &= own(t->br2_idx[3 * c + k]);
    for (int c = 0; c < d.nbr3; c++)
      for (int k = 0; k < 4; k++) disjoint &= own(t->br3_idx[4 * c + k]);
    if (!disjoint) return fail(set_err(BM_EINVAL, "an atom belongs to two constraint groups"));
    std::vector<int> fr;
    for (int a = 0; a < n; a++)
      if (!owner[a]) fr.push_back(a);
    d.nfree = (int)fr.size();
    UPLOAD(d.freeatom, fr);
    // virtual sites: a site belongs to no constraint group and hosts no
    // other site (the spreading kernel reads site forces while it adds to
    // host forces)
    d.nvirt2 = t->n_virt2;
    d.nvirt3 = t->n_virt3;
    std::vector<int4> v2i(d.nvirt2), v3i(d.nvirt3);
    std::vector<double2> v2p(d.nvirt2);
    std::vector<double4> v3p(2 * (size_t)d.nvirt3);
    std::vector<char> site(n, 0);
    auto in_range = [&](int a) { return a >= 0 && a < n; };
    bool ok = true;
    for (int k = 0; k < d.nvirt2; k++) {
      const int *ix = t->virt2_idx + 3 * k;
      for (int j = 0; j < 3; j++) ok &= in_range(ix[j]);
      if (!ok) break;
      ok &= !owner[ix[0]] && !site[ix[0]] && ix[0] != ix[1] && ix[0] != ix[2] && ix[1] != ix[2];
      site[ix[0]] = 1;
      v2i[k] = make_int4(ix[0], ix[1], ix[2], 0);
      v2p[k] = make_double2(t->virt2_par[2 * k], t->virt2_par[2 * k + 1]);
    }
    for (int k = 0; ok && k < d.nvirt3; k++) {
      const int *ix = t->virt3_idx + 4 * k;
      for (int j = 0; j < 4; j++) ok &= in_range(ix[j]);
      if (!ok) break;
      ok &= !owner[ix[0]] && !site[ix[0]] && ix[1] != ix[2] && ix[1] != ix[3] && ix[2] != ix[3] &&
            ix[0] != ix[1] && ix[0] != ix[2] && ix[0] != ix[3];
      site[ix[0]] = 1;
      const double *p = t->virt3_par + 3 * k;
      v3i[k] = make_int4(ix[0], ix[1], ix[2], ix[3]);
      v3p[2 * k] = make_double4(p[0], std::cos(p[1]), std::sin(p[1]), 0);
      v3p[2 * k + 1] = make_double4(std::cos(p[2]), std::sin(p[2]), 0, 0);
    }
    for (int k = 0; ok && k < d.nvirt2; k++) ok &= !site[v2i[k].y] && !site[v2i[k].z];
    for (int k = 0; ok && k < d.nvirt3; k++) ok &= !site[v3i[k].y] && !site[v3i[k].z] && !site[v3i[k].w];
    if (!ok)
      return fail(set_err(BM_EINVAL, "virtual site indices out of range, repeated, constrained or hosted by a site"));
    UPLOAD(d.virt2_i, v2i); UPLOAD(d.virt2_p, v2p);
    UPLOAD(d.virt3_i, v3i); UPLOAD(d.virt3_p, v3p);
  }
  // msld
  {
    std::vector<int> bs(nl, 0), sbd(d.nsites + 1, 0), bf(nl, 0);
    std::vector<double> fb(nl, 0.0), lf(nl), ln(nl);
    const double kT = kBoltzD * rp->temperature;
    for (int b = 0; b < t->n_blocks; b++) {
      bs[b] = t->block_site[b];
      bf[b] = t->block_fixed ? t->block_fixed[b] : 0;
      fb[b] = t->fixed_bias ? t->fixed_bias[b] : 0.0;
    }
    if (t->n_blocks >= 1 && t->site_bound)
      for (int q = 0; q <= t->n_sites; q++) sbd[q] = t->site_bound[q];
    else
      sbd[1] = 1;
    for (int b = 0; b < nl; b++) {
      double g = (t->theta_friction && b < t->n_blocks) ? t->theta_friction[b] : rp->msld_gamma;
      double a = std::exp(-g * rp->dt);
      lf[b] = a;
      ln[b] = std::sqrt((1 - a * a) * kT);
    }
    UPLOAD(d.block_site, bs); UPLOAD(d.site_bound, sbd); UPLOAD(d.block_fixed, bf);
    UPLOAD(d.fixed_bias, fb); UPLOAD(d.lfric, lf); UPLOAD(d.lnoise, ln);
    d.crk1 = t->cr_k1; d.crk2 = t->cr_k2; d.crk3 = t->cr_k3; d.crq0 = t->cr_q0; d.crw = t->cr_w;
    d.ncr = nl > 1 && (d.crk1 != 0 || d.crk2 != 0 || d.crk3 != 0);
    if (d.ncr) {
      std::vector<double> bq(nl, 0.0);
      for (int b = 0; b < t->n_blocks && t->block_charge; b++) bq[b] = t->block_charge[b];
      UPLOAD(d.blockcharge, bq);
    }
    d.nvb = t->n_varbias;
    std::vector<int4> vi(d.nvb);
    std::vector<double2> vp(d.nvb);
    for (int v = 0; v < d.nvb; v++) {
      vi[v] = make_int4(t->vb_i[v], t->vb_j[v], t->vb_type[v], t->vb_n[v]);
      vp[v] = make_double2(t->vb_l0[v], t->vb_k[v]);
    }
    UPLOAD(d.vb_i, vi); UPLOAD(d.vb_p, vp);
    // theta biases (one entry per site) and the piecewise constraint
    if (t->tcoll_k && t->tcoll_n) {
      std::vector<double> k(t->tcoll_k, t->tcoll_k + d.nsites), nn(t->tcoll_n, t->tcoll_n + d.nsites);
      double *pk = nullptr, *pn = nullptr;
      UPLOAD(pk, k);
      UPLOAD(pn, nn);
      d.tcoll_k = pk;
      d.tcoll_n = pn;
    }
    if (t->tinde_k) {
      std::vector<double> k(t->tinde_k, t->tinde_k + d.nsites);
      double *pk = nullptr;
      UPLOAD(pk, k);
      d.tinde_k = pk;
    }
    d.piecewise = t->piecewise;
    d.fixlam = t->fix_lambda;
    d.pw_width = t->pw_width;
    d.pw_k = t->pw_k;
    d.edge_k = t->edge_k;
    d.edge_n = t->edge_n;
    d.edge_alpha = t->edge_alpha;
    d.edge_phi = t->edge_phi;
    ALLOC(d.theta0, d.nsites);
    ALLOC(d.dtheta, nl);
    d.fnex = rp->fnex;
    d.softcore = rp->soft_core;
    d.softcore14 = rp->soft_core14;
  }
  // integrator parameters (system/state.cxx:282-297)
  {
    const double kT = kBoltzD * rp->temperature;
    d.halfdt = 0.5 * rp->dt;
    double a = std::exp(-rp->gamma * rp->dt);
    d.fric = a;
    d.noise = std::sqrt((1 - a * a) * kT);
    d.shaketol = rp->shake_tol;
    d.seed = rp->seed;
  }
#undef ALLOC
#undef UPLOAD
  if (s->alch) launch_theta_to_lambda(d, s->st, true);
  if (hipStreamSynchronize(s->st) != hipSuccess) return fail(set_err(BM_EHIP, "init sync"));
  *out = s;
  return BM_OK;
}

int bm_system_destroy(bm_system *s) {
  if (!s) return BM_EINVAL;
  (void)hipSetDevice(s->device);
  delete s;
  return BM_OK;
}

int bm_sync(bm_system *s) {
  if (!s) return BM_EINVAL;
  HIPCHK(hipStreamSynchronize(s->st));
  int flag = 0, nan = -1;
  HIPCHK(hipMemcpy(&flag, s->d.errflag, sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&nan, s->d.nanflag, sizeof(int), hipMemcpyDeviceToHost));
  if (!flag && nan < 0) return BM_OK;
  // every condition that was raised, the first non-finite value first (a
  // blow-up also scatters atoms into one column)
  std::string msg;
  if (nan >= 0) msg += "non-finite position or velocity (first at atom or DOF " + std::to_string(nan) + "); ";
  if (flag & 4) msg += "the tile list exceeded its capacity; ";
  if (flag & 2) msg += "a PME spread brick exceeded its atom capacity; ";
  if (flag & 1) msg += "a domdec column exceeded its capacity; ";
  msg.resize(msg.size() - 2);
  return set_err(nan >= 0 ? BM_ENAN : BM_EOVERFLOW, msg.c_str());
}

int bm_domdec_update(bm_system *s, int rebuild) {
  if (!s) return BM_EINVAL;
  blade::launch_virtual_place(s->d, s->st);
  if (rebuild) blade::launch_domdec_rebuild(s->d, s->maxBlocks, s->st);
  blade::launch_domdec_pack(s->d, s->maxBlocks, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

int bm_reset_force(bm_system *s, unsigned) {
  if (!s) return BM_EINVAL;
  blade::launch_domdec_pack(s->d, s->maxBlocks, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

// Each term launch is followed by an assembly of that term's forces into
// the force array; bm_calc_force assembles all terms in one pass.
int bm_getforce_nbdirect(bm_system *s, unsigned flags) {
  if (!s) return BM_EINVAL;
  blade::launch_nbdirect(s->d, flags & BM_CALC_ENERGY, s->alch, s->soft, s->nwgNb, s->st);
  if (s->alch) blade::launch_nb_dudl(s->d, s->soft, s->st);
  blade::launch_force_assemble(s->d, blade::kAsmNb, s->st);
  blade::launch_virtual_spread(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

int bm_getforce_ewald(bm_system *s, unsigned flags) {
  if (!s) return BM_EINVAL;
  if (s->d.elecmethod == BM_PME) {
    blade::launch_pme_grid(s->d, flags & BM_CALC_ENERGY, s->st);
    blade::launch_force_assemble(s->d, blade::kAsmPme, s->st);
    blade::launch_virtual_spread(s->d, s->st);
  }
  HIPCHK(hipGetLastError());
  return BM_OK;
}

// The bonded launch's term families one at a time: the launch evaluates
// the families in fam (the others write zero forces to their slots) and the
// assembly adds the slots to the force array.
static int bonded_families(bm_system *s, unsigned flags, unsigned fam) {
  if (!s) return BM_EINVAL;
  blade::launch_bonded(s->d, flags & BM_CALC_ENERGY, fam, s->st);
  blade::launch_force_assemble(s->d, blade::kAsmBonded, s->st);
  blade::launch_virtual_spread(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

// getforce_ewaldself (nbrecip/nbrecip.cu:52-77): PME self energy and its
// dU/dlambda
int bm_getforce_ewaldself(bm_system *s, unsigned flags) {
  if (s && s->d.elecmethod != BM_PME) return BM_OK;
  return bonded_families(s, flags, blade::kFamSelf);
}

// bond (+Urey-Bradley), angle, dihedral, improper and CMAP terms
int bm_getforce_bonded(bm_system *s, unsigned flags) {
  return bonded_families(s, flags, blade::kFamBond | blade::kFamAngle | blade::kFamDihe | blade::kFamImpr |
                                       blade::kFamCmap);
}
int bm_getforce_nb14(bm_system *s, unsigned flags) { return bonded_families(s, flags, blade::kFamNb14); }
int bm_getforce_nbex(bm_system *s, unsigned flags) {
  if (s && s->d.elecmethod != BM_PME) return BM_OK;
  return bonded_families(s, flags, blade::kFamNbex);
}
// fixed, variable and theta biases, charge restraints and MSLD atom
// restraints (msld/msld.cu:1000-1551)
int bm_getforce_msldbias(bm_system *s, unsigned flags) {
  return bonded_families(s, flags, blade::kFamBias | blade::kFamCats);
}
// NOE, harmonic, bond/angle/dihedral and RESD restraints (restrain/restrain.h:7-12)
int bm_getforce_restrain(bm_system *s, unsigned flags) { return bonded_families(s, flags, blade::kFamRestrain); }

int bm_calc_force(bm_system *s, long, unsigned flags) {
  if (!s) return BM_EINVAL;
  s->force_launches(flags & BM_CALC_ENERGY);
  blade::launch_virtual_spread(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

int bm_update(bm_system *s, long step, int energy_step) {
  if (!s) return BM_EINVAL;
  blade::launch_set_step(s->d, step, s->st);
  blade::launch_update(s->d, energy_step != 0, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

int bm_holonomic_position(bm_system *s) {
  if (!s) return BM_EINVAL;
  blade::launch_constrain_positions(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}
int bm_holonomic_rectify(bm_system *s) {
  if (!s) return BM_EINVAL;
  blade::launch_rectify_init(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}
int bm_holonomic_velocity(bm_system *s) {
  if (!s) return BM_EINVAL;
  blade::launch_constrain_velocities(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

int bm_msld_theta_to_lambda(bm_system *s) {
  if (!s) return BM_EINVAL;
  blade::launch_theta_to_lambda(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}
int bm_msld_chain_rule(bm_system *s) {
  if (!s) return BM_EINVAL;
  blade::launch_chain_rule(s->d, s->st);
  HIPCHK(hipGetLastError());
  return BM_OK;
}

// ---------------------------------------------------------------------
// list capacities: the reference grows its partner list x1.5 and retries
// (domdec/cull.cu:337-375, at most 12x the first size, domdec.cu:129-130);
// here an overflow halts the state updates on the device and bm_dynamics
// grows the buffer that overflowed and runs the steps again from the one
// that raised it
// ---------------------------------------------------------------------

// tile list arrays for maxTiles directed tiles (their contents are lost:
// the next step must rebuild the list)
static int set_tile_capacity(bm_system *s, long long maxTiles) {
  blade::Dev &d = s->d;
  int rc = BM_OK;
  s->maxTiles = maxTiles;
  d.maxtiles = maxTiles;
  d.maxchunks = (int)(maxTiles / d.tchunk + s->maxBlocks + 1);
  const size_t T = (size_t)maxTiles + 1;
  if ((rc = s->realloc_dev(d.tile_i, T)) || (rc = s->realloc_dev(d.tile_j, T)) ||
      (rc = s->realloc_dev(d.tile_ex, T)) || (rc = s->realloc_dev(d.tile_flag, T)) ||
      (rc = s->realloc_dev(d.chunk, (size_t)d.maxchunks)) ||
      (rc = s->realloc_dev(d.ichunk, (size_t)d.maxchunks * 3 * blade::kBlockAtoms)))
    return rc;
  s->capacityGen++;
  s->drop_graphs();
  return BM_OK;
}

// PME brick bins of pcap spline records each (emptied)
static int set_brick_capacity(bm_system *s, int pcap) {
  blade::Dev &d = s->d;
  if (d.elecmethod != BM_PME) return BM_OK;
  const size_t nbr = (size_t)d.pnb[0] * d.pnb[1] * d.pnb[2];
  d.pcap = pcap;
  if (int rc = s->realloc_dev(d.pbrec, nbr * (size_t)pcap * blade::pme_spline_rec_words(d.order))) return rc;
  if (int rc = s->realloc_dev(d.pbhdr, nbr * (size_t)pcap)) return rc;
  HIPCHK(hipMemset(d.pbcount, 0, nbr * sizeof(int)));
  s->capacityGen++;
  s->drop_graphs();
  return BM_OK;
}

int bm_set_list_capacity(bm_system *s, long long max_tiles, int brick_atoms) {
  if (!s) return BM_EINVAL;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->st));
  if (max_tiles > 0) {
    const long long MB = s->maxBlocks;
    if (int rc = set_tile_capacity(s, std::min(max_tiles, MB * MB))) return rc;
  }
  if (brick_atoms > 0)
    if (int rc = set_brick_capacity(s, brick_atoms)) return rc;
  return BM_OK;
}

int bm_get_list_capacity(bm_system *s, long long *max_tiles, int *brick_atoms) {
  if (!s) return BM_EINVAL;
  if (max_tiles) *max_tiles = s->maxTiles;
  if (brick_atoms) *brick_atoms = s->d.elecmethod == BM_PME ? s->d.pcap : 0;
  return BM_OK;
}

// grow what the raised flag names (x1.5), clear the flag; BM_EOVERFLOW when
// the flag cannot be cured by growth or growth passed 12x the first size
static int regrow(bm_system *s, int flag) {
  if (flag & 1) return BM_EOVERFLOW;   // a column beyond the sort's LDS capacity
  if (flag & 4) {
    const long long MB = s->maxBlocks;
    const long long mt = std::min(MB * MB, (long long)std::ceil(1.5 * (double)s->maxTiles));
    if (mt <= s->maxTiles || mt > 12 * s->maxTiles0) return BM_EOVERFLOW;
    if (int rc = set_tile_capacity(s, mt)) return rc;
  }
  if (flag & 2) {
    const int pc = (int)std::ceil(1.5 * (double)s->d.pcap);
    if ((long long)pc > 12LL * std::max(1, s->pcap0)) return BM_EOVERFLOW;
    if (int rc = set_brick_capacity(s, pc)) return rc;
  }
  // a halted step left its brick counts in place (the assembly that clears
  // them returned early)
  if (s->d.elecmethod == BM_PME)
    HIPCHK(hipMemset(s->d.pbcount, 0, (size_t)s->d.pnb[0] * s->d.pnb[1] * s->d.pnb[2] * sizeof(int)));
  HIPCHK(hipMemset(s->d.errflag, 0, sizeof(int)));
  return BM_OK;
}

// capture and instantiate the step graph of one (rebuild, energy) variant
static int capture_step_graph(bm_system *s, int rb, int en) {
  if (s->graph[rb][en]) return BM_OK;
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  s->step_launches(rb, en);
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(s->st, &g));
  s->graphSrc[rb][en] = g;
  HIPCHK(hipGraphInstantiate(&s->graph[rb][en], g, nullptr, nullptr, 0));
  return BM_OK;
}

int bm_capture_graphs(bm_system *s) {
  if (!s) return BM_EINVAL;
  for (int rb = 0; rb < 2; rb++)
    for (int en = 0; en < 2; en++)
      if (int rc = capture_step_graph(s, rb, en)) return rc;
  return BM_OK;
}

// True under rocprofv3 hardware-counter collection (it exports
// ROCPROF_COUNTER_COLLECTION).  Steps then run as eager launches instead of
// graph launches: with counters on, hipGraphLaunch of these step graphs has
// ended in a host segmentation fault inside the counter tool's own copy
// routine, while the same steps launched one by one collect normally.
static bool counter_collection() {
  static const bool on = [] {
    const char *v = std::getenv("ROCPROF_COUNTER_COLLECTION");
    return v && *v && std::strcmp(v, "0") != 0 && std::strcmp(v, "False") != 0 && std::strcmp(v, "false") != 0;
  }();
  return on;
}

// the launches of steps [step0, step0 + nsteps)
static int dynamics_launches(bm_system *s, long step0, long nsteps, int freq_nrg, int use_graph) {
  blade::launch_set_step(s->d, step0, s->st);
  for (long k = 0; k < nsteps; k++) {
    const long step = step0 + k;
    const int rb = (step % s->freqDomdec) == 0, en = (step % freq_nrg) == 0;
    if (!use_graph) {
      s->step_launches(rb, en);
      HIPCHK(hipGetLastError());
      continue;
    }
    if (int rc = capture_step_graph(s, rb, en)) return rc;
    HIPCHK(hipGraphLaunch(s->graph[rb][en], s->st));
  }
  return BM_OK;
}

int bm_dynamics(bm_system *s, long step0, long nsteps, int freq_nrg, int use_graph) {
  if (!s || nsteps < 0) return BM_EINVAL;
  if (freq_nrg <= 0) freq_nrg = 1 << 30;
  if (counter_collection()) use_graph = 0;
  const long end = step0 + nsteps;
  for (;;) {
    if (int rc = dynamics_launches(s, step0, end - step0, freq_nrg, use_graph)) return rc;
    // one synchronisation per call: a capacity overflow halted the state
    // at the step that raised it; grow the buffer and run from that step
    HIPCHK(hipStreamSynchronize(s->st));
    int flag = 0, nan = -1;
    HIPCHK(hipMemcpy(&flag, s->d.errflag, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&nan, s->d.nanflag, sizeof(int), hipMemcpyDeviceToHost));
    if (!flag || nan >= 0) break;
    long long at = 0;
    HIPCHK(hipMemcpy(&at, s->d.step, sizeof(long long), hipMemcpyDeviceToHost));
    if (at < step0 || at >= end) break;
    if (regrow(s, flag)) break;
    step0 = (long)at;
  }
  return bm_sync(s);
}

// ---------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------
}  // extern "C"

// One launch of a kernel type over several members: the members, the
// workgroup prefix table and device copies of the members' Dev.
struct MultiPlan {
  std::vector<int> members;
  std::vector<int> wg0;            // host copy, [n+1]
  blade::Multi m{};
  blade::Dev *devs = nullptr;      // device
  int *wg0_d = nullptr;            // device
  // the group's template keys
  int order = 0, K0 = 0, K1 = 0, elec = 0, vdw = 0;
  bool alch = false, soft = false;
  size_t lds = 0;
};

struct bm_batch {
  std::vector<bm_system *> sys;
  int device = 0;
  hipStream_t st = nullptr, st2 = nullptr, st3 = nullptr;
  hipEvent_t fork = nullptr, fork2 = nullptr, join2 = nullptr, join3 = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraph_t graphSrc[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  // fused launches of a non-energy step
  MultiPlan pack, normals, bonded, tail;
  std::vector<MultiPlan> bin, spread, fftp, fftx, mfp, mfx, nb, nbd, asmb;
  MultiPlan rebuild[blade::kRebuildKernels];   // the list rebuild, one plan per kernel
  std::vector<int> fftSingle;      // PME members on neither fused FFT path (other sizes or shapes)
  std::vector<int> gens;           // members' capacity generations the launch plans were sized for
  ~bm_batch() {
    if (st) (void)hipStreamSynchronize(st);
    if (st2) (void)hipStreamSynchronize(st2);
    if (st3) (void)hipStreamSynchronize(st3);
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        if (graph[a][b]) (void)hipGraphExecDestroy(graph[a][b]);
        if (graphSrc[a][b]) (void)hipGraphDestroy(graphSrc[a][b]);
      }
    for (auto e : join)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {fork, fork2, join2, join3})
      if (e) (void)hipEventDestroy(e);
    for (MultiPlan *p : plans()) {
      if (p->devs) (void)hipFree(p->devs);
      if (p->wg0_d) (void)hipFree(p->wg0_d);
    }
    if (st3) (void)hipStreamDestroy(st3);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }
  std::vector<MultiPlan *> plans() {
    std::vector<MultiPlan *> v = {&pack, &normals, &bonded, &tail};
    for (MultiPlan &p : rebuild) v.push_back(&p);
    for (auto *g : {&bin, &spread, &fftp, &fftx, &mfp, &mfx, &nb, &nbd, &asmb})
      for (MultiPlan &p : *g) v.push_back(&p);
    return v;
  }
  // the members' current Dev into every plan's device copies
  int upload_devs() {
    for (MultiPlan *p : plans()) {
      std::vector<blade::Dev> h;
      for (int k : p->members) h.push_back(sys[k]->d);
      if (hipMemcpy(p->devs, h.data(), h.size() * sizeof(blade::Dev), hipMemcpyHostToDevice) != hipSuccess)
        return set_err(BM_EHIP, "batch: Dev upload");
    }
    return BM_OK;
  }
  // one non-energy step with one launch per kernel type and group
  void fused_step(bool rebuild) {
    if (rebuild) {
      // the list rebuild of every member, one launch per rebuild kernel
      blade::Multi rm[blade::kRebuildKernels];
      for (int k = 0; k < blade::kRebuildKernels; k++) rm[k] = this->rebuild[k].m;
      blade::launch_domdec_rebuild_multi(rm, st);
    }
    blade::launch_pack_multi(pack.m, st);
    (void)hipEventRecord(fork2, st);
    // nbdirect first: the largest kernel takes the CUs before the second
    // stream's smaller ones fill them
    for (MultiPlan &p : nb) blade::launch_nbdirect_multi(p.m, p.elec, p.vdw, false, p.alch, p.soft, p.lds, st);
    (void)hipStreamWaitEvent(st2, fork2, 0);
    (void)hipStreamWaitEvent(st3, fork2, 0);
    // the FP64 dU/dlambda of the members' alchemical atoms, the normals and
    // the bonded terms on the third stream, the PME chain on the second
    for (MultiPlan &p : nbd) blade::launch_nb_dudl_multi(p.m, p.elec, p.vdw, p.soft, st3);
    blade::launch_normals_multi(normals.m, st3);
    blade::launch_bonded_multi(bonded.m, false, st3);
    (void)hipEventRecord(join3, st3);
    for (size_t g = 0; g < bin.size(); g++) {
      blade::launch_pme_bin_multi(bin[g].m, bin[g].order, st2);
      blade::launch_pme_spread_multi(spread[g].m, spread[g].order, spread[g].lds, st2);
    }
    // FFT groups in turn on st2 (one stream per branch: the process has
    // four hardware queues)
    for (size_t g = 0; g < fftp.size(); g++)
      blade::launch_pme_fft_multi(fftp[g].m, fftx[g].m, fftp[g].K0, fftp[g].K1, false, st2);
    for (size_t g = 0; g < mfp.size(); g++)
      blade::launch_pme_fft_mfma_multi(mfp[g].m, mfx[g].m, mfp[g].K0, mfp[g].K1, false, st2);
    for (int k : fftSingle) blade::launch_pme_fft(sys[k]->d, false, st2);
    (void)hipEventRecord(join2, st2);
    (void)hipStreamWaitEvent(st, join2, 0);
    (void)hipStreamWaitEvent(st, join3, 0);
    for (MultiPlan &p : asmb)
      blade::launch_assemble_multi(p.m, p.order, blade::kAsmNb | blade::kAsmBonded | blade::kAsmPme, true, st);
    blade::launch_tail_multi(tail.m, st);
  }
};

// members -> plan with per-member workgroup counts from grid()
template <class F>
static int make_plan(bm_batch *b, MultiPlan &p, const std::vector<int> &members, F grid) {
  p.members = members;
  p.wg0.assign(1, 0);
  for (int k : members) p.wg0.push_back(p.wg0.back() + grid(b->sys[k]->d));
  if (hipMalloc(&p.devs, std::max<size_t>(1, members.size()) * sizeof(blade::Dev)) != hipSuccess ||
      hipMalloc(&p.wg0_d, p.wg0.size() * sizeof(int)) != hipSuccess ||
      hipMemcpy(p.wg0_d, p.wg0.data(), p.wg0.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return set_err(BM_EHIP, "batch: plan tables");
  p.m.devs = p.devs;
  p.m.wg0 = p.wg0_d;
  p.m.nsys = (int)members.size();
  p.m.nwg = p.wg0.back();
  return BM_OK;
}

// members grouped by key(), one plan per key
template <class K, class F>
static int make_groups(bm_batch *b, std::vector<MultiPlan> &out, const std::vector<int> &members, K key, F grid) {
  std::vector<std::pair<decltype(key(0)), std::vector<int>>> groups;
  for (int k : members) {
    auto kk = key(k);
    auto it = std::find_if(groups.begin(), groups.end(), [&](const auto &g) { return g.first == kk; });
    if (it == groups.end()) groups.push_back({kk, {k}});
    else it->second.push_back(k);
  }
  out.resize(groups.size());
  for (size_t g = 0; g < groups.size(); g++)
    if (int rc = make_plan(b, out[g], groups[g].second, grid)) return rc;
  return BM_OK;
}

static int build_plans(bm_batch *b) {
  using blade::Dev;
  const int n = (int)b->sys.size();
  std::vector<int> all(n), pme, plane, mplane;
  for (int k = 0; k < n; k++) {
    all[k] = k;
    const Dev &dk = b->sys[k]->d;
    if (dk.elecmethod == BM_PME) {
      pme.push_back(k);
      if (dk.fft_mode == 0 && blade::fft_plane_path(dk)) plane.push_back(k);
      else if (dk.fft_mode == 1 && dk.K[1] == dk.K[2]) mplane.push_back(k);
      else b->fftSingle.push_back(k);
    }
  }
  auto D = [&](int k) -> const Dev & { return b->sys[k]->d; };
  int rc;
  if ((rc = make_plan(b, b->pack, all, blade::grid_pack))) return rc;
  if ((rc = make_plan(b, b->normals, all, blade::grid_normals))) return rc;
  if ((rc = make_plan(b, b->bonded, all, blade::grid_bonded))) return rc;
  if ((rc = make_plan(b, b->tail, all, blade::grid_tail))) return rc;
  for (int r = 0; r < blade::kRebuildKernels; r++)
    if ((rc = make_plan(b, b->rebuild[r], all, [r](const Dev &d) { return blade::grid_rebuild(d, r); }))) return rc;
  if ((rc = make_groups(b, b->bin, pme, [&](int k) { return D(k).order; }, blade::grid_pme_bin))) return rc;
  if ((rc = make_groups(b, b->spread, pme, [&](int k) { return D(k).order; }, blade::grid_pme_spread))) return rc;
  for (size_t g = 0; g < b->bin.size(); g++) {
    b->bin[g].order = b->spread[g].order = D(b->bin[g].members[0]).order;
    for (int k : b->spread[g].members) b->spread[g].lds = std::max(b->spread[g].lds, blade::lds_pme_spread(D(k)));
  }
  auto fkey = [&](int k) { return std::make_pair(D(k).K[0], D(k).K[1]); };
  if ((rc = make_groups(b, b->fftp, plane, fkey, blade::grid_fft_planes))) return rc;
  if ((rc = make_groups(b, b->fftx, plane, fkey, blade::grid_fft_x))) return rc;
  for (size_t g = 0; g < b->fftp.size(); g++) {
    b->fftp[g].K0 = D(b->fftp[g].members[0]).K[0];
    b->fftp[g].K1 = D(b->fftp[g].members[0]).K[1];
  }
  if ((rc = make_groups(b, b->mfp, mplane, fkey, blade::grid_fft_mfma_planes))) return rc;
  if ((rc = make_groups(b, b->mfx, mplane, fkey, blade::grid_fft_mfma_x))) return rc;
  for (size_t g = 0; g < b->mfp.size(); g++) {
    b->mfp[g].K0 = D(b->mfp[g].members[0]).K[0];
    b->mfp[g].K1 = D(b->mfp[g].members[0]).K[1];
  }
  // nonbonded launches grouped by method and lambda flags; the FP64
  // dU/dlambda of members with lambda sites grouped by method and soft core
  auto nkey = [&](int k) {
    const bm_system *s = b->sys[k];
    return std::make_tuple(s->d.elecmethod, s->d.vdwmethod, s->alch, s->soft, blade::lds_nbdirect(s->d) == 0);
  };
  std::vector<int> alchm;
  for (int k : all)
    if (b->sys[k]->alch) alchm.push_back(k);
  if ((rc = make_groups(b, b->nb, all, nkey, blade::grid_nbdirect))) return rc;
  auto dkey = [&](int k) {
    const bm_system *s = b->sys[k];
    return std::make_tuple(s->d.elecmethod, s->d.vdwmethod, s->soft);
  };
  if ((rc = make_groups(b, b->nbd, alchm, dkey, blade::grid_nb_dudl))) return rc;
  for (std::vector<MultiPlan> *grp : {&b->nb, &b->nbd})
    for (MultiPlan &p : *grp) {
      const bm_system *s = b->sys[p.members[0]];
      p.elec = s->d.elecmethod;
      p.vdw = s->d.vdwmethod;
      p.alch = s->alch;
      p.soft = s->soft;
      p.lds = 0;
      if (blade::lds_nbdirect(s->d))
        for (int k : p.members) p.lds = std::max(p.lds, blade::lds_nbdirect(D(k)));
    }
  auto okey = [&](int k) { return D(k).elecmethod == BM_PME ? D(k).order : 0; };
  if ((rc = make_groups(b, b->asmb, all, okey, blade::grid_assemble))) return rc;
  for (MultiPlan &p : b->asmb) p.order = okey(p.members[0]);
  return b->upload_devs();
}

extern "C" {

int bm_batch_create(int n, bm_system *const *systems, bm_batch **out) {
  if (n <= 0 || !systems || !out) return set_err(BM_EINVAL, "batch arguments");
  *out = nullptr;
  for (int k = 0; k < n; k++) {
    if (!systems[k]) return set_err(BM_EINVAL, "null batch member");
    if (systems[k]->device != systems[0]->device) return set_err(BM_EINVAL, "batch members on different devices");
    if (systems[k]->freqDomdec != systems[0]->freqDomdec)
      return set_err(BM_EINVAL, "batch members with different list rebuild cadence");
    if (systems[k]->d.nvirt2 + systems[k]->d.nvirt3 > 0)
      return set_err(BM_ENOTSUP, "batch members with virtual sites (run them as single systems)");
  }
  HIPCHK(hipSetDevice(systems[0]->device));
  bm_batch *b = new bm_batch();
  b->device = systems[0]->device;
  b->sys.assign(systems, systems + n);
  for (int k = 0; k < n; k++) b->gens.push_back(systems[k]->capacityGen);
  b->join.assign(n, nullptr);
  bool ok = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->st3, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->join2, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->join3, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; ok && k < n; k++) ok = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    delete b;
    return set_err(BM_EHIP, "batch stream/events");
  }
  if (int rc = build_plans(b)) {
    delete b;
    return rc;
  }
  *out = b;
  return BM_OK;
}

// the launches of one batch step on the batch stream
static int batch_step_launches(bm_batch *b, int rb, int en) {
  if (!en) {
    // non-energy steps: one launch per kernel type over all members
    b->fused_step(rb);
  } else {
    // energy steps: every member's own step as a parallel branch
    HIPCHK(hipEventRecord(b->fork, b->st));
    for (size_t m = 0; m < b->sys.size(); m++) {
      bm_system *s = b->sys[m];
      HIPCHK(hipStreamWaitEvent(s->st, b->fork, 0));
      s->step_launches(rb, en);
      HIPCHK(hipEventRecord(b->join[m], s->st));
      HIPCHK(hipStreamWaitEvent(b->st, b->join[m], 0));
    }
  }
  return BM_OK;
}

// capture and instantiate the batch step graph of one (rebuild, energy)
// variant
static int capture_batch_graph(bm_batch *b, int rb, int en) {
  if (b->graph[rb][en]) return BM_OK;
  HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  if (int rc = batch_step_launches(b, rb, en)) return rc;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(b->st, &g));
  b->graphSrc[rb][en] = g;
  HIPCHK(hipGraphInstantiate(&b->graph[rb][en], g, nullptr, nullptr, 0));
  return BM_OK;
}

int bm_batch_capture_graphs(bm_batch *b) {
  if (!b) return BM_EINVAL;
  HIPCHK(hipSetDevice(b->device));
  if (int rc = b->upload_devs()) return rc;
  for (int rb = 0; rb < 2; rb++)
    for (int en = 0; en < 2; en++)
      if (int rc = capture_batch_graph(b, rb, en)) return rc;
  return BM_OK;
}

int bm_batch_dynamics(bm_batch *b, long step0, long nsteps, int freq_nrg) {
  if (!b || nsteps < 0) return BM_EINVAL;
  // the fused launches were sized for the members' list capacities: a
  // member that grew since (bm_dynamics on it alone) needs a new batch
  for (size_t k = 0; k < b->sys.size(); k++)
    if (b->sys[k]->capacityGen != b->gens[k])
      return set_err(BM_EINVAL, "a batch member's list capacity changed after bm_batch_create");
  if (freq_nrg <= 0) freq_nrg = 1 << 30;
  HIPCHK(hipSetDevice(b->device));
  if (int rc = b->upload_devs()) return rc;
  // every member's step counter, ordered before the first graph launch;
  // each member's own work is joined into the batch stream by the graph
  for (bm_system *s : b->sys) {
    HIPCHK(hipEventRecord(b->fork, s->st));
    HIPCHK(hipStreamWaitEvent(b->st, b->fork, 0));
  }
  for (bm_system *s : b->sys) blade::launch_set_step(s->d, step0, b->st);
  const int freqDomdec = b->sys[0]->freqDomdec;
  const bool eager = counter_collection();
  for (long k = 0; k < nsteps; k++) {
    const long step = step0 + k;
    const int rb = (step % freqDomdec) == 0, en = (step % freq_nrg) == 0;
    if (eager) {
      if (int rc = batch_step_launches(b, rb, en)) return rc;
      HIPCHK(hipGetLastError());
      continue;
    }
    if (int rc = capture_batch_graph(b, rb, en)) return rc;
    HIPCHK(hipGraphLaunch(b->graph[rb][en], b->st));
  }
  // members' later work (getters, eager steps) follows the batch
  HIPCHK(hipEventRecord(b->fork, b->st));
  for (bm_system *s : b->sys) HIPCHK(hipStreamWaitEvent(s->st, b->fork, 0));
  return BM_OK;
}

double bm_uniform(uint64_t seed, long step, uint32_t k) {
  const uint4 o = blade::philox(make_uint4(k, (unsigned)step, (unsigned)((unsigned long long)step >> 32), 0x554E4946u),
                                make_uint2((unsigned)seed, (unsigned)(seed >> 32)));
  return ((double)((((unsigned long long)o.x) << 21) ^ (o.y >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
}

// potential energy of a member at its current positions (a list rebuild,
// then one energy evaluation)
static int member_energy(bm_system *s, double &e) {
  double E[BM_NTERMS];
  int rc = bm_domdec_update(s, 1);
  if (!rc) rc = bm_calc_force(s, 0, BM_CALC_ENERGY);
  if (!rc) rc = bm_sync(s);
  if (!rc) rc = bm_get_energy(s, E);
  e = E[BM_EPOTENTIAL];
  return rc;
}

// swap a device array of two members through a scratch buffer
static int swap_dev(bm_system *a, bm_system *b, double *pa, double *pb, double *tmp, size_t n, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(tmp, pa, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(pa, pb, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(pb, tmp, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  (void)a;
  (void)b;
  return BM_OK;
}

// Replica exchange over the members of a batch, with `tmp` (at least
// 3N + nL doubles of device memory) as the staging buffer of the swaps.
static int batch_exchange(bm_batch *b, long step, int parity, uint64_t seed, int32_t *replica,
                          int32_t *accepted, double *dW, double *tmp) {
  const int n = (int)b->sys.size();
  const size_t N = 3 * (size_t)b->sys[0]->d.n + b->sys[0]->d.nl;
  std::vector<double> e0(n, 0.0), e1(n, 0.0);
  int rc = BM_OK;
  for (int k = 0; k < n && !rc; k++) rc = member_energy(b->sys[k], e0[k]);
  auto swap_pos = [&](bm_system *x, bm_system *y) {
    int r = swap_dev(x, y, x->d.pos, y->d.pos, tmp, N, b->st);
    if (!r && x->alch) blade::launch_theta_to_lambda(x->d, b->st, true);
    if (!r && y->alch) blade::launch_theta_to_lambda(y->d, b->st, true);
    if (!r && hipStreamSynchronize(b->st) != hipSuccess) r = BM_EHIP;
    return r;
  };
  for (int k = parity; k + 1 < n && !rc; k += 2) rc = swap_pos(b->sys[k], b->sys[k + 1]);
  for (int k = 0; k < n && !rc; k++)
    if (k >= parity && k - parity < 2 * ((n - parity) / 2)) rc = member_energy(b->sys[k], e1[k]);
  for (int k = parity; k + 1 < n && !rc; k += 2) {
    bm_system *x = b->sys[k], *y = b->sys[k + 1];
    const double w = (e1[k] - e0[k]) / x->kT + (e1[k + 1] - e0[k + 1]) / y->kT;
    // uniform in (0,1) from (seed, step, pair); stream tag "REX"
    const double u = bm_uniform(seed, step, 0x52455800u | (unsigned)k);
    const bool acc = u < std::exp(-w);
    if (dW) dW[k] = w;
    if (accepted) accepted[k] = acc ? 1 : 0;
    if (acc) {
      rc = swap_dev(x, y, x->d.vel, y->d.vel, tmp, N, b->st);
      if (replica) std::swap(replica[k], replica[k + 1]);
    } else {
      rc = swap_pos(x, y);
    }
  }
  // every member's list from its final positions, for the steps that follow
  for (int k = 0; k < n && !rc; k++) {
    rc = bm_domdec_update(b->sys[k], 1);
    if (!rc) rc = bm_sync(b->sys[k]);
  }
  return rc;
}

// members of one exchange must have equal atom and block counts and equal
// boxes (positions are exchanged, boxes are not)
static int exchange_check(bm_batch *b, int parity) {
  if (!b || (parity != 0 && parity != 1)) return set_err(BM_EINVAL, "exchange arguments");
  const int n = (int)b->sys.size();
  for (int k = 1; k < n; k++) {
    const blade::Dev &d0 = b->sys[0]->d, &dk = b->sys[k]->d;
    if (dk.n != d0.n || dk.nl != d0.nl)
      return set_err(BM_EINVAL, "replica exchange needs members of equal atom and block counts");
    if (dk.box[0] != d0.box[0] || dk.box[1] != d0.box[1] || dk.box[2] != d0.box[2])
      return set_err(BM_EINVAL, "replica exchange needs members with equal boxes");
  }
  return BM_OK;
}

int bm_batch_exchange(bm_batch *b, long step, int parity, uint64_t seed, int32_t *replica, int32_t *accepted,
                      double *dW) {
  if (int rc = exchange_check(b, parity)) return rc;
  if (int rc = bm_batch_sync(b)) return rc;
  HIPCHK(hipSetDevice(b->device));
  const size_t N = 3 * (size_t)b->sys[0]->d.n + b->sys[0]->d.nl;
  double *tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, N * sizeof(double)));
  const int rc = batch_exchange(b, step, parity, seed, replica, accepted, dW, tmp);
  (void)hipFree(tmp);
  return rc;
}

// ---------------------------------------------------------------------
// Replica exchange that allocates nothing (bm_batch_exchange_noalloc).
// Its swap staging lives in a file-scope device array of
// kExchangeScratch doubles (16 MiB, held by every device the library runs
// on until the process ends).  Host code reaches each device's copy only
// through the address hipGetSymbolAddress returns on that device, fetched
// on the first call on the device and kept per device.  Every call on one
// device shares that array: calls that overlap on one device race.  The
// staging is written in full before it is read, so nothing is cleared.
// ---------------------------------------------------------------------
constexpr size_t kExchangeScratch = size_t(1) << 21;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kExchangeScratch];

// this device's copy of the scratch array (the device must be current)
static int exchange_scratch(int device, double **out) {
  static std::mutex mu;
  static double *ptrs[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return set_err(BM_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(BM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptrs[device] = static_cast<double *>(p);
  }
  *out = ptrs[device];
  return BM_OK;
}

int bm_batch_exchange_noalloc(bm_batch *b, long step, int parity, uint64_t seed, int32_t *replica,
                              int32_t *accepted, double *dW) {
  if (int rc = exchange_check(b, parity)) return rc;
  const size_t N = 3 * (size_t)b->sys[0]->d.n + b->sys[0]->d.nl;
  if (N > kExchangeScratch)
    return set_err(BM_EINVAL, "bm_batch_exchange_noalloc: members of " + std::to_string(N) +
                                  " degrees of freedom exceed its scratch of " + std::to_string(kExchangeScratch));
  if (int rc = bm_batch_sync(b)) return rc;
  HIPCHK(hipSetDevice(b->device));
  double *tmp = nullptr;
  if (int rc = exchange_scratch(b->device, &tmp)) return rc;
  return batch_exchange(b, step, parity, seed, replica, accepted, dW, tmp);
}

int bm_batch_sync(bm_batch *b) {
  if (!b) return BM_EINVAL;
  HIPCHK(hipStreamSynchronize(b->st));
  for (bm_system *s : b->sys) {
    if (int rc = bm_sync(s)) return rc;
    int nan = -1;
    if (int rc = bm_get_nan_flag(s, &nan)) return rc;
    if (nan != -1) return set_err(BM_ENAN, "NaN/Inf in a batch member");
  }
  return BM_OK;
}

int bm_batch_destroy(bm_batch *b) {
  if (!b) return BM_EINVAL;
  (void)hipSetDevice(b->device);
  delete b;
  return BM_OK;
}

// Per-phase timing with events on the system stream: nsteps eager steps
// (rebuild every freq_domdec steps, no energies) with events around the
// nbdirect launch and around the PME grid (bin, spread, FFT) plus the force
// assembly with the PME gather.  out = {ms nbdirect, ms pme, ms step}.
int bm_profile_steps(bm_system *s, long step0, long nsteps, double *out) {
  if (!s || nsteps <= 0 || !out) return BM_EINVAL;
  double tn = 0, tp = 0, tt = 0;
  blade::launch_set_step(s->d, step0, s->st);
  for (long k = 0; k < nsteps; k++) {
    const long step = step0 + k;
    const bool rb = (step % s->freqDomdec) == 0;
    HIPCHK(hipEventRecord(s->ev[0], s->st));
    blade::launch_virtual_place(s->d, s->st);
    if (rb) blade::launch_domdec_rebuild(s->d, s->maxBlocks, s->st);
    blade::launch_domdec_pack(s->d, s->maxBlocks, s->st);
    HIPCHK(hipEventRecord(s->ev[1], s->st));
    blade::launch_nbdirect(s->d, false, s->alch, s->soft, s->nwgNb, s->st);
    HIPCHK(hipEventRecord(s->ev[2], s->st));
    if (s->alch) blade::launch_nb_dudl(s->d, s->soft, s->st);
    blade::launch_bonded(s->d, false, blade::kFamAll, s->st);
    HIPCHK(hipEventRecord(s->ev[3], s->st));
    if (s->d.elecmethod == BM_PME) blade::launch_pme_grid(s->d, false, s->st);
    blade::launch_force_assemble(s->d, blade::kAsmNb | blade::kAsmBonded | blade::kAsmPme, s->st);
    blade::launch_virtual_spread(s->d, s->st);
    HIPCHK(hipEventRecord(s->ev[5], s->st));
    blade::launch_update(s->d, false, s->st);
    HIPCHK(hipEventRecord(s->ev[4], s->st));
    HIPCHK(hipEventSynchronize(s->ev[4]));
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, s->ev[1], s->ev[2]));
    HIPCHK(hipEventElapsedTime(&b, s->ev[3], s->ev[5]));
    HIPCHK(hipEventElapsedTime(&c, s->ev[0], s->ev[4]));
    tn += a; tp += b; tt += c;
  }
  out[0] = tn / nsteps;
  out[1] = tp / nsteps;
  out[2] = tt / nsteps;
  s->msNb = out[0]; s->msPme = out[1]; s->msTotal = out[2]; s->nMeas = nsteps;
  return BM_OK;
}

int bm_time_nbdirect(bm_system *s, int reps, double *out) {
  if (!s || reps <= 0 || !out) return BM_EINVAL;
  HIPCHK(hipEventRecord(s->ev[0], s->st));
  for (int r = 0; r < reps; r++) blade::launch_nbdirect(s->d, false, s->alch, s->soft, s->nwgNb, s->st);
  HIPCHK(hipEventRecord(s->ev[1], s->st));
  HIPCHK(hipEventSynchronize(s->ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
  out[0] = ms / reps;
  return BM_OK;
}

int bm_time_pme(bm_system *s, int reps, double *out) {
  if (!s || reps <= 0 || !out) return BM_EINVAL;
  if (s->d.elecmethod != BM_PME) return set_err(BM_ENOTSUP, "bm_time_pme: no PME");
  double t[3] = {0, 0, 0};
  for (int r = 0; r < reps; r++) {
    HIPCHK(hipEventRecord(s->ev[0], s->st));
    blade::launch_pme_spread(s->d, s->st);
    HIPCHK(hipEventRecord(s->ev[1], s->st));
    blade::launch_pme_fft(s->d, false, s->st);
    HIPCHK(hipEventRecord(s->ev[2], s->st));
    blade::launch_force_assemble(s->d, blade::kAsmPme, s->st);
    HIPCHK(hipEventRecord(s->ev[3], s->st));
    HIPCHK(hipEventSynchronize(s->ev[3]));
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
    HIPCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
    HIPCHK(hipEventElapsedTime(&c, s->ev[2], s->ev[3]));
    t[0] += a; t[1] += b; t[2] += c;
  }
  for (int k = 0; k < 3; k++) out[k] = t[k] / reps;
  return BM_OK;
}

int bm_get_kernel_times(bm_system *s, double *ms_nb, double *ms_pme, double *ms_total, long *n) {
  if (!s) return BM_EINVAL;
  if (ms_nb) *ms_nb = s->msNb;
  if (ms_pme) *ms_pme = s->msPme;
  if (ms_total) *ms_total = s->msTotal;
  if (n) *n = s->nMeas;
  return BM_OK;
}

}  // extern "C"

__global__ void bm_calibration_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" {

int bm_calibration_copy(int device, size_t bytes, int reps, double *out) {
  if (!out || reps <= 0 || bytes < 16) return set_err(BM_EINVAL, "calibration copy arguments");
  HIPCHK(hipSetDevice(device));
  const size_t n = bytes / 16;
  uint4 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = BM_OK;
  float ms = 0;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = set_err(BM_EHIP, "calibration copy allocation");
  } else {
    const int grid = 256 * 32;
    hipMemsetAsync(a, 1, n * 16, st);
    bm_calibration_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);   // warm
    hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) bm_calibration_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);
    hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = set_err(BM_EHIP, "calibration copy timing");
    out[0] = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

static int d2h(bm_system *s, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return BM_OK;
}

int bm_get_positions(bm_system *s, double *pos, double *theta) {
  if (!s) return BM_EINVAL;
  const int n = s->d.n;
  std::vector<double> x(3 * (size_t)n + s->d.nl);
  if (int rc = d2h(s, x.data(), s->d.pos, x.size() * sizeof(double))) return rc;
  if (pos) std::memcpy(pos, x.data(), 3 * (size_t)n * sizeof(double));
  if (theta) std::memcpy(theta, x.data() + 3 * (size_t)n, s->d.nl * sizeof(double));
  return BM_OK;
}

// the box-dependent members of the device system, as bm_system_create sets
// them; captured step graphs hold the old values, so they are dropped and
// captured again on next use
static void set_box(bm_system *s, const double *box) {
  blade::Dev &d = s->d;
  for (int k = 0; k < 3; k++) {
    d.box[k] = box[k];
    d.invbox[k] = 1.0 / box[k];
    d.boxf[k] = (float)box[k];
    d.fracscale[k] = (float)(box[k] / 4294967296.0);
    d.ddscale[k] = box[k] * (1.0 / 4294967296.0);
  }
  const double V = box[0] * box[1] * box[2];
  d.pmefacd = 0.5 * blade::kElectricD / M_PI / V;
  d.pmefac = (float)d.pmefacd;
  s->drop_graphs();
}

int bm_volume_move(bm_system *s, const double *scale) {
  if (!s || !scale) return BM_EINVAL;
  HIPCHK(hipSetDevice(s->device));
  double oldbox[3], box[3];
  for (int k = 0; k < 3; k++) {
    oldbox[k] = s->d.box[k];
    box[k] = s->d.box[k] * scale[k];
    if (!(scale[k] > 0) || !(box[k] > s->d.rc)) return set_err(BM_EINVAL, "volume move: box edge must exceed rcut");
  }
  blade::launch_volume_move(s->d, scale, oldbox, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  set_box(s, box);
  return BM_OK;
}

int bm_volume_restore(bm_system *s, const double *box) {
  if (!s || !box) return BM_EINVAL;
  HIPCHK(hipSetDevice(s->device));
  blade::launch_volume_restore(s->d, s->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->st));
  set_box(s, box);
  return BM_OK;
}

int bm_get_box(bm_system *s, double *box) {
  if (!s || !box) return BM_EINVAL;
  for (int k = 0; k < 3; k++) box[k] = s->d.box[k];
  return BM_OK;
}

int bm_set_positions(bm_system *s, const double *pos, const double *theta) {
  if (!s) return BM_EINVAL;
  const int n = s->d.n;
  if (pos) HIPCHK(hipMemcpyAsync(s->d.pos, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->st));
  if (theta) {
    HIPCHK(hipMemcpyAsync(s->d.pos + 3 * (size_t)n, theta, s->d.nl * sizeof(double), hipMemcpyHostToDevice, s->st));
    if (s->alch) blade::launch_theta_to_lambda(s->d, s->st, true);
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return BM_OK;
}

int bm_set_velocities(bm_system *s, const double *vel, const double *theta_vel) {
  if (!s) return BM_EINVAL;
  const int n = s->d.n;
  if (vel) HIPCHK(hipMemcpyAsync(s->d.vel, vel, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->st));
  if (theta_vel)
    HIPCHK(hipMemcpyAsync(s->d.vel + 3 * (size_t)n, theta_vel, s->d.nl * sizeof(double), hipMemcpyHostToDevice,
                          s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return BM_OK;
}

int bm_get_velocities(bm_system *s, double *vel, double *theta_vel) {
  if (!s) return BM_EINVAL;
  const int n = s->d.n;
  std::vector<double> v(3 * (size_t)n + s->d.nl);
  if (int rc = d2h(s, v.data(), s->d.vel, v.size() * sizeof(double))) return rc;
  if (vel) std::memcpy(vel, v.data(), 3 * (size_t)n * sizeof(double));
  if (theta_vel) std::memcpy(theta_vel, v.data() + 3 * (size_t)n, s->d.nl * sizeof(double));
  return BM_OK;
}

int bm_get_forces(bm_system *s, double *force, double *theta_force) {
  if (!s) return BM_EINVAL;
  const int n = s->d.n;
  std::vector<long long> f(3 * (size_t)n + s->d.nl);
  if (int rc = d2h(s, f.data(), s->d.force, f.size() * sizeof(long long))) return rc;
  if (force)
    for (size_t i = 0; i < 3 * (size_t)n; i++) force[i] = (double)f[i] * blade::kFixInvD;
  if (theta_force)
    for (int b = 0; b < s->d.nl; b++) theta_force[b] = (double)f[3 * (size_t)n + b] * blade::kFixInvD;
  return BM_OK;
}

int bm_get_force_fixed(bm_system *s, long long *force) {
  if (!s || !force) return BM_EINVAL;
  return d2h(s, force, s->d.force, (3 * (size_t)s->d.n + s->d.nl) * sizeof(long long));
}

int bm_get_dudl(bm_system *s, double *dudl) {
  if (!s || !dudl) return BM_EINVAL;
  std::vector<long long> f(s->d.nl);
  if (int rc = d2h(s, f.data(), s->d.dudl, f.size() * sizeof(long long))) return rc;
  for (int b = 0; b < s->d.nl; b++) dudl[b] = (double)f[b] * blade::kFixInvD;
  return BM_OK;
}

int bm_get_lambda(bm_system *s, double *lambda) {
  if (!s || !lambda) return BM_EINVAL;
  return d2h(s, lambda, s->d.lambda, s->d.nl * sizeof(double));
}

int bm_get_energy(bm_system *s, double *energy) {
  if (!s || !energy) return BM_EINVAL;
  long long e[BM_NTERMS];
  if (int rc = d2h(s, e, s->d.energy, sizeof(e))) return rc;
  double pot = 0;
  for (int k = 0; k < BM_NTERMS; k++) energy[k] = (double)e[k] * blade::kFixInvD;
  for (int k = 0; k < BM_EPOTENTIAL; k++) pot += energy[k];
  energy[BM_EPOTENTIAL] = pot;
  energy[BM_ETOTAL] = pot + energy[BM_EKINETIC];
  return BM_OK;
}

int bm_get_nan_flag(bm_system *s, int *flag) {
  if (!s || !flag) return BM_EINVAL;
  return d2h(s, flag, s->d.nanflag, sizeof(int));
}

int bm_get_domdec_sizes(bm_system *s, int *nb, int *nt, int *ne) {
  if (!s) return BM_EINVAL;
  int v[3] = {0, 0, 0};
  if (int rc = d2h(s, &v[0], s->d.nblocks, sizeof(int))) return rc;
  if (int rc = d2h(s, &v[1], s->d.ntiles, sizeof(int))) return rc;
  if (int rc = d2h(s, &v[2], s->d.nexcltiles, sizeof(int))) return rc;
  if (nb) *nb = v[0];
  if (nt) *nt = v[1];
  if (ne) *ne = v[2];
  return BM_OK;
}

int bm_get_domdec(bm_system *s, int *l2g, int *bcount, int *ti, int *tj, unsigned *tmask) {
  if (!s) return BM_EINVAL;
  int nb, nt, ne;
  if (int rc = bm_get_domdec_sizes(s, &nb, &nt, &ne)) return rc;
  std::vector<int> L((size_t)nb * blade::kBlockAtoms), I(nt), J(nt), X(nt);
  std::vector<unsigned> M((size_t)std::max(ne, 1) * blade::kBlockAtoms);
  if (int rc = d2h(s, L.data(), s->d.l2g, L.size() * sizeof(int))) return rc;
  if (nt) {
    if (int rc = d2h(s, I.data(), s->d.tile_i, nt * sizeof(int))) return rc;
    if (int rc = d2h(s, J.data(), s->d.tile_j, nt * sizeof(int))) return rc;
    if (int rc = d2h(s, X.data(), s->d.tile_ex, nt * sizeof(int))) return rc;
  }
  if (ne)
    if (int rc = d2h(s, M.data(), s->d.emask, (size_t)ne * blade::kBlockAtoms * sizeof(unsigned))) return rc;
  if (l2g) std::memcpy(l2g, L.data(), L.size() * sizeof(int));
  if (bcount)
    for (int b = 0; b < nb; b++) {
      int c = 0;
      for (int l = 0; l < blade::kBlockAtoms; l++) c += L[(size_t)b * blade::kBlockAtoms + l] >= 0;
      bcount[b] = c;
    }
  for (int t = 0; t < nt; t++) {
    if (ti) ti[t] = I[t];
    if (tj) tj[t] = J[t];
    if (tmask)
      for (int l = 0; l < blade::kBlockAtoms; l++)
        tmask[(size_t)t * blade::kBlockAtoms + l] = X[t] >= 0 ? M[(size_t)X[t] * blade::kBlockAtoms + l] : 0u;
  }
  return BM_OK;
}

}  // extern "C"
