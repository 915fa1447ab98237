// autotune.cpp - the step autotuner (bm_autotune).  The reference fixes its
// list cadence and kernel shapes; here the fastest of a few candidate
// configurations is measured on the system itself: PME FFT form, neighbour
// list rebuild cadence and directed tiles per nonbonded work unit.  It uses
// only the public system API: every candidate is a fresh device system made
// from the caller's topology, run parameters and state.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/blade_mi355x.h"

namespace {

struct Tuner {
  int device;
  const bm_topology *top;
  const bm_runparams *rp0;
  const double *pos, *vel, *theta, *theta_vel;
  int steps;
  bm_tune_result *out;

  const bm_tune_trial *find(int fft, int freq, int chunk) const {
    for (int i = 0; i < out->n_trials; i++) {
      const bm_tune_trial &t = out->trial[i];
      if (t.fft_mode == fft && t.freq_domdec == freq && t.tile_chunk == chunk) return &t;
    }
    return nullptr;
  }

  // ms per step of one candidate (infinite when it fails)
  double run(int fft, int freq, int chunk) {
    if (const bm_tune_trial *t = find(fft, freq, chunk)) return t->status ? INFINITY : t->ms_per_step;
    bm_runparams r = *rp0;
    r.fft_mode = fft;
    r.freq_domdec = freq;
    r.tile_chunk = chunk;
    // the list pad is proportional to the cadence (domdec/domdec.cu:138)
    const int f0 = rp0->freq_domdec > 0 ? rp0->freq_domdec : 10;
    r.cull_pad = rp0->cull_pad * freq / f0;
    double ms = INFINITY;
    bm_system *s = nullptr;
    int rc = bm_system_create(device, top, &r, pos, vel, theta, theta_vel, &s);
    if (!rc) {
      // step 0 rebuilds; the warmup captures both step graphs, and the
      // timed steps are whole rebuild periods
      const long warm = 2L * freq;
      const long n = (long)((steps + freq - 1) / freq) * freq;
      rc = bm_dynamics(s, 0, warm, 0, 1);
      if (!rc) rc = bm_sync(s);
      const auto t0 = std::chrono::steady_clock::now();
      if (!rc) rc = bm_dynamics(s, warm, n, 0, 1);
      if (!rc) rc = bm_sync(s);
      const auto t1 = std::chrono::steady_clock::now();
      int flag = -1;
      if (!rc && bm_get_nan_flag(s, &flag) == BM_OK && flag != -1) rc = BM_ENAN;
      if (!rc) ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / n;
      bm_system_destroy(s);
    }
    if (out->n_trials < BM_TUNE_MAX_TRIALS) {
      bm_tune_trial &t = out->trial[out->n_trials++];
      t.fft_mode = fft;
      t.freq_domdec = freq;
      t.tile_chunk = chunk;
      t.status = rc;
      t.ms_per_step = rc ? 0.0 : ms;
    }
    return ms;
  }
};

// batched candidates: every member with the same cadence and tile chunk
struct BatchTuner {
  int device, n;
  const bm_topology *const *tops;
  const bm_runparams *const *rps;
  const double *const *pos, *const *vel, *const *theta, *const *theta_vel;
  int steps;
  bm_tune_result *out;

  double run(int freq, int chunk) {
    for (int i = 0; i < out->n_trials; i++) {
      const bm_tune_trial &t = out->trial[i];
      if (t.freq_domdec == freq && t.tile_chunk == chunk) return t.status ? INFINITY : t.ms_per_step;
    }
    std::vector<bm_system *> sys(n, nullptr);
    int rc = BM_OK;
    for (int k = 0; k < n && !rc; k++) {
      bm_runparams r = *rps[k];
      const int f0 = r.freq_domdec > 0 ? r.freq_domdec : 10;
      r.cull_pad = r.cull_pad * freq / f0;
      r.freq_domdec = freq;
      r.tile_chunk = chunk;
      rc = bm_system_create(device, tops[k], &r, pos[k], vel[k], theta[k], theta_vel[k], &sys[k]);
    }
    double ms = INFINITY;
    bm_batch *b = nullptr;
    if (!rc) rc = bm_batch_create(n, sys.data(), &b);
    if (!rc) {
      const long warm = 2L * freq;
      const long m = (long)((steps + freq - 1) / freq) * freq;
      rc = bm_batch_dynamics(b, 0, warm, 0);
      if (!rc) rc = bm_batch_sync(b);
      const auto t0 = std::chrono::steady_clock::now();
      if (!rc) rc = bm_batch_dynamics(b, warm, m, 0);
      if (!rc) rc = bm_batch_sync(b);
      const auto t1 = std::chrono::steady_clock::now();
      for (int k = 0; k < n && !rc; k++) {
        int flag = -1;
        if (bm_get_nan_flag(sys[k], &flag) == BM_OK && flag != -1) rc = BM_ENAN;
      }
      if (!rc) ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / m;
    }
    if (b) bm_batch_destroy(b);
    for (bm_system *s : sys)
      if (s) bm_system_destroy(s);
    if (out->n_trials < BM_TUNE_MAX_TRIALS) {
      bm_tune_trial &t = out->trial[out->n_trials++];
      t.fft_mode = -1;
      t.freq_domdec = freq;
      t.tile_chunk = chunk;
      t.status = rc;
      t.ms_per_step = rc ? 0.0 : ms;
    }
    return ms;
  }
};

bool mfma_grid(const bm_runparams *rp) {
  if (rp->elec_method != BM_PME) return false;
  for (int k = 0; k < 3; k++)
    if (rp->grid[k] < 16 || rp->grid[k] > 128 || rp->grid[k] % 16) return false;
  return true;
}

}  // namespace

extern "C" int bm_autotune(int device, const bm_topology *top, const bm_runparams *rp, const double *pos,
                           const double *vel, const double *theta, const double *theta_vel, int steps,
                           bm_tune_result *out) {
  if (!top || !rp || !pos || !vel || !out || steps < 1) return BM_EINVAL;
  *out = bm_tune_result{};
  Tuner T{device, top, rp, pos, vel, theta, theta_vel, steps, out};
  int fft = rp->fft_mode, freq = rp->freq_domdec > 0 ? rp->freq_domdec : 10;
  int chunk = rp->tile_chunk > 0 ? rp->tile_chunk : 16;
  double best = T.run(fft, freq, chunk);
  auto consider = [&](int f, int q, int c) {
    const double ms = T.run(f, q, c);
    if (ms < best) {
      best = ms;
      fft = f;
      freq = q;
      chunk = c;
    }
  };
  if (rp->elec_method == BM_PME) {
    consider(0, freq, chunk);
    if (mfma_grid(rp)) consider(1, freq, chunk);
  }
  const int f1 = fft, c1 = chunk;
  for (int q : {5, 10, 15, 20}) consider(f1, q, c1);
  const int f2 = fft, q2 = freq;
  for (int c : {4, 8, 16}) consider(f2, q2, c);
  if (!std::isfinite(best)) return out->trial[0].status ? out->trial[0].status : BM_EINVAL;
  out->fft_mode = fft;
  out->freq_domdec = freq;
  out->tile_chunk = chunk;
  out->ms_per_step = best;
  return BM_OK;
}

extern "C" int bm_batch_autotune(int device, int n, const bm_topology *const *tops, const bm_runparams *const *rps,
                                 const double *const *pos, const double *const *vel, const double *const *theta,
                                 const double *const *theta_vel, int steps, bm_tune_result *out) {
  if (n < 1 || !tops || !rps || !pos || !vel || !theta || !theta_vel || !out || steps < 1) return BM_EINVAL;
  *out = bm_tune_result{};
  BatchTuner T{device, n, tops, rps, pos, vel, theta, theta_vel, steps, out};
  int freq = rps[0]->freq_domdec > 0 ? rps[0]->freq_domdec : 10;
  int chunk = rps[0]->tile_chunk > 0 ? rps[0]->tile_chunk : 16;
  double best = T.run(freq, chunk);
  auto consider = [&](int q, int c) {
    const double ms = T.run(q, c);
    if (ms < best) {
      best = ms;
      freq = q;
      chunk = c;
    }
  };
  const int q1 = freq;
  for (int c : {16, 8, 4}) consider(q1, c);
  const int c2 = chunk;
  for (int q : {10, 15, 20}) consider(q, c2);
  if (!std::isfinite(best)) return out->trial[0].status ? out->trial[0].status : BM_EINVAL;
  out->fft_mode = -1;
  out->freq_domdec = freq;
  out->tile_chunk = chunk;
  out->ms_per_step = best;
  return BM_OK;
}
