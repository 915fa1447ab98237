"""Throughput of the per-step MSLD force + integration path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 23,558-atom MSLD box (62.23 A cube: 7,023 TIP3P waters, a
2,489-atom synthetic polymer with one lambda site of 3 x 8-atom
substituents; switch 9-10 A, PME 64^3 order 6, 2 fs, SHAKE on H bonds,
Langevin 1/ps at 298 K).  The system is generated from a fixed seed and
relaxed on the GPU before timing.  One step = list maintenance (rebuild
every 10 steps), all force terms, the VRORV update with constraints and
theta->lambda, replayed from captured hipGraphs; energies every 1000 steps.

Prints one JSON line (metric ns/day at 2 fs) with a `roofline` object for
the nbdirect kernel (FP32 VALU, 64 FLOP per in-cutoff pair; its duration
from HIP events on the system stream) and a `cpu_baseline` object (the FP64
OpenMP oracle on this host's cores, a bounded number of steps).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 64            # algorithmic FP32 work per in-cutoff pair
PEAK_FP32 = 157.3e12          # MI355X vector FP32, FLOP/s
PEAK_HBM = 8.0e12
FREQ_NRG = 1000
BATCH_STEPS, BATCH_WARMUP = 300, 30   # the default line's batch figure
PREC_STEPS, PREC_WARMUP = 1000, 100   # the default line's precision-form figure
COMB_STEPS, COMB_WARMUP = 400, 40     # the default line's combined-form figure
DHFR_STEPS, DHFR_WARMUP = 3000, 200   # the default line's DHFR deck figure
CAL_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per nbdirect launch from the committed rocprofv3 summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by a calibrated copy kernel), or
    None when there is no summary"""
    import re
    path = os.path.join(ROOT, "profiles", "traffic_msld23558.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^nbdirect_bytes_per_launch:\s*(\d+)", open(path).read(), re.M)
    return int(m.group(1)) if m else None


def profiled_valu_insts():
    """VALU wave-instructions per nbdirect launch from the committed
    rocprofv3 SQ counter summary, or None"""
    import re
    path = os.path.join(ROOT, "profiles", "valu_nbdirect_msld23558.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^nbdirect_valu_insts_per_launch:\s*(\d+)", open(path).read(), re.M)
    return int(m.group(1)) if m else None


PEAK_VALU_ISSUE = 256 * 2.4e9   # wave64 VALU instructions per second: one per CU per cycle


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="oracle steps timed for the CPU baseline (about 15 s on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="launch kernels one by one instead of replaying captured graphs "
                         "(for kernel tracers that do not follow graph launches)")
    ap.add_argument("--state", default=None,
                    help="npz file holding the relaxed state: loaded when present, written otherwise")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead: N independent MSLD systems (3,000-25,000 atoms, 2-6 substituents) "
                         "advanced by one graph launch per step, against the same systems run one by one")
    ap.add_argument("--state-dir", default=None, help="relaxed batch member states (npz per member)")
    ap.add_argument("--batch-fft", choices=("fp32", "mfma"), default="fp32",
                    help="PME FFT form of the batched members (mfma: grids rounded up to the MFMA sizes); "
                         "the one-by-one form is always the plain FP32 one")
    ap.add_argument("--fft", choices=("fp32", "mfma"), default="fp32",
                    help="PME FFT form: FP32 Stockham, or DFT contractions on the matrix cores in split-BF16x3")
    ap.add_argument("--precision", action="store_true",
                    help="print only the precision-form line (92,224 atoms, PME 96^3 on the MFMA FFT)")
    ap.add_argument("--prec-state", default=None,
                    help="relaxed precision-form state (npz): loaded if present, else written after relaxing")
    ap.add_argument("--prec-fft", choices=("both", "mfma", "fp32"), default="both",
                    help="FFT forms timed in the precision-form run (errors are measured on the MFMA one)")
    ap.add_argument("--combined", action="store_true",
                    help="print only the combined-form line (8 lambda-replicas of a 60,000-atom four-site box)")
    ap.add_argument("--no-combined", action="store_true",
                    help="leave the combined-form figure out of the default line")
    ap.add_argument("--no-precision", action="store_true",
                    help="leave the precision-form figure out of the default line")
    ap.add_argument("--no-tune", action="store_true",
                    help="no autotuning: the FFT form of --fft, the cadence of --freq-domdec, the chunk of --tile-chunk")
    ap.add_argument("--freq-domdec", type=int, default=10, help="list rebuild cadence without autotuning")
    ap.add_argument("--tile-chunk", type=int, default=16, help="tiles per nonbonded work unit without autotuning")
    ap.add_argument("--no-dhfr", action="store_true", help="skip the reference's DHFR deck figure")
    ap.add_argument("--no-batch", action="store_true",
                    help="leave out the 32-system batched figure that the default line carries")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one system on one GPU; --gpus %d is not supported "
              "(independent systems or replicas shard one per GPU with no per-step exchange)" % a.gpus,
              file=sys.stderr)
        return 2

    from blade_amd import synthetic, system, _ffi
    from blade_amd._ffi import check, ptr
    if a.batch:
        return batch_main(a)
    if a.combined:
        print(json.dumps(combined_run(a.steps, a.warmup)), flush=True)
        return 0
    if a.precision:
        print(json.dumps(precision_run(a.steps, a.warmup, graph=not a.eager, state=a.prec_state,
                                       modes={"both": (1, 0), "mfma": (1,), "fp32": (0,)}[a.prec_fft])),
              flush=True)
        return 0

    t0 = time.time()
    spec = synthetic.config2()
    spec.fft_mode = 1 if a.fft == "mfma" else 0
    spec.freq_domdec, spec.tile_chunk = a.freq_domdec, a.tile_chunk
    log("generated %s: %d atoms in %.1f s" % (spec.name, spec.n, time.time() - t0))
    if a.state and os.path.exists(a.state):
        z = np.load(a.state)
        spec.pos, spec.vel = z["pos"], z["vel"]
        spec.theta0, spec.theta_v = list(z["theta"]), list(z["theta_v"])
        log("loaded relaxed state from %s" % a.state)
    else:
        t0 = time.time()
        spec = system.relax(spec, log=log, use_graph=not a.eager)
        log("relaxed in %.1f s" % (time.time() - t0))
        if a.state:
            np.savez(a.state, pos=spec.pos, vel=spec.vel, theta=np.array(spec.theta0),
                     theta_v=np.array(spec.theta_v))

    s = system.build(spec)
    tune = None
    if not a.no_tune:
        # FFT form, list cadence and tile chunk timed from this state; the
        # chosen ones build the device system below
        t0 = time.time()
        tune = s.autotune(300)
        log("autotune in %.1f s: fft_mode %d, rebuild every %d steps, %d tiles per work unit (%.4f ms/step)"
            % (time.time() - t0, tune["fft_mode"], tune["freq_domdec"], tune["tile_chunk"], tune["ms_per_step"]))
    s.run_energy()                       # device system + first list build
    dev = s.device()
    L = s.L
    dt_fs = spec.dt / synthetic.PICOSECOND * 1000.0

    # every step graph (rebuild / energy variants) is captured before the
    # warmup, so none is captured inside the timed steps
    graph = 0 if a.eager else 1
    if graph:
        check(L.bm_capture_graphs(dev), "graph capture")
    check(L.bm_dynamics(dev, 1, a.warmup, FREQ_NRG, graph), "warmup")
    check(L.bm_sync(dev), "sync")
    step0 = 1 + a.warmup
    t1 = time.perf_counter()
    check(L.bm_dynamics(dev, step0, a.steps, FREQ_NRG, graph), "timed steps")
    check(L.bm_sync(dev), "sync")
    t2 = time.perf_counter()
    ms_step = (t2 - t1) * 1e3 / a.steps
    nsday = (1e3 / ms_step) * dt_fs * 86400e-6
    import ctypes as C
    flag = C.c_int(-1)
    check(L.bm_get_nan_flag(dev, C.byref(flag)), "nan flag")
    if flag.value != -1:
        raise RuntimeError("NaN during the timed steps (atom %d)" % flag.value)

    # the nonbonded kernel's average launch: back-to-back launches between
    # HIP events on the stream it runs on (the roofline's time)
    tnb = np.zeros(1)
    check(L.bm_time_nbdirect(dev, 20, ptr(tnb)), "nbdirect timing")
    ms_nb = float(tnb[0])
    # per-phase times with HIP events on the system stream (eager steps)
    prof = np.zeros(3)
    check(L.bm_profile_steps(dev, step0 + a.steps, 50, ptr(prof)), "profile")
    _, ms_pme, ms_eager = prof

    # the PME phases alone (spread, FFT, gather), back to back on the stream
    pme_phases = pme_fractions(L, dev, spec.n, s)

    # HBM copy calibration: a plain copy kernel over 256 MiB
    cal = np.zeros(1)
    check(L.bm_calibration_copy(0, CAL_BYTES, 20, ptr(cal)), "calibration copy")
    copy_gbps = 2 * CAL_BYTES / (cal[0] * 1e-3) / 1e9

    # in-cutoff pair count at the current configuration
    n = spec.n
    x = np.zeros(3 * n)
    th = np.zeros(spec.nblocks)
    check(L.bm_get_positions(dev, ptr(x), ptr(th)), "positions")
    pairs = in_cutoff_pairs(s.topology().contents, x, np.array(spec.box, float), spec.rcut)
    flops = FLOP_PER_PAIR * pairs
    achieved = flops / (ms_nb * 1e-3)
    # the issue-rate view of the same launch: VALU wave-instructions per
    # launch (SQ_INSTS_VALU from the committed counter summary) over this
    # run's launch time, against one instruction per CU per cycle
    vi = profiled_valu_insts()
    valu_issue = None if vi is None else {
        "insts_per_launch": vi, "achieved_G_per_s": round(vi / (ms_nb * 1e-3) / 1e9, 1),
        "peak_G_per_s": round(PEAK_VALU_ISSUE / 1e9, 1), "frac": round(vi / (ms_nb * 1e-3) / PEAK_VALU_ISSUE, 4)}

    cpu = None
    if not a.no_cpu and a.cpu_steps > 0:
        # the CPU baseline: the FP64 oracle's own step, timed
        from oracle import oracle as orc
        top, rp = s.topology(), s.runparams()
        lam = np.zeros(spec.nblocks)
        check(L.bm_get_lambda(dev, ptr(lam)), "lambda")
        xs = np.concatenate([x, th])
        vv = np.zeros(3 * n)
        tv = np.zeros(spec.nblocks)
        check(L.bm_get_velocities(dev, ptr(vv), ptr(tv)), "velocities")
        vs = np.concatenate([vv, tv])
        lam_c = lam.copy()
        orc.step(top, rp, 0, 1, xs, vs, lam_c)          # warm the caches
        c0 = time.perf_counter()
        for k in range(a.cpu_steps):
            orc.step(top, rp, k + 1, 0, xs, vs, lam_c)
        c1 = time.perf_counter()
        sps = a.cpu_steps / (c1 - c0)
        cpu = {"value": round(sps * dt_fs * 86400e-6, 4), "unit": "ns/day", "cores": orc.threads(),
               "kind": "port",
               "sample": "%d VRORV steps of the same 23,558-atom system with the FP64 OpenMP oracle "
                         "(%.2f s per step)" % (a.cpu_steps, (c1 - c0) / a.cpu_steps)}

    line = {
        "metric": "ns/day at 2 fs (23.6k-atom MSLD box)",
        "value": round(nsday, 2),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair/spline math, fp64 integration, 64-bit fixed-point accumulation",
        "data": "synthetic (seeded generator, GPU-relaxed)",
        "config": {"workload": "config2-msld-23558", "atoms": n, "box_A": spec.box[0],
                   "cutoff_A": spec.rcut, "switch_A": spec.rswitch, "pme_grid": list(spec.grid),
                   "pme_order": spec.order, "dt_fs": dt_fs, "lambda_sites": 1, "substituents": 3,
                   "list_rebuild_steps": tune["freq_domdec"] if tune else a.freq_domdec, "energy_every": FREQ_NRG,
                   "graph": not a.eager,
                   "pme_fft": "MFMA split-BF16x3 DFT" if (tune["fft_mode"] if tune else a.fft == "mfma")
                   else "fp32 Stockham",
                   "tiles_per_work_unit": tune["tile_chunk"] if tune else a.tile_chunk},
        "autotune": tune,
        "pair_interactions_per_s": round(pairs / (ms_step * 1e-3), 1),
        "kernel_ms": {"nbdirect": round(ms_nb, 5), "pme": round(ms_pme, 5), "eager_step": round(ms_eager, 5)},
        "hbm_copy_GBps": round(copy_gbps, 1),
        "pme_phases": pme_phases,
        "roofline": {"kernel": "nbdirect", "bound": "valu", "achieved": round(achieved / 1e12, 4),
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32, 5),
                     "traffic": profiled_traffic(),
                     "note": "FP32 vector-ALU bound (VALU, not the matrix cores): the peak is MI355X's "
                             "published vector FP32 figure",
                     "work": "%d in-cutoff pairs x %d FLOP per launch" % (pairs, FLOP_PER_PAIR),
                     "valu_issue": valu_issue},
        "cpu_baseline": cpu,
    }
    s.close()
    if not a.no_batch:
        # the metric's second figure: 32 uneven MSLD systems advanced together
        # (one launch per kernel type and step) against the same systems one
        # by one, both timed over the same steps
        try:
            bl = batch_run(32, BATCH_STEPS, BATCH_WARMUP, a.state_dir, autotune=not a.no_tune)
            line["batch_32"] = {"metric": bl["metric"], "value": bl["value"], "unit": "ns/day",
                                "ms_per_step": bl["ms_per_step"], "steps": bl["steps"],
                                "one_by_one": bl["one_by_one"]["value"],
                                "batched_over_one_by_one": bl["batched_over_one_by_one"],
                                "config": bl["config"]}
        except Exception as e:   # the main figure stands; say why the batch one is missing
            line["batch_32"] = {"error": str(e)[:300]}
    if not a.no_dhfr:
        try:
            line["dhfr_deck"] = dhfr_run(DHFR_STEPS, DHFR_WARMUP)
        except Exception as e:
            line["dhfr_deck"] = {"error": str(e)[:300]}
    if not a.no_precision:
        try:
            line["precision_92224"] = precision_run(PREC_STEPS, PREC_WARMUP)
        except Exception as e:
            line["precision_92224"] = {"error": str(e)[:300]}
    if not a.no_combined:
        try:
            line["combined_8x60000"] = combined_run(COMB_STEPS, COMB_WARMUP)
        except Exception as e:
            line["combined_8x60000"] = {"error": str(e)[:300]}
    print(json.dumps(line), flush=True)
    return 0


def _timed_ms(s, steps, warmup, graph=True):
    """ms per step of `steps` steps (graph-replayed, or eager launches)
    after `warmup` more"""
    from blade_amd._ffi import check
    L, dev = s.L, s.device()
    if graph:
        check(L.bm_capture_graphs(dev), "graph capture")
    check(L.bm_dynamics(dev, 1, warmup, FREQ_NRG, int(graph)), "warmup")
    check(L.bm_sync(dev), "sync")
    t1 = time.perf_counter()
    check(L.bm_dynamics(dev, 1 + warmup, steps, FREQ_NRG, int(graph)), "timed steps")
    check(L.bm_sync(dev), "sync")
    return (time.perf_counter() - t1) * 1e3 / steps


def dhfr_run(steps, warmup):
    """The reference's own benchmark deck, test/dhfr/input (DHFR in water,
    23,558 atoms, CHARMM22, rcut 9 A with a 7.5 A switch, PME order 4 on
    64^3, gamma 5/ps, SHAKE tolerance 1e-8), set up by the script
    interpreter from the committed fixtures exactly as written (its
    500,000-step run and output files left out), then graph-replayed and
    timed like the main figure, at the deck's settings (no tuning)."""
    import gzip
    import shutil
    import tempfile
    from blade_amd import system
    from blade_amd._ffi import check
    gold = os.path.join(ROOT, "tests", "golden", "dhfr")
    tmp = tempfile.mkdtemp(prefix="dhfr")
    cwd = os.getcwd()
    try:
        t = os.path.join(tmp, "test", "dhfr")
        run = os.path.join(tmp, "run")
        os.makedirs(t)
        os.makedirs(run)
        for f in ("5dfr.psf", "5dfr.crd", "par_all22_prot.inp"):
            with gzip.open(os.path.join(gold, f + ".gz"), "rb") as src, open(os.path.join(t, f), "wb") as dst:
                shutil.copyfileobj(src, dst)
        deck = open(os.path.join(gold, "input")).read()
        k = deck.index("run dynamics")
        setup = deck[:k] + ("run setvariable nsteps 1\nrun setvariable freqxtc 0\nrun setvariable freqnrg 0\n"
                            "run setvariable freqlmd 0\nrun setvariable fnmcpo {}\nrun dynamics\n"
                            .format(os.path.join(tmp, "start.cpt")))
        open(os.path.join(run, "setup.inp"), "w").write(setup)
        os.chdir(run)
        t0 = time.time()
        s = system.BladeSystem((0,))
        s.L.blade_interpretter(b"setup.inp", s.h)
        err = s.L.blade_system_error(s.h).decode()
        if err:
            raise RuntimeError(err)
        log("dhfr deck set up in %.1f s" % (time.time() - t0))
        L, dev = s.L, s.device()
        rp = s.runparams().contents
        check(L.bm_capture_graphs(dev), "graph capture")
        check(L.bm_dynamics(dev, 1, warmup, FREQ_NRG, 1), "warmup")
        check(L.bm_sync(dev), "sync")
        t1 = time.perf_counter()
        check(L.bm_dynamics(dev, 1 + warmup, steps, FREQ_NRG, 1), "timed steps")
        check(L.bm_sync(dev), "sync")
        ms = (time.perf_counter() - t1) * 1e3 / steps
        s.close()
        return {"metric": "ns/day at 2 fs, the reference's DHFR benchmark deck (test/dhfr/input)",
                "value": round(0.002 * 86400.0 / ms, 2), "unit": "ns/day", "ms_per_step": round(ms, 5),
                "steps": steps, "warmup": warmup, "atoms": 23558,
                "config": {"rcut_A": rp.rcut, "switch_A": rp.rswitch, "pme_grid": list(rp.grid),
                           "pme_order": rp.order, "list_rebuild_steps": rp.freq_domdec or 10,
                           "energy_every": FREQ_NRG}}
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp, ignore_errors=True)


PEAK_BF16 = 2.5e15      # dense BF16 MFMA FLOP/s


def pme_fractions(L, dev, n, s, reps=20):
    """the PME phases timed alone (bm_time_pme) against their bounds:
    spread and gather against HBM with the algorithmic bytes N*(16+4) +
    G*4 and N*20 + G*4 + N*24; the FFT's useful work 2 * 2.5 G log2 G FLOP
    against the FP32 vector peak (FP32 FFT) or the dense BF16 MFMA peak
    (MFMA FFT, which also reports its executed work, 3 split products of
    8 K FLOP per grid point and complex pass, the real-data z passes at
    half)"""
    from blade_amd._ffi import check, ptr
    rp = s.runparams().contents
    K = [int(rp.grid[0]), int(rp.grid[1]), int(rp.grid[2])]
    G = K[0] * K[1] * K[2]
    t = np.zeros(3)
    check(L.bm_time_pme(dev, reps, ptr(t)), "PME phase timing")
    ms_sp, ms_fft, ms_ga = (float(v) for v in t)
    sp_bytes = n * 20 + G * 4
    ga_bytes = n * 20 + G * 4 + n * 24
    useful = 2 * 2.5 * G * np.log2(G)
    mfma = int(rp.fft_mode) == 1
    out = {
        "spread": {"ms": round(ms_sp, 5), "bytes": sp_bytes, "GBps": round(sp_bytes / (ms_sp * 1e-3) / 1e9, 1),
                   "frac_hbm": round(sp_bytes / (ms_sp * 1e-3) / PEAK_HBM, 5)},
        "gather": {"ms": round(ms_ga, 5), "bytes": ga_bytes, "GBps": round(ga_bytes / (ms_ga * 1e-3) / 1e9, 1),
                   "frac_hbm": round(ga_bytes / (ms_ga * 1e-3) / PEAK_HBM, 5)},
        "fft": {"ms": round(ms_fft, 5), "form": "MFMA split-BF16x3 DFT" if mfma else "fp32 Stockham",
                "useful_flop": int(useful), "useful_TFLOPs": round(useful / (ms_fft * 1e-3) / 1e12, 3),
                "peak_TFLOPs": (PEAK_BF16 if mfma else PEAK_FP32) / 1e12,
                "frac_useful": round(useful / (ms_fft * 1e-3) / (PEAK_BF16 if mfma else PEAK_FP32), 6)},
        "grid": K,
    }
    if mfma:
        executed = 3 * 8 * G * (2 * K[0] + 2 * K[1] + K[2])
        out["fft"]["executed_flop"] = int(executed)
        out["fft"]["frac_executed"] = round(executed / (ms_fft * 1e-3) / PEAK_BF16, 5)
    return out


def in_cutoff_pairs(t, x, box, rcut):
    """non-excluded atom pairs closer than rcut in the minimum image: a
    periodic k-d tree count of all pairs, less the excluded pairs inside
    the cutoff"""
    from scipy.spatial import cKDTree
    n = t.n_atoms
    p = np.ascontiguousarray(x.reshape(3, n).T) % box
    p[p >= box] = 0.0
    tree = cKDTree(p, boxsize=box)
    allp = (int(tree.count_neighbors(tree, rcut)) - n) // 2
    ex = t.array("excl_idx", t.n_excl, 2)
    d = p[ex[:, 0]] - p[ex[:, 1]]
    d -= box * np.round(d / box)
    return allp - int(np.sum(np.einsum("ij,ij->i", d, d) < rcut * rcut))


def precision_run(steps, warmup, graph=True, state=None, modes=(1, 0)):
    """The precision form: the 92,224-atom box (two lambda sites x 4
    substituents, PME 96^3 order 6) with the PME FFT as split-BF16x3 DFT
    contractions on the matrix cores, GPU-relaxed, timed like the main
    figure, and the same state with the FP32 FFT beside it.  (The MFMA
    form's errors against the FP64 oracle on this box are measured by the
    GPU tests.)"""
    import copy
    from blade_amd import synthetic, system
    t0 = time.time()
    spec = synthetic.config3()
    log("generated %s: %d atoms in %.1f s" % (spec.name, spec.n, time.time() - t0))
    if state and os.path.exists(state):
        z = np.load(state)
        spec.pos, spec.vel = z["pos"], z["vel"]
        spec.theta0, spec.theta_v = list(z["theta"]), list(z["theta_v"])
        log("loaded relaxed state from %s" % state)
    else:
        t0 = time.time()
        spec = system.relax(spec, use_graph=graph)
        log("relaxed in %.1f s" % (time.time() - t0))
        if state:
            np.savez(state, pos=spec.pos, vel=spec.vel, theta=np.array(spec.theta0),
                     theta_v=np.array(spec.theta_v))
    dt_fs = spec.dt / synthetic.PICOSECOND * 1000.0
    out, phases = {}, {}
    for mode in modes:
        sp = copy.copy(spec)
        sp.fft_mode = mode
        s = system.build(sp)
        s.run_energy()
        ms = _timed_ms(s, steps, warmup, graph)
        phases[mode] = pme_fractions(s.L, s.device(), sp.n, s)
        s.close()
        out[mode] = ms
    nsday = lambda m: round((1e3 / out[m]) * dt_fs * 86400e-6, 2) if m in out else None
    ms = out[modes[0]]
    return {"metric": "ns/day at 2 fs (92,224-atom precision-form box, %s FFT)" % ("MFMA" if modes[0] else "FP32"),
            "value": nsday(modes[0]), "unit": "ns/day", "ms_per_step": round(ms, 5),
            "steps": steps, "warmup": warmup, "fp32_fft_value": nsday(0),
            "pme_phases": phases.get(modes[0]), "pme_phases_fp32_fft": phases.get(0) if modes[0] else None,
            "config": {"workload": spec.name, "atoms": spec.n, "box_A": list(spec.box), "lambda_sites": 2,
                       "substituents": [4, 4], "pme_grid": list(spec.grid), "pme_order": spec.order,
                       "pme_fft": "MFMA split-BF16x3 DFT" if modes[0] else "fp32 Stockham",
                       "cutoff_A": spec.rcut, "switch_A": spec.rswitch,
                       "dt_fs": dt_fs}}


def _member(args):
    from blade_amd import synthetic
    natoms, nsubs, seed = args
    return synthetic.batch_member(seed, natoms, nsubs)


def _relaxed(spec, path, log_fn):
    from blade_amd import system
    if path and os.path.exists(path):
        z = np.load(path)
        spec.pos, spec.vel = z["pos"], z["vel"]
        spec.theta0, spec.theta_v = list(z["theta"]), list(z["theta_v"])
        return spec
    spec = system.relax(spec)
    if path:
        np.savez(path, pos=spec.pos, vel=spec.vel, theta=np.array(spec.theta0), theta_v=np.array(spec.theta_v))
    return spec


def batch_main(a):
    """print the batched line (--batch N)"""
    print(json.dumps(batch_run(a.batch, a.steps, a.warmup, a.state_dir, autotune=not a.no_tune,
                               fft=a.batch_fft)), flush=True)
    return 0


MFMA_GRIDS = (16, 32, 48, 64, 80, 96, 128)


def batch_run(nsys, steps, warmup, state_dir=None, autotune=True, fft="fp32"):
    """aggregate ns/day of nsys systems: one graph launch per step for all
    of them, against the same systems advanced one after another in the
    plain per-system FP32-FFT form.  fft="mfma": the batched members run
    the PME FFT on the matrix cores, each grid edge rounded up to an MFMA
    size (a finer grid than the plain form's)."""
    import copy
    import concurrent.futures as cf
    from blade_amd import synthetic, system
    from blade_amd._ffi import check
    members = synthetic.batch_set(nsys)
    t0 = time.time()
    with cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        specs = list(ex.map(_member, members))
    log("generated %d systems (%d atoms) in %.1f s" % (len(specs), sum(s.n for s in specs), time.time() - t0))
    if state_dir:
        os.makedirs(state_dir, exist_ok=True)
    t0 = time.time()
    specs = [_relaxed(sp, os.path.join(state_dir, "m%02d.npz" % k) if state_dir else None, log)
             for k, sp in enumerate(specs)]
    log("relaxed in %.1f s" % (time.time() - t0))
    dt_fs = specs[0].dt / synthetic.PICOSECOND * 1000.0
    # one by one: the plain per-system form (default cadence 10, 16 tiles
    # per work unit), every member's own graphs, members in turn
    systems = [system.build(sp) for sp in specs]
    for s in systems:
        s.run_energy()
    for s in systems:
        check(s.L.bm_capture_graphs(s.device()), "graph capture")
        check(s.L.bm_dynamics(s.device(), 1, warmup, FREQ_NRG, 1), "warmup")
        check(s.L.bm_sync(s.device()), "sync")
    t1 = time.perf_counter()
    for s in systems:   # each member alone: finish it before the next starts
        check(s.L.bm_dynamics(s.device(), 1 + warmup, steps, FREQ_NRG, 1), "plain steps")
        check(s.L.bm_sync(s.device()), "sync")
    t_plain = time.perf_counter() - t1
    for s in systems:
        s.close()
    # batched: the same starting states, one launch per kernel type and step
    # for every member, with the cadence and tile chunk timed on the batch
    bspecs = specs
    if fft == "mfma":
        bspecs = []
        for sp in specs:
            q = copy.copy(sp)
            q.fft_mode = 1
            q.grid = tuple(next(g for g in MFMA_GRIDS if g >= int(np.ceil(L / sp.grid_space))) for L in sp.box)
            bspecs.append(q)
    systems = [system.build(sp) for sp in bspecs]
    tune = None
    if autotune:
        t0 = time.time()
        tune = system.autotune_batch(systems, 150)
        log("batch autotune in %.1f s: rebuild every %d steps, %d tiles per work unit"
            % (time.time() - t0, tune["freq_domdec"], tune["tile_chunk"]))
    for s in systems:
        s.run_energy()
    b = system.Batch(systems)
    b.capture_graphs()
    b.dynamics(1, warmup, FREQ_NRG)
    b.sync()
    t1 = time.perf_counter()
    b.dynamics(1 + warmup, steps, FREQ_NRG)
    b.sync()
    t_batch = time.perf_counter() - t1
    n = len(systems)
    agg = lambda t: n * steps / t * dt_fs * 86400e-6
    line = {
        "metric": "aggregate ns/day at 2 fs, %d batched MSLD systems" % n,
        "value": round(agg(t_batch), 2), "unit": "ns/day", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(t_batch * 1e3 / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 pair/spline math, fp64 integration, 64-bit fixed-point accumulation",
        "data": "synthetic (seeded generator, GPU-relaxed)",
        "config": {"workload": "batch-%d-msld" % n, "atoms_total": sum(s.n for s in specs),
                   "pme_fft_batched": "MFMA split-BF16x3 DFT, grids rounded up to MFMA sizes" if fft == "mfma"
                   else "fp32 Stockham",
                   "atoms_range": [min(s.n for s in specs), max(s.n for s in specs)],
                   "substituents_range": [min(m[1] for m in members), max(m[1] for m in members)],
                   "cutoff_A": 10.0, "grid_spacing_A": 1.0, "graph": True,
                   "list_rebuild_steps": tune["freq_domdec"] if tune else 10,
                   "tiles_per_work_unit": tune["tile_chunk"] if tune else 16},
        "autotune": tune,
        "one_by_one": {"value": round(agg(t_plain), 2), "ms_per_step_all": round(t_plain * 1e3 / steps, 5)},
        "batched_over_one_by_one": round(t_plain / t_batch, 3),
    }
    b.close()
    for s in systems:
        s.close()
    return line



def combined_run(steps, warmup, nrep=8, frames=20, frame_every=10):
    """The combined form: nrep lambda-replicas of the 60,000-atom box with
    four lambda sites of 4-8 substituents (27 blocks; soft core, atom
    restraints, per-block fixed biases, PME 80^3 on the MFMA FFT), the
    structure relaxed once on the GPU and each replica started from its own
    theta, biases and velocities.  Timed one by one (each replica's own
    graphs, in turn) and batched (one graph launch per step for all), the
    same steps; then frames every frame_every steps give per-block lambda
    means over all replicas.  (The production run with lambda histograms
    and oracle checks of sampled frames is tests/production_c5.py.)"""
    from blade_amd import synthetic, system
    t0 = time.time()
    base = synthetic.config5(0)
    log("generated %s: %d atoms, %d blocks in %.1f s" % (base.name, base.n, base.nblocks, time.time() - t0))
    t0 = time.time()
    base = system.relax(base)
    log("relaxed in %.1f s" % (time.time() - t0))
    specs = [synthetic.replica_of(base, r) for r in range(nrep)]
    dt_fs = base.dt / synthetic.PICOSECOND * 1000.0

    def timed(run_one_by_one):
        systems = [system.build(sp) for sp in specs]
        for s in systems:
            s.run_energy()
        if run_one_by_one:
            for s in systems:
                check(s.L.bm_capture_graphs(s.device()), "graph capture")
                check(s.L.bm_dynamics(s.device(), 1, warmup, FREQ_NRG, 1), "warmup")
                check(s.L.bm_sync(s.device()), "sync")
            t1 = time.perf_counter()
            for s in systems:
                check(s.L.bm_dynamics(s.device(), 1 + warmup, steps, FREQ_NRG, 1), "plain steps")
                check(s.L.bm_sync(s.device()), "sync")
            dt = time.perf_counter() - t1
            for s in systems:
                s.close()
            return dt, None, None
        b = system.Batch(systems)
        b.capture_graphs()
        b.dynamics(1, warmup, FREQ_NRG)
        b.sync()
        t1 = time.perf_counter()
        b.dynamics(1 + warmup, steps, FREQ_NRG)
        b.sync()
        return time.perf_counter() - t1, b, systems

    from blade_amd._ffi import check
    t_plain, _, _ = timed(True)
    t_batch, b, systems = timed(False)
    # lambda frames: every frame_every steps, all replicas
    lam = []
    step = 1 + warmup + steps
    for f in range(frames):
        b.dynamics(step, frame_every, 0)
        b.sync()
        step += frame_every
        lam += [s.lambdas() for s in systems]
    lam = np.array(lam)
    b.close()
    for s in systems:
        s.close()
    agg = lambda t: nrep * steps / t * dt_fs * 86400e-6
    return {"metric": "aggregate ns/day at 2 fs, %d lambda-replicas of the 60,000-atom four-site box" % nrep,
            "value": round(agg(t_batch), 2), "unit": "ns/day", "ms_per_step": round(t_batch * 1e3 / steps, 5),
            "steps": steps, "warmup": warmup,
            "one_by_one": round(agg(t_plain), 2), "batched_over_one_by_one": round(t_plain / t_batch, 3),
            "lambda_mean_per_block": [round(float(m), 3) for m in lam.mean(axis=0)],
            "lambda_frames": int(lam.shape[0]),
            "config": {"workload": "config5-combined-8x60000", "atoms_per_replica": base.n, "replicas": nrep,
                       "lambda_sites": 4, "blocks": base.nblocks, "box_A": list(base.box),
                       "pme_grid": list(base.grid), "pme_fft": "MFMA split-BF16x3 DFT",
                       "cutoff_A": base.rcut, "switch_A": base.rswitch, "dt_fs": dt_fs}}

if __name__ == "__main__":
    sys.exit(main())
