"""Production run of the combined workload (8 lambda-replicas of the
60,000-atom four-site MSLD box, batched on one GPU), with lambda
histograms and FP64 oracle checks at sampled frames.

    python tests/production_c5.py --steps 100000 > production_c5.md

The structure is relaxed once, every replica starts from its own theta,
fixed biases and velocities, and the batch advances all replicas with one
graph launch per step.  Every `--sample` steps the lambdas of all replicas
are read; at `--frames` evenly spaced steps one replica's state (the
replicas taken in turn) is rebuilt as a fresh device system and its forces,
energy and dU/dlambda are compared with the oracle.  The report (markdown,
on stdout) gives the throughput, the lambda histograms per site, the
fraction of samples with each block at lambda > 0.99 (the reference
analysis' physical-state cut), transitions between such states, and the
oracle errors per frame.  Progress lines go to stderr.

Not collected by pytest: a measurement, run on the GPU by hand.
"""
import argparse
import copy
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from blade_amd import synthetic, system  # noqa: E402
from oracle import oracle as orc  # noqa: E402

E_POT = 20


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def oracle_frame(spec, state):
    """device and oracle forces, energy and dU/dlambda at `state`"""
    x, v, th, tv = state
    sp = copy.copy(spec)
    sp.pos, sp.vel, sp.theta0, sp.theta_v = x, v, list(th[:spec.nblocks]), list(tv[:spec.nblocks])
    s = system.build(sp)
    s.run_energy()
    top, rp = s.topology(), s.runparams()
    lam = orc.theta_to_lambda(top, rp, np.array(sp.theta0, float))
    t0 = time.time()
    fo, do, eo, _ = orc.forces(top, rp, s.positions(), lam)
    t_orc = time.time() - t0
    fg, eg, dg = s.forces(), s.energy(), s.dudl()
    s.close()
    return {"force_rel_rms": float(np.sqrt(np.mean((fg - fo) ** 2)) / np.sqrt(np.mean(fo ** 2))),
            "energy_rel": float(abs(eg[E_POT] - eo[E_POT]) / abs(eo[E_POT])),
            "dudl_max_abs": float(np.max(np.abs(dg - do))),
            "epot": float(eo[E_POT]), "oracle_s": t_orc}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--sample", type=int, default=100, help="steps between lambda samples")
    ap.add_argument("--frames", type=int, default=10, help="oracle-checked frames")
    ap.add_argument("--replicas", type=int, default=8)
    a = ap.parse_args(argv)
    if a.steps % a.sample or (a.steps // a.sample) % a.frames:
        raise SystemExit("--steps must be a multiple of --sample * --frames")
    t0 = time.time()
    base = synthetic.config5(0)
    log("generated %s: %d atoms, %d blocks in %.1f s" % (base.name, base.n, base.nblocks, time.time() - t0))
    base = system.relax(base)
    specs = [synthetic.replica_of(base, r) for r in range(a.replicas)]
    systems = [system.build(sp) for sp in specs]
    for s in systems:
        s.run_energy()
    b = system.Batch(systems)
    b.capture_graphs()
    dt_fs = base.dt / synthetic.PICOSECOND * 1000.0
    nb = base.nblocks
    block_site = np.array(base.block_site)
    samples = []
    frames = []
    per_frame = a.steps // a.frames
    step = 1
    t_run = 0.0
    log("running %d steps, %d replicas, lambda every %d steps" % (a.steps, a.replicas, a.sample))
    for f in range(a.frames):
        for _ in range(per_frame // a.sample):
            t1 = time.perf_counter()
            b.dynamics(step, a.sample, 0)
            b.sync()
            t_run += time.perf_counter() - t1
            step += a.sample
            samples.append([s.lambdas()[:nb] for s in systems])
        r = f % a.replicas
        st = system.state(systems[r], base.n, nb)
        res = oracle_frame(specs[r], st)
        res.update(step=step - 1, replica=r)
        frames.append(res)
        log("step %d: %.1f ns/day aggregate so far; frame replica %d force relRMS %.2e energy rel %.2e "
            "dU/dl %.2e (oracle %.1f s)" % (step - 1, a.replicas * (step - 1) / t_run * dt_fs * 86400e-6, r,
                                            res["force_rel_rms"], res["energy_rel"], res["dudl_max_abs"],
                                            res["oracle_s"]))
    b.close()
    for s in systems:
        s.close()
    lam = np.array(samples)                      # [samples, replicas, blocks]
    nsamp = lam.shape[0]
    agg = a.replicas * a.steps / t_run * dt_fs * 86400e-6

    out = []
    out.append("# Production run: 8 lambda-replicas of the 60,000-atom four-site MSLD box")
    out.append("")
    out.append("Command (on the GPU):")
    out.append("")
    out.append("    python tests/production_c5.py --steps %d --sample %d --frames %d" % (a.steps, a.sample, a.frames))
    out.append("")
    out.append("| quantity | value |")
    out.append("|---|---|")
    out.append("| atoms per replica | %d |" % base.n)
    out.append("| blocks (4 sites + environment) | %d |" % nb)
    out.append("| replicas | %d |" % a.replicas)
    out.append("| steps per replica | %d (%.1f ps at %.1f fs) |" % (a.steps, a.steps * dt_fs / 1000.0, dt_fs))
    out.append("| timed dynamics (graph-replayed batch) | %.1f s |" % t_run)
    out.append("| aggregate throughput | %.1f ns/day |" % agg)
    out.append("| lambda samples per replica | %d (every %d steps) |" % (nsamp, a.sample))
    out.append("")
    out.append("## Lambda populations (fraction of samples with lambda > 0.99, all replicas)")
    out.append("")
    out.append("| site | block | P(lambda > 0.99) | mean lambda | transitions into lambda > 0.99 |")
    out.append("|---|---|---|---|---|")
    phys = lam > 0.99
    for b_ in range(1, nb):
        ph = phys[:, :, b_]
        into = int(np.sum(ph[1:] & ~ph[:-1]))
        out.append("| %d | %d | %.3f | %.3f | %d |" % (block_site[b_], b_, ph.mean(), lam[:, :, b_].mean(), into))
    out.append("")
    out.append("Per site, the fraction of samples with some block above 0.99 (the rest are intermediate):")
    out.append("")
    out.append("| site | substituents | P(one block > 0.99) |")
    out.append("|---|---|---|")
    for site in sorted(set(block_site[1:])):
        bl = [b_ for b_ in range(1, nb) if block_site[b_] == site]
        anyp = phys[:, :, bl].any(axis=2).mean()
        out.append("| %d | %d | %.3f |" % (site, len(bl), anyp))
    out.append("")
    out.append("## Lambda histogram (all substituent blocks, all replicas, 10 bins)")
    out.append("")
    h, edges = np.histogram(lam[:, :, 1:].ravel(), bins=10, range=(0.0, 1.0))
    out.append("| bin | count | fraction |")
    out.append("|---|---|---|")
    for k in range(10):
        out.append("| %.1f-%.1f | %d | %.4f |" % (edges[k], edges[k + 1], h[k], h[k] / h.sum()))
    out.append("")
    out.append("## Sampled frames against the FP64 oracle")
    out.append("")
    out.append("Each frame's replica state is rebuilt as a fresh device system (lists rebuilt, sites placed); bars: "
               "forces 1e-5 relative RMS, energy 1e-6 relative, dU/dlambda 1e-3 kcal/mol.")
    out.append("")
    out.append("| step | replica | force rel RMS | energy rel | max abs dU/dlambda (kcal/mol) | Epot (kcal/mol) |")
    out.append("|---|---|---|---|---|---|")
    ok = True
    for fr in frames:
        out.append("| %d | %d | %.2e | %.2e | %.2e | %.1f |" % (fr["step"], fr["replica"], fr["force_rel_rms"],
                                                              fr["energy_rel"], fr["dudl_max_abs"], fr["epot"]))
        ok = ok and fr["force_rel_rms"] <= 1e-5 and fr["energy_rel"] <= 1e-6 and fr["dudl_max_abs"] <= 1e-3
    out.append("")
    out.append("All frames within the bars: %s" % ("yes" if ok else "NO"))
    print("\n".join(out), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
