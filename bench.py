#!/usr/bin/env python3
"""Benchmark of the MDACP LJ step on one MI355X.

Workload (the configuration the metric is quoted on): 4,000,000 particles on
an FCC lattice at density 0.5 in a periodic cube of side 200, smoothed LJ cut
at rc = 2.5, Verlet margin 0.5, dt = 0.001, NVE, |v0| = sqrt(3) (T = 1),
observables (T, P, E) every 100 steps as benchmark.cc:32-40 does.  A step is
one MDManager::Calculate: expiry test, pair-list rebuild when expired, half
drift, halo refresh, force kick, half drift.

Reported: particle-updates/s of the FP64 path (`value`) and of the mixed FP32
path, the energy drift of both over --drift-steps, the pair-force kernel's
roofline position (HIP events on the domain's stream) and the oracle timed on
the host as the CPU baseline.

--workload ensemble: the batched-replica workload instead -- 256 independent
Langevin replicas (N = 4 k^3 drawn log-uniform in [2048, 65536], density in
[0.3, 0.9], T in [0.7, 1.5], gamma = 1, own key each), rc = 2.5, margin 0.5,
dt = 0.005, all advanced by one pair-force launch per step; particle-updates/s
summed over the replicas, FP64 and mixed.

--workload units: the default workload on the units of a process grid (--grid, default
2,2,2: MDManager's decomposition for OMP_NUM_THREADS=8), all on this GPU, with
device-to-device migration and halo between them.

--workload quench: 16,384,000 particles (FCC, 4*160^3) at density 0.75,
Langevin at T = 0.9 (gamma = 1, dt = 0.005) for --equil steps, then
MDManager::ChangeScale to density 0.60 and NVE at dt = 0.001 from that state,
FP64 and mixed; particle-updates/s of the NVE steps and the energy drift.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s, 4M LJ ρ=0.5: FP64 vs mixed FP32 at equal energy drift"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP64 force kernel on this workload, from the
# FETCH_SIZE and WRITE_SIZE counters (rocprofv3 --pmc) scaled by a copy of known
# size; written from the profile summaries.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_brick_fp64_4M_traffic.json")
ENS_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_brick_fp64_ensemble256_traffic.json")
QUENCH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_force_brick_fp64_quench16M_traffic.json")
FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3


def algorithmic_per_update(rho, rc, margin):
    """Half-list model: bytes and flops per particle-update."""
    n_L = 2.0 * math.pi / 3.0 * rho * (rc + margin) ** 3
    n_C = 2.0 * math.pi / 3.0 * rho * rc ** 3
    return 96.0 + 4.0 + 4.0 * n_L, 26.0 * n_C + 8.0 * (n_L - n_C) + 15.0, n_L, n_C


def thermalize(d, steps):
    """Steps taken before tuning and timing, as the reference's Benchmark project does
    (ThermalizeLoop, benchmark.cc:20-27), so that the timed window sees the melted fluid
    and not the starting lattice.  Returns the constructions per step over the second
    half of these steps: the Verlet cadence the timed window is amortised at."""
    first = steps - steps // 2
    if first:
        d.calculate(first)
    rebuilds = d.calculate(steps // 2) if steps // 2 else 0
    return rebuilds / max(1, steps // 2), rebuilds, steps // 2


def construction_ms(d, reps=3, wait=None):
    """Wall time of one pair-list construction (MDManager::MakePairList on the device, host
    round trips included): the median of `reps` constructions, each followed by a
    reduction that waits for the stream (`wait`, default the kinetic energy), less that
    reduction alone."""
    wait = wait or d.kinetic_energy
    wait()
    ts, ks = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        d.make_pair_list()
        wait()
        t1 = time.perf_counter()
        wait()
        t2 = time.perf_counter()
        ts.append(t1 - t0)
        ks.append(t2 - t1)
    return 1e3 * (float(np.median(ts)) - float(np.median(ks)))


def observation_ms(d, reps=3):
    """Wall time of one observation (T, P, E: the three observers and their reductions)."""
    d.observables()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        d.observables()
        ts.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(ts))


def amortised(n, steps, elapsed_s, rebuilds, n_obs, rate, t_build_ms, t_obs_ms, observe_every=100):
    """Particle-updates/s of the timed window with constructions and observations at the
    workload's own cadence: the window's own constructions and observations are replaced
    by steps * rate constructions and steps / observe_every observations (MUPS,
    benchmark.cc:45-51, over a run long enough to hold them).  A long window holds about
    that many already, and the correction vanishes."""
    t = elapsed_s + 1e-3 * ((steps * rate - rebuilds) * t_build_ms + (steps / observe_every - n_obs) * t_obs_ms)
    return steps * n / t, 1e3 * t / steps


def run_gpu(precision, x, p, L, args, observe_every=100):
    from mdacp_amd import mdgpu

    n = len(x)
    d = mdgpu.Domain(L, dt=args.dt, rc=args.rc, margin=args.margin, precision=precision, capacity=n)
    d.add_particles(x, p)
    d.make_pair_list()
    rate, therm_rebuilds, therm_steps = thermalize(d, args.thermalize)
    d.calculate(args.warmup)
    tuned = tune(d, args)
    d.calculate(20)
    d.observables()
    # the timed steps: fused steps queued ahead of the host while the validity budgets
    # allow (the library's defaults), observed after the first step
    # and every observe_every steps (benchmark.cc:32-40)
    rebuilds = 0
    n_obs = 0
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = 1 if done == 0 else min(observe_every, args.steps - done)
        rebuilds += d.calculate(k)
        done += k
        if (done - 1) % observe_every == 0:
            d.observables()
            n_obs += 1
    t1 = time.perf_counter()
    elapsed = t1 - t0
    t_build = construction_ms(d)
    t_obs = observation_ms(d)
    pu_s, ms_step = amortised(n, args.steps, elapsed, rebuilds, n_obs, rate, t_build, t_obs, observe_every)
    # host buffers at the boundary: the state read back into and written from host
    # arrays (mdgpu_get_particles, mdgpu_set_positions / mdgpu_set_momenta), as a
    # caller holding the state on the host pays once per timed window
    t2 = time.perf_counter()
    qh, ph, _ = d.particles()
    t3 = time.perf_counter()
    d.set_positions(qh)
    d.set_momenta(ph)
    t4 = time.perf_counter()
    pcie = dict(download_ms=1e3 * (t3 - t2), upload_ms=1e3 * (t4 - t3), bytes_each_way=int(qh.nbytes + ph.nbytes),
                pu_s=args.steps * n / (args.steps * ms_step * 1e-3 + t4 - t2))
    del qh, ph
    # the same number of steps again with HIP events around each pair-force launch and
    # each construction (launched directly, so slower per step than the timed pass; each
    # kick leaves the final positions to a kernel of its own, as in the timed pass's runs,
    # so the events time the kernel the timed steps launch)
    d.enable_timing(True)
    d.calculate(max(args.steps, 100))
    tm = d.timing()
    d.enable_timing(False)
    res = dict(pu_s=pu_s, ms_per_step=ms_step, window_pu_s=args.steps * n / elapsed,
               window_ms_per_step=1e3 * elapsed / args.steps,
               force_avg_ms=tm["force_ms"] / max(1, tm["force_launches"]), force_launches=tm["force_launches"],
               rebuilds=rebuilds, observations=n_obs, construction_ms=t_build, observation_ms=t_obs,
               rebuilds_per_step=rate, cadence_sample=f"{therm_rebuilds} constructions in the last {therm_steps} "
                                                      f"of {args.thermalize} thermalisation steps",
               list=d.list_stats(), tuned=tuned, pcie=pcie)
    return d, res


def tune(d, args):
    """Launch shapes from the autotuner (timed on a clone, outside the timed steps)."""
    if args.no_autotune:
        return None
    from mdacp_amd.autotune import autotune

    t0 = time.perf_counter()
    r = autotune(d)
    r["choice"]["seconds"] = time.perf_counter() - t0
    print(f"# autotune {r['choice']}", file=sys.stderr, flush=True)
    return r["choice"]


def energy_drift(d, steps, every=10, blocks=20):
    """Energy drift of E/N over `steps` steps: E sampled every `every` steps, averaged
    over `blocks` consecutive blocks; the slope of a least-squares line through the block
    means, scaled to 10^4 steps, with its standard error from the scatter of the block
    means about the line (block averages absorb the correlation of nearby samples).  The
    detection floor is twice the standard error: a slope below it is not distinguishable
    from zero on this run.  rms_fluct is the RMS of the samples about the line."""
    e, t = [], []
    for k in range(0, steps + 1, every):
        if k:
            d.calculate(every)
        e.append(d.observables()["E"])
        t.append(k)
    e, t = np.array(e), np.array(t, dtype=float)
    nb = len(e) // blocks
    eb = e[: nb * blocks].reshape(blocks, nb).mean(axis=1)
    tb = t[: nb * blocks].reshape(blocks, nb).mean(axis=1)
    A = np.stack([tb - tb.mean(), np.ones_like(tb)], 1)
    (a, b), res, _, _ = np.linalg.lstsq(A, eb, rcond=None)
    resid = eb - A @ np.array([a, b])
    se = math.sqrt(float(resid @ resid) / (blocks - 2) / float(np.sum((tb - tb.mean()) ** 2)))
    a_all, b_all = np.polyfit(t, e, 1)
    fl = e - (a_all * t + b_all)
    return dict(slope_per_1e4_steps=float(a * 1e4), stderr_per_1e4_steps=float(se * 1e4),
                floor_per_1e4_steps=float(2.0 * se * 1e4), rms_fluct=float(np.sqrt(np.mean(fl ** 2))),
                e_start=float(e[0]), e_end=float(e[-1]), steps=steps, sample_every=every, blocks=blocks,
                estimator="least-squares slope of E/N block means; stderr from block residuals")


def drift_comparison(d64, d32):
    """The mixed-precision path's drift target, |slope(mixed)| <= 2 |slope(FP64)|, judged with
    the FP64 slope replaced by its detection floor when it is below that floor."""
    if not d64 or not d32:
        return None
    ref = max(abs(d64["slope_per_1e4_steps"]), d64["floor_per_1e4_steps"])
    return dict(fp64_reference_per_1e4_steps=ref,
                fp64_below_floor=abs(d64["slope_per_1e4_steps"]) < d64["floor_per_1e4_steps"],
                mixed_over_fp64=abs(d32["slope_per_1e4_steps"]) / ref,
                mixed_within_2x=abs(d32["slope_per_1e4_steps"]) <= 2.0 * ref)


def reference_cpu():
    """The reference's own CPU program timed in the build container (oracle/time_reference.py:
    its Benchmark project, unmodified, rc 3.0 / margin 0.3 as compiled), for the record next
    to the port timed here."""
    path = os.path.join(ROOT, "profiles", "reference_cpu_benchmark.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        r = json.load(f)
    return [dict(value=run["particle_updates_per_s"], unit="particle-updates/s", cores=run["threads"],
                 kind="reference", host=r["host"], particles=run["particles"],
                 sample=(f"kaityo256/mdacp Benchmark project compiled unmodified, {run['threads']} OpenMP "
                         f"thread(s), N={run['particles']} vacancy FCC rho 0.5, rc {r['rc']} margin {r['margin']} "
                         f"(compiled in), {run['total_loop']} timed steps, MUPS line of benchmark.cc:45-51; "
                         f"timed in the build container, not on this machine; the fastest of "
                         f"{len(run.get('samples_mups', [run['mups']]))} runs on shared virtual CPUs "
                         f"(MUPS {run.get('samples_mups', [run['mups']])})"))
            for run in r["runs"]]


def cpu_baseline(x, p, L, args, gpu_rebuilds_per_step):
    from oracle import oracle as O

    s = O.OracleSystem(L, args.dt, args.rc, args.margin)
    s.add(x, p)
    t0 = time.perf_counter()
    s.make_pair_list()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.cpu_steps):
        s.step()
    t_steps = (time.perf_counter() - t0) / args.cpu_steps
    per_step = t_steps + t_build * gpu_rebuilds_per_step
    n = len(x)
    return dict(value=n / per_step, unit="particle-updates/s", cores=1, kind="port",
                sample=(f"oracle (C restatement of the reference step, 1 thread) on the same "
                        f"{n}-particle workload: {args.cpu_steps} steps timed ({t_steps:.2f} s/step) "
                        f"plus one pair-list build ({t_build:.2f} s) amortised at the GPU run's "
                        f"rebuild rate ({gpu_rebuilds_per_step:.4f}/step)"),
                reference=reference_cpu())


def run_ensemble_gpu(precision, table, built, args):
    from mdacp_amd import mdgpu

    boxes, counts, q, p, ids = built
    e = mdgpu.Ensemble(boxes, counts, q, p, ids, T=[r["T"] for r in table], gamma=1.0,
                       seed=[r["seed"] for r in table], thermostat=mdgpu.LANGEVIN, dt=args.dt, rc=args.rc,
                       margin=args.margin, precision=precision)
    rate, therm_rebuilds, therm_steps = thermalize(e, args.thermalize)
    e.calculate(args.warmup)
    tuned = tune(e, args)
    e.calculate(20)
    e.observables()
    n = int(counts.sum())
    rebuilds = 0
    n_obs = 0
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = 1 if done == 0 else min(100, args.steps - done)
        rebuilds += e.calculate(k)
        done += k
        if (done - 1) % 100 == 0:
            e.observables()
            n_obs += 1
    elapsed = time.perf_counter() - t0
    t_build = construction_ms(e, wait=e.observe_replicas_raw)
    t_obs = observation_ms(e)
    pu_s, ms_step = amortised(n, args.steps, elapsed, rebuilds, n_obs, rate, t_build, t_obs)
    # the same number of steps again with HIP events around each pair-force launch
    e.enable_timing(True)
    e.calculate(max(args.steps, 100))
    tm = e.timing()
    e.enable_timing(False)
    res = dict(pu_s=pu_s, ms_per_step=ms_step, window_pu_s=args.steps * n / elapsed, rebuilds=rebuilds,
               rebuilds_per_step=rate, construction_ms=t_build, observation_ms=t_obs,
               cadence_sample=f"{therm_rebuilds} constructions in the last {therm_steps} of {args.thermalize} "
                              f"thermalisation steps",
               force_avg_ms=tm["force_ms"] / max(1, tm["force_launches"]),
               rebuild_avg_ms=t_build, list=e.list_stats(),
               mean_T_over_target=float(np.mean(e.observables()["T"] / np.array([r["T"] for r in table]))),
               tuned=tuned)
    e.close()
    return res


def cpu_baseline_ensemble(table, built, args, sample=16, steps=3):
    """The oracle on the first `sample` replicas, one after another, `steps` Langevin steps each plus one
    pair-list build amortised at the GPU run's rebuild rate."""
    from oracle import oracle as O

    boxes, counts, q, p, ids = built
    offs = np.concatenate([[0], np.cumsum(counts)])
    t_steps = t_build = 0.0
    n = 0
    for r in range(sample):
        s = O.OracleSystem(float(boxes[r, 0]), args.dt, args.rc, args.margin)
        a, b = offs[r], offs[r + 1]
        s.add(q[a:b], p[a:b], ids[a:b])
        s.set_thermostat(O.LANGEVIN, table[r]["T"], 1.0, 0.1)
        s.set_noise(O.NOISE_PHILOX, table[r]["seed"])
        t0 = time.perf_counter()
        s.make_pair_list()
        t_build += time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step()
        t_steps += time.perf_counter() - t0
        n += int(b - a)
    return n, t_steps / steps, t_build


def measured_traffic(path, **match):
    """Measured L2-to-memory bytes per force launch from a committed profile, if it was taken
    on this workload (its recorded particle and replica counts match), else None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t["hbm_bytes_per_launch"] if all(t.get(k) == v for k, v in match.items()) else None


def ensemble_traffic(n, nrep):
    return measured_traffic(ENS_TRAFFIC_FILE, particles=n, replicas=nrep)


def main_ensemble(args):
    from mdacp_amd import mdgpu
    from mdacp_amd.ensemble import build, replica_table

    table = replica_table(nrep=args.replicas)
    built = build(table)
    n = int(built[1].sum())
    print(f"# ensemble: {len(table)} replicas, {n} particles", file=sys.stderr, flush=True)
    r64 = run_ensemble_gpu(mdgpu.FP64, table, built, args)
    print(f"# fp64 {r64}", file=sys.stderr, flush=True)
    r32 = run_ensemble_gpu(mdgpu.MIXED, table, built, args)
    print(f"# mixed {r32}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        ns, per_step, t_build = cpu_baseline_ensemble(table, built, args)
        per = per_step + t_build * r64["rebuilds_per_step"]
        cpu = dict(value=ns / per, unit="particle-updates/s", cores=1, kind="port",
                   sample=(f"oracle (C restatement of the reference step, 1 thread), the first 16 replicas "
                           f"({ns} particles) one after another: 3 Langevin steps each ({per_step:.3f} s per "
                           f"step of the 16) plus one pair-list build of each ({t_build:.3f} s) amortised at the "
                           f"GPU run's rebuild rate"))
    rho_mean = float(np.mean([r["rho"] for r in table]))
    bytes_pu, _, _, _ = algorithmic_per_update(rho_mean, args.rc, args.margin)
    avg_s = r64["force_avg_ms"] * 1e-3
    achieved = bytes_pu * n / avg_s / 1e9
    line = {
        "metric": "particle-updates/s, batched Langevin replicas (256, N 2048-62500, rho 0.3-0.9)",
        "value": r64["pu_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: per-replica FCC lattices (confmaker.cc:8-44), |v0|=sqrt(3T) (variables.cc:82-108)",
        "config": {"workload": (f"{len(table)} Langevin replicas, {n} particles, rc={args.rc} margin={args.margin} "
                                f"dt={args.dt} gamma=1, one pair-force launch per step"),
                   "particles": n, "replicas": len(table), "thermalize": args.thermalize,
                   "rebuilds_in_timed_steps": r64["rebuilds"]},
        "amortisation": {"rebuilds_per_step": r64["rebuilds_per_step"], "cadence_sample": r64["cadence_sample"],
                         "construction_ms": r64["construction_ms"], "observation_ms": r64["observation_ms"],
                         "window_pu_s": r64["window_pu_s"]},
        "roofline": dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=achieved / HBM_PEAK_GBS, traffic=ensemble_traffic(n, len(table)), kernel="k_force_brick_fp64",
                         kernel_avg_ms=r64["force_avg_ms"], bytes_per_particle_update=bytes_pu,
                         note="algorithmic bytes at the replicas' mean density; kernel time from HIP events over "
                              "a second pass of the same number of steps"),
        "cpu_baseline": cpu,
        "mixed": {"value": r32["pu_s"], "ms_per_step": r32["ms_per_step"],
                  "speedup_vs_fp64": r32["pu_s"] / r64["pu_s"], "force_kernel_avg_ms": r32["force_avg_ms"],
                  "rebuilds_in_timed_steps": r32["rebuilds"]},
        "rebuild_avg_ms": r64["rebuild_avg_ms"],
        "autotune": {"fp64": r64["tuned"], "mixed": r32["tuned"]},
        "mean_T_over_target": r64["mean_T_over_target"],
        "pair_list": r64["list"],
    }
    print(json.dumps(line), flush=True)


def main_quench(args):
    from mdacp_amd import mdgpu
    from mdacp_amd.lattice import fcc, initial_velocity

    from mdacp_amd.ensemble import box_for

    k = args.fcc_cells
    L = box_for(k, 0.75)
    x = fcc(L, 0.75)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0 * 0.9), 0)
    print(f"# quench: N={n} L={L} rho=0.75", file=sys.stderr, flush=True)
    g = mdgpu.Domain(L, dt=0.005, rc=args.rc, margin=args.margin, capacity=n)
    g.add_particles(x, p)
    g.set_thermostat(mdgpu.LANGEVIN, 0.9, 1.0, 0.1)
    g.set_seed(20261202)
    g.make_pair_list()
    t0 = time.perf_counter()
    g.calculate(args.equil)
    t_eq = time.perf_counter() - t0
    alpha = (0.75 / 0.6) ** (1.0 / 3.0)
    g.change_scale(alpha)
    print(f"# equilibrated {args.equil} Langevin steps in {t_eq:.1f} s; scaled by {alpha}", file=sys.stderr,
          flush=True)
    out = {}
    for name, prec in (("fp64", mdgpu.FP64), ("mixed", mdgpu.MIXED)):
        d = g.clone(dt=args.dt, precision=prec)
        d.set_thermostat(mdgpu.NVE)
        d.calculate(args.warmup)
        tuned = tune(d, args)
        d.calculate(20)
        d.observables()
        rebuilds = 0
        t0 = time.perf_counter()
        done = 0
        while done < args.steps:
            kk = min(100, args.steps - done)
            rebuilds += d.calculate(kk)
            done += kk
            d.observables()
        elapsed = time.perf_counter() - t0
        d.enable_timing(True)
        d.calculate(200)
        tm = d.timing()
        d.enable_timing(False)
        drift = energy_drift(d, args.drift_steps) if args.drift_steps else None
        out[name] = dict(pu_s=args.steps * n / elapsed, ms_per_step=1e3 * elapsed / args.steps, rebuilds=rebuilds,
                         force_avg_ms=tm["force_ms"] / max(1, tm["force_launches"]),
                         rebuild_avg_ms=tm["rebuild_ms"] / max(1, tm["rebuilds"]), drift=drift, tuned=tuned)
        print(f"# {name} {out[name]}", file=sys.stderr, flush=True)
        d.close()
    g.close()
    bytes_pu, _, _, _ = algorithmic_per_update(0.6, args.rc, args.margin)
    avg_s = out["fp64"]["force_avg_ms"] * 1e-3
    achieved = bytes_pu * n / avg_s / 1e9
    line = {
        "metric": "particle-updates/s, quench 16.4M LJ rho 0.75 -> 0.60, NVE after ChangeScale",
        "value": out["fp64"]["pu_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["fp64"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: FCC lattice (confmaker.cc:8-44) at rho 0.75, Langevin T=0.9, ChangeScale to rho 0.60",
        "config": {"workload": (f"N={n} rho 0.75 -> 0.60 (alpha={alpha:.7f}), {args.equil} Langevin steps "
                                f"(T=0.9, gamma=1, dt=0.005), then NVE dt={args.dt} rc={args.rc} "
                                f"margin={args.margin}"), "particles": n,
                   "rebuilds_in_timed_steps": out["fp64"]["rebuilds"]},
        "roofline": dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                         traffic=measured_traffic(QUENCH_TRAFFIC_FILE, particles=n), kernel="k_force_brick_fp64",
                         kernel_avg_ms=out["fp64"]["force_avg_ms"],
                         bytes_per_particle_update=bytes_pu),
        "cpu_baseline": None,
        "mixed": {"value": out["mixed"]["pu_s"], "ms_per_step": out["mixed"]["ms_per_step"],
                  "speedup_vs_fp64": out["mixed"]["pu_s"] / out["fp64"]["pu_s"],
                  "force_kernel_avg_ms": out["mixed"]["force_avg_ms"], "rebuilds_in_timed_steps": out["mixed"]["rebuilds"]},
        "energy_drift": {"fp64": out["fp64"]["drift"], "mixed": out["mixed"]["drift"]},
        "rebuild_avg_ms": out["fp64"]["rebuild_avg_ms"],
        "autotune": {"fp64": out["fp64"]["tuned"], "mixed": out["mixed"]["tuned"]},
    }
    print(json.dumps(line), flush=True)


def run_units_gpu(precision, x, p, L, grid, args, observe_every=100):
    """The default workload on the units of a process grid (mdgpu_create_unit): MDManager's
    decomposition with OMP_NUM_THREADS = gx*gy*gz threads, every unit on this GPU, stepped
    by mdgpu_group_calculate; particle-updates/s amortised as for the single domain."""
    from mdacp_amd import mdgpu

    n = len(x)
    g = mdgpu.UnitGroup(L, grid, dt=args.dt, rc=args.rc, margin=args.margin, precision=precision, capacity=n)
    g.add_particles(x, p)
    g.make_pair_list()
    rate, therm_rebuilds, therm_steps = thermalize(g, args.thermalize)
    g.calculate(args.warmup)
    g.calculate(20)
    g.observables()
    rebuilds = 0
    n_obs = 0
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = 1 if done == 0 else min(observe_every, args.steps - done)
        rebuilds += g.calculate(k)
        done += k
        if (done - 1) % observe_every == 0:
            g.observables()
            n_obs += 1
    elapsed = time.perf_counter() - t0
    t_build = construction_ms(g)
    t_obs = observation_ms(g)
    pu_s, ms_step = amortised(n, args.steps, elapsed, rebuilds, n_obs, rate, t_build, t_obs)
    g.enable_timing(True)
    g.calculate(args.steps)
    tm = g.timing()
    g.enable_timing(False)
    res = dict(pu_s=pu_s, ms_per_step=ms_step, window_pu_s=args.steps * n / elapsed, rebuilds=rebuilds,
               rebuilds_per_step=rate, construction_ms=t_build, observation_ms=t_obs,
               cadence_sample=f"{therm_rebuilds} constructions in the last {therm_steps} of {args.thermalize} "
                              f"thermalisation steps",
               force_ms_per_step=tm["force_ms"] / args.steps, force_launches_per_step=tm["force_launches"] / args.steps,
               step_wall_ms_timed_pass=tm["total_ms"] / args.steps, list=g.list_stats(),
               units=[u.counts()[0] for u in g.units])
    g.close()
    return res


def main_units(args):
    from mdacp_amd import mdgpu
    from mdacp_amd.ensemble import box_for
    from mdacp_amd.lattice import fcc, initial_velocity

    grid = tuple(int(v) for v in args.grid.split(","))
    k = args.fcc_cells or 100
    L = box_for(k, args.density)
    x = fcc(L, args.density)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0), 0)
    print(f"# units {grid}: N={n} L={L} rho={args.density}", file=sys.stderr, flush=True)
    r64 = run_units_gpu(mdgpu.FP64, x, p, L, grid, args)
    print(f"# fp64 {r64}", file=sys.stderr, flush=True)
    r32 = run_units_gpu(mdgpu.MIXED, x, p, L, grid, args)
    print(f"# mixed {r32}", file=sys.stderr, flush=True)
    bytes_pu, _, _, _ = algorithmic_per_update(args.density, args.rc, args.margin)
    achieved = bytes_pu * n / (r64["force_ms_per_step"] * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=None, kernel="k_force_group_fp64 (the bricks of all units in one launch per step; per-unit launches summed when a group cannot batch)",
                    kernel_avg_ms=r64["force_ms_per_step"], bytes_per_particle_update=bytes_pu)
    ref = reference_cpu()
    ref8 = [r for r in (ref or []) if r["cores"] == 8]
    cpu = None
    if ref8:
        cpu = dict(ref8[0], sample="the reference's own program with the same decomposition: " + ref8[0]["sample"])
    line = {
        "metric": METRIC,
        "value": r64["pu_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: FCC lattice (confmaker.cc:8-44), |v0|=sqrt(3) random directions (variables.cc:82-108)",
        "config": {"workload": (f"LJ NVE N={n} FCC rho={args.density} L={L:.6g} rc={args.rc} margin={args.margin} "
                                f"dt={args.dt} on {grid[0]}x{grid[1]}x{grid[2]} units of a process grid on one GPU "
                                f"(MDManager with OMP_NUM_THREADS={grid[0] * grid[1] * grid[2]}: device-to-device "
                                f"migration and halo); {args.thermalize} thermalisation steps"),
                   "particles": n, "grid": list(grid), "units": r64["units"],
                   "rebuilds_in_timed_steps": r64["rebuilds"]},
        "amortisation": {"rebuilds_per_step": r64["rebuilds_per_step"], "cadence_sample": r64["cadence_sample"],
                         "construction_ms": r64["construction_ms"], "observation_ms": r64["observation_ms"],
                         "window_pu_s": r64["window_pu_s"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "force_launches_per_step": r64["force_launches_per_step"],
        "mixed": {"value": r32["pu_s"], "ms_per_step": r32["ms_per_step"], "speedup_vs_fp64": r32["pu_s"] / r64["pu_s"],
                  "force_ms_per_step": r32["force_ms_per_step"]},
        "pair_list": r64["list"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["single", "ensemble", "quench", "units"], default="single")
    ap.add_argument("--grid", default="2,2,2", help="units: units per axis (the reference's OMP_NUM_THREADS=8 grid)")
    ap.add_argument("--equil", type=int, default=200, help="quench: Langevin steps before the rescale")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--thermalize", type=int, default=None,
                    help="steps before tuning and timing (default 3000 single, 400 ensemble, 0 quench)")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes")
    ap.add_argument("--fcc-cells", type=int, default=None, help="N = 4 k^3 (100; quench 160)")
    ap.add_argument("--density", type=float, default=0.5)
    ap.add_argument("--dt", type=float, default=None, help="0.001 (single), 0.005 (ensemble)")
    ap.add_argument("--rc", type=float, default=2.5)
    ap.add_argument("--margin", type=float, default=0.5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: one domain runs on one GPU; --gpus {args.gpus} is not supported "
              f"(several GPUs would each own one box of a spatial domain decomposition, exchanging "
              f"ghost images every step, which is not built)", file=sys.stderr)
        sys.exit(2)
    if args.thermalize is None:
        args.thermalize = {"single": 3000, "ensemble": 400, "quench": 0, "units": 3000}[args.workload]
    if args.workload == "ensemble":
        if args.dt is None:
            args.dt = 0.005
        if args.steps == 2000:
            args.steps = 1000
        main_ensemble(args)
        return
    if args.dt is None:
        args.dt = 0.001
    if args.workload == "units":
        main_units(args)
        return
    if args.workload == "quench":
        if args.fcc_cells is None:
            args.fcc_cells = 160
        if args.steps == 2000:
            args.steps = 500
        if args.drift_steps == 10000:
            args.drift_steps = 2000
        main_quench(args)
        return
    if args.fcc_cells is None:
        args.fcc_cells = 100

    from mdacp_amd import mdgpu
    from mdacp_amd.lattice import fcc, initial_velocity

    from mdacp_amd.ensemble import box_for

    k = args.fcc_cells
    L = box_for(k, args.density)
    x = fcc(L, args.density)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0), 0)
    print(f"# workload N={n} L={L} rho={args.density}", file=sys.stderr, flush=True)

    d64, r64 = run_gpu(mdgpu.FP64, x, p, L, args)
    print(f"# fp64 {r64}", file=sys.stderr, flush=True)
    drift64 = energy_drift(d64, args.drift_steps) if args.drift_steps else None
    print(f"# fp64 drift {drift64}", file=sys.stderr, flush=True)
    # force error of the mixed kick on the FP64 trajectory's state
    k64 = d64.kick()
    d64.set_precision(mdgpu.MIXED)
    k32 = d64.kick()
    force_rms = float(math.sqrt(np.sum((k32 - k64) ** 2) / np.sum(k64 ** 2)))
    d64.close()
    del k64, k32

    d32, r32 = run_gpu(mdgpu.MIXED, x, p, L, args)
    print(f"# mixed {r32}", file=sys.stderr, flush=True)
    drift32 = energy_drift(d32, args.drift_steps) if args.drift_steps else None
    print(f"# mixed drift {drift32}", file=sys.stderr, flush=True)
    d32.close()

    bytes_pu, flops_pu, n_L, n_C = algorithmic_per_update(args.density, args.rc, args.margin)
    avg_s = r64["force_avg_ms"] * 1e-3
    achieved = bytes_pu * n / avg_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and n == 4_000_000:
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=traffic, kernel="k_force_brick_fp64",
                    kernel_avg_ms=r64["force_avg_ms"], bytes_per_particle_update=bytes_pu,
                    vector_flop_frac=flops_pu * n / avg_s / 1e12 / FP64_VECTOR_PEAK_TFLOPS)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(x, p, L, args, r64["rebuilds_per_step"])

    avg32 = r32["force_avg_ms"] * 1e-3
    line = {
        "metric": METRIC,
        "value": r64["pu_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: FCC lattice (confmaker.cc:8-44), |v0|=sqrt(3) random directions (variables.cc:82-108)",
        "config": {"workload": (f"LJ NVE N={n} FCC rho={args.density} L={L:.6g} rc={args.rc} margin={args.margin} "
                                f"dt={args.dt}, T/P/E observed every 100 steps, rebuilds by the Verlet budget; "
                                f"{args.thermalize} thermalisation steps before tuning and timing"),
                   "particles": n, "thermalize": args.thermalize, "rebuilds_in_timed_steps": r64["rebuilds"]},
        "amortisation": {"rebuilds_per_step": r64["rebuilds_per_step"], "cadence_sample": r64["cadence_sample"],
                         "construction_ms": r64["construction_ms"], "observation_ms": r64["observation_ms"],
                         "observations_in_timed_steps": r64["observations"],
                         "window_pu_s": r64["window_pu_s"], "window_ms_per_step": r64["window_ms_per_step"],
                         "note": "value = particle-updates/s of the timed steps with constructions at the measured "
                                 "Verlet cadence and observations every 100 steps (the window's own replaced by "
                                 "that many); window_pu_s is the timed steps as they fell"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": {"value": r32["pu_s"], "ms_per_step": r32["ms_per_step"], "speedup_vs_fp64": r32["pu_s"] / r64["pu_s"],
                  "force_kernel_avg_ms": r32["force_avg_ms"],
                  "force_kernel_speedup_vs_fp64": r64["force_avg_ms"] / r32["force_avg_ms"],
                  "force_kernel_frac_hbm": bytes_pu * n / avg32 / 1e9 / HBM_PEAK_GBS,
                  "force_rms_rel_err_vs_fp64": force_rms, "rebuilds_in_timed_steps": r32["rebuilds"],
                  "construction_ms": r32["construction_ms"], "rebuilds_per_step": r32["rebuilds_per_step"],
                  "window_pu_s": r32["window_pu_s"]},
        "energy_drift": {"fp64": drift64, "mixed": drift32, "criterion": drift_comparison(drift64, drift32)},
        "pair_list": r64["list"],
        "rebuild_avg_ms": r64["construction_ms"],
        "pcie_inclusive": r64["pcie"],
        "autotune": {"fp64": r64["tuned"], "mixed": r32["tuned"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
