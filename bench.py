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
dt = 0.005, all advanced by one pair-force launch per step replayed from a
captured graph; particle-updates/s summed over the replicas, FP64 and mixed.

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


def run_gpu(precision, x, p, L, args, observe_every=100):
    from mdacp_amd import mdgpu

    n = len(x)
    d = mdgpu.Domain(L, dt=args.dt, rc=args.rc, margin=args.margin, precision=precision, capacity=n)
    d.add_particles(x, p)
    d.make_pair_list()
    d.calculate(args.warmup)
    tuned = tune(d, args)
    d.calculate(20)
    d.observables()
    # the timed steps: graph-replayed fused steps, queued ahead of the host while the
    # validity budgets allow (the library's defaults), observed every 100 steps
    rebuilds = 0
    trace = []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(observe_every, args.steps - done)
        rebuilds += d.calculate(k)
        done += k
        trace.append(d.observables())
    t1 = time.perf_counter()
    elapsed = t1 - t0
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
                pu_s=args.steps * n / (elapsed + t4 - t2))
    del qh, ph
    # the same number of steps again with HIP events around each pair-force launch and
    # each construction (launched directly, so slower per step than the timed pass)
    d.enable_timing(True)
    d.calculate(args.steps)
    tm = d.timing()
    d.enable_timing(False)
    res = dict(pu_s=args.steps * n / elapsed, ms_per_step=1e3 * elapsed / args.steps,
               force_avg_ms=tm["force_ms"] / max(1, tm["force_launches"]), force_launches=tm["force_launches"],
               rebuilds=rebuilds, rebuild_avg_ms=tm["rebuild_ms"] / max(1, tm["rebuilds"]), list=d.list_stats(),
               tuned=tuned, pcie=pcie)
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


def energy_drift(d, steps, every=100):
    """Slope of E/N (linear fit) scaled to 10^4 steps, and the RMS residual."""
    e, t = [], []
    for k in range(0, steps + 1, every):
        if k:
            d.calculate(every)
        ob = d.observables()
        e.append(ob["E"])
        t.append(k)
    e, t = np.array(e), np.array(t, dtype=float)
    a, b = np.polyfit(t, e, 1)
    resid = e - (a * t + b)
    return dict(slope_per_1e4_steps=float(a * 1e4), rms_fluct=float(np.sqrt(np.mean(resid ** 2))),
                e_start=float(e[0]), e_end=float(e[-1]), steps=steps)


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
                        f"rebuild rate ({gpu_rebuilds_per_step:.4f}/step)"))


def run_ensemble_gpu(precision, table, built, args):
    from mdacp_amd import mdgpu

    boxes, counts, q, p, ids = built
    e = mdgpu.Ensemble(boxes, counts, q, p, ids, T=[r["T"] for r in table], gamma=1.0,
                       seed=[r["seed"] for r in table], thermostat=mdgpu.LANGEVIN, dt=args.dt, rc=args.rc,
                       margin=args.margin, precision=precision)
    e.calculate(args.warmup)
    tuned = tune(e, args)
    e.calculate(20)
    e.observables()
    n = int(counts.sum())
    rebuilds = 0
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(100, args.steps - done)
        rebuilds += e.calculate(k)
        done += k
        e.observables()
    elapsed = time.perf_counter() - t0
    # the same number of steps again with HIP events around each pair-force launch
    e.enable_timing(True)
    e.calculate(args.steps)
    tm = e.timing()
    e.enable_timing(False)
    res = dict(pu_s=args.steps * n / elapsed, ms_per_step=1e3 * elapsed / args.steps, rebuilds=rebuilds,
               force_avg_ms=tm["force_ms"] / max(1, tm["force_launches"]),
               rebuild_avg_ms=tm["rebuild_ms"] / max(1, tm["rebuilds"]), list=e.list_stats(),
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
        per = per_step + t_build * r64["rebuilds"] / args.steps
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
                                f"dt={args.dt} gamma=1, one pair-force launch per step (graph replay)"),
                   "particles": n, "replicas": len(table), "rebuilds_in_timed_steps": r64["rebuilds"]},
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["single", "ensemble", "quench"], default="single")
    ap.add_argument("--equil", type=int, default=200, help="quench: Langevin steps before the rescale")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
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
    if args.workload == "ensemble":
        if args.dt is None:
            args.dt = 0.005
        if args.steps == 2000:
            args.steps = 1000
        main_ensemble(args)
        return
    if args.dt is None:
        args.dt = 0.001
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
        cpu = cpu_baseline(x, p, L, args, r64["rebuilds"] / args.steps)

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
                                f"dt={args.dt}, T/P/E observed every 100 steps, rebuilds by the Verlet budget"),
                   "particles": n, "rebuilds_in_timed_steps": r64["rebuilds"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": {"value": r32["pu_s"], "ms_per_step": r32["ms_per_step"], "speedup_vs_fp64": r32["pu_s"] / r64["pu_s"],
                  "force_kernel_avg_ms": r32["force_avg_ms"],
                  "force_kernel_frac_hbm": bytes_pu * n / avg32 / 1e9 / HBM_PEAK_GBS,
                  "force_rms_rel_err_vs_fp64": force_rms, "rebuilds_in_timed_steps": r32["rebuilds"]},
        "energy_drift": {"fp64": drift64, "mixed": drift32},
        "pair_list": r64["list"],
        "rebuild_avg_ms": r64["rebuild_avg_ms"],
        "pcie_inclusive": r64["pcie"],
        "autotune": {"fp64": r64["tuned"], "mixed": r32["tuned"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
