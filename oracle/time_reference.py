"""Times the reference's own CPU program (oracle/_ref/mdacp_cpu: kaityo256/mdacp
compiled unmodified from its sources with its default flags, makefile:14) on its
Benchmark project (benchmark.cc:12-52), and writes the MUPS it prints
(benchmark.cc:45-51: TotalLoop * N / seconds / 1e6) to
profiles/reference_cpu_benchmark.json.  TEST AND MEASUREMENT INFRASTRUCTURE: runs
in the build container, where the reference sources are; never on the product path.

The reference fixes rc = 3.0 (mdconfig.h:15) and margin 0.3 (simulationinfo.cc:6)
at compile time and holds at most 10^6 particles per unit (mdconfig.h:11); its
Benchmark starts from the vacancy-diluted FCC of ExtractConfigurationMaker
(confmaker.cc:173-252) at the given density.  With OMP_NUM_THREADS = 8 the
4,000,000-particle box (SystemSize 200, density 0.5) splits into 2x2x2 units of
500,000; one thread runs one unit of SystemSize 100 (500,000 particles).

    python oracle/time_reference.py [--loops 300]
"""
import argparse
import json
import os
import platform
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "_ref", "mdacp_cpu")
OUT = os.path.join(ROOT, "profiles", "reference_cpu_benchmark.json")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run(threads, size, loops, thermalize, density=0.5, v0=3.0 ** 0.5):
    cfg = (f"Mode=Benchmark\nThermalizeLoop={thermalize}\nSystemSize={size}\nDensity={density}\n"
           f"TimeStep=0.001\nTotalLoop={loops}\nObserveLoop=100\nInitialVelocity={v0!r}\n")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "bench.cfg")
        with open(path, "w") as f:
            f.write(cfg)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        out = subprocess.run([BIN, path], cwd=td, env=env, capture_output=True, text=True, check=True).stdout
    m = re.search(r"# N = (\d+) ([0-9.eE+-]+) \[SEC\] ([0-9.eE+-]+) \[MUPS\]", out)
    if not m:
        raise RuntimeError("no MUPS line in the reference's output:\n" + out[-2000:])
    n, sec, mups = int(m.group(1)), float(m.group(2)), float(m.group(3))
    observe = [line for line in out.splitlines() if line.endswith("#observe")]
    return dict(threads=threads, system_size=size, density=density, particles=n, total_loop=loops,
                thermalize_loop=thermalize, seconds=sec, mups=mups, particle_updates_per_s=mups * 1e6,
                observe=observe, config=cfg)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--loops", type=int, default=300)
    ap.add_argument("--thermalize", type=int, default=150)
    ap.add_argument("--repeat", type=int, default=2,
                    help="runs per configuration; the build container's virtual CPUs are shared, so the "
                         "fastest run is the one reported")
    args = ap.parse_args()
    if not os.path.exists(BIN):
        raise SystemExit("oracle/_ref/mdacp_cpu is not built: make -C oracle programs")
    runs = []
    for threads, size in ((1, 100), (8, 200)):
        samples = [run(threads, size, args.loops, args.thermalize) for _ in range(args.repeat)]
        best = max(samples, key=lambda r: r["mups"])
        best = dict(best, samples_mups=[r["mups"] for r in samples])
        runs.append(best)
    res = dict(program="kaityo256/mdacp Benchmark project, compiled unmodified (g++ -O3 -fopenmp, makefile:14), "
                       "CalculateForceNext kernel (default build)",
               rc=3.0, margin=0.3, dt=0.001, host=cpu_model(), host_cores=os.cpu_count(),
               note="virtual CPUs of a shared host: runs of one configuration vary; the fastest is reported",
               runs=runs)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    for r in runs:
        print(f"threads={r['threads']} N={r['particles']} {r['seconds']:.2f} s {r['mups']:.3f} MUPS "
              f"(samples {r['samples_mups']})", flush=True)


if __name__ == "__main__":
    main()
