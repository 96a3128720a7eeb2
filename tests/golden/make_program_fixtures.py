"""Fixtures from the reference's own program on the CPU (oracle/_ref/mdacp_cpu:
kaityo256/mdacp's main.cc, MDManager, ProjectManager and projects, compiled
unmodified from its sources by `make -C oracle programs`).  Runs in the build
container only; tests/test_gpu_programs.py runs the same program on the MI355X
facade (oracle/_ref/mdacp_gpu) and compares with these, so the reference itself
never runs on the GPU machine.

For each project configuration in PROGRAMS (OMP_NUM_THREADS = its `threads`, 1 by
default): the rows the program prints
(time, T[, P, E] per observation) and every cdview dump it writes
(MDManager::SaveAsCdviewSequential, mdmanager.cc:81-105; lines
"0 type x y z px py pz", mdunit.cc:52-67), as float arrays in
tests/golden/programs.npz, with the configurations in programs.json.

    python tests/golden/make_program_fixtures.py
"""
import glob
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CPU = os.path.join(ROOT, "oracle", "_ref", "mdacp_cpu")

PROGRAMS = {
    # Benchmark project (NVE, vacancy FCC, |v0| = 1), t T P E every 100 steps
    "benchmark": dict(marker="#observe", cfg="""Mode=Benchmark
ThermalizeLoop=20
SystemSize=20
TimeStep=0.001
TotalLoop=400
ObserveLoop=100
InitialVelocity=1.0
"""),
    # Benchmark with the Nose-Hoover heat bath (two temperature reductions a step)
    "nosehoover": dict(marker="#observe", cfg="""Mode=Benchmark
ThermalizeLoop=10
SystemSize=20
TimeStep=0.002
TotalLoop=300
ObserveLoop=100
InitialVelocity=1.0
ControlTemperature=yes
AimedTemperature=0.9
HeatbathType=NoseHoover
"""),
    # Langevin project (the GPU noise is another stream: compared in distribution)
    "langevin": dict(marker=None, cfg="""Mode=Langevin
InitialVelocity=1.8
TimeStep=0.005
UnitLength=16
ObserveLoop=100
TotalLoop=3000
AimedTemperature=0.9
Density=0.70
HeatbathTau=0.1
HeatbathGamma=1.0
HeatbathType=Langevin
Seed=12345
"""),
    # Collision project: two droplets flying at each other, cdview dumps
    "collision": dict(marker="# observe", cfg="""Mode=Collision
SystemSizeX=48
SystemSizeY=24
SystemSizeZ=24
TimeStep=0.002
ObserveLoop=100
TotalLoop=400
SaveCdviewFile=yes
"""),
    # Burst project: a droplet thermalised by Nose-Hoover, then NVE, cdview dumps
    "burst": dict(marker="# observe", cfg="""Mode=Burst
Density=0.8
ThermalizeLoop=200
AimedTemperature=1.5
SystemSize=30
Radius=8
TimeStep=0.005
TotalLoop=300
ObserveLoop=100
InitialVelocity=0.9
"""),
}

# The same programs with several units in one process (OMP_NUM_THREADS > 1: ParaInfo
# splits the box into one unit per thread; each unit draws its own initial velocities
# with its id as the seed, so these runs differ from the one-unit runs above)
for _name, _threads in (("benchmark", 8), ("nosehoover", 8), ("collision", 4), ("burst", 8)):
    PROGRAMS[f"{_name}_omp{_threads}"] = dict(PROGRAMS[_name], threads=_threads)


def observe_rows(out, marker):
    """Rows of numbers the program printed (the observation lines)."""
    rows = []
    for line in out.splitlines():
        if line.startswith("#") or not line.strip():
            continue
        if marker:
            if marker not in line:
                continue
            line = line.replace(marker, " ")
        rows.append([float(v) for v in line.split()[:4]])
    return np.array(rows)


def read_cd(path):
    """A cdview dump as an (n, 8) array."""
    return np.loadtxt(path, ndmin=2)


def run_program(binary, cfg_text, workdir, env_extra=None, timeout=600):
    cfg = os.path.join(workdir, "input.cfg")
    with open(cfg, "w") as f:
        f.write(cfg_text)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    env.update(env_extra or {})
    r = subprocess.run([binary, cfg], cwd=workdir, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{binary} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    dumps = {os.path.basename(p): read_cd(p) for p in sorted(glob.glob(os.path.join(workdir, "conf*.cd")))}
    return r.stdout, dumps


def main():
    if not os.path.exists(CPU):
        sys.exit("oracle/_ref/mdacp_cpu is not built: make -C oracle programs")
    arrays, meta = {}, {}
    for name, spec in PROGRAMS.items():
        threads = spec.get("threads", 1)
        with tempfile.TemporaryDirectory() as td:
            out, dumps = run_program(CPU, spec["cfg"], td, {"OMP_NUM_THREADS": str(threads)})
        rows = observe_rows(out, spec["marker"])
        arrays[f"{name}/rows"] = rows
        for fn, a in dumps.items():
            arrays[f"{name}/{fn}"] = a
        meta[name] = dict(cfg=spec["cfg"], marker=spec["marker"], rows=list(rows.shape), dumps=sorted(dumps),
                          threads=threads)
        print(name, rows.shape, {k: v.shape for k, v in dumps.items()}, flush=True)
    np.savez_compressed(os.path.join(HERE, "programs.npz"), **arrays)
    with open(os.path.join(HERE, "programs.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
