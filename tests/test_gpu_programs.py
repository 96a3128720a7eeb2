"""The reference's own program (main.cc, MDManager, ProjectManager and its
projects, unmodified) running on the MI355X facade (mdacp_amd/facade ->
libmdgpu.so), against the same program on the CPU.

Both binaries are built from the reference sources by `make -C oracle programs`
(oracle/_ref/mdacp_gpu and oracle/_ref/mdacp_cpu); the tests are skipped where
they were not built.  They read no reference file at run time: the configs are
written here.  The programs print observables with 6 significant digits
(std::ostream defaults), which bounds the comparison.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
GPU = os.path.join(REF, "mdacp_gpu")
CPU = os.path.join(REF, "mdacp_cpu")

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not (os.path.exists(GPU) and os.path.exists(CPU)),
                       reason="oracle/_ref programs not built (make -C oracle programs)"),
]


def run(binary, cfg_text, tmp_path, tag, env_extra=None, timeout=240):
    d = tmp_path / tag
    d.mkdir()
    cfg = d / "input.cfg"
    cfg.write_text(cfg_text)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    env.update(env_extra or {})
    print(f"running {os.path.basename(binary)} {tag}", flush=True)
    r = subprocess.run([binary, str(cfg)], cwd=d, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"{binary} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return r.stdout


def observe_rows(out, marker):
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


BENCH = """Mode=Benchmark
ThermalizeLoop=20
SystemSize=20
TimeStep=0.001
TotalLoop=400
ObserveLoop=100
InitialVelocity=1.0
"""


def test_benchmark_project_matches_reference_program(tmp_path):
    """Benchmark project (NVE, vacancy FCC, |v0| = 1): t, T, P, E every 100 steps."""
    cpu = observe_rows(run(CPU, BENCH, tmp_path, "cpu"), "#observe")
    gpu = observe_rows(run(GPU, BENCH, tmp_path, "gpu"), "#observe")
    assert cpu.shape == gpu.shape == (4, 4), (cpu, gpu)
    # 6 printed digits: agreement to the last printed digit (relative 1e-5 of each value's scale)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


def test_benchmark_project_mixed_precision(tmp_path):
    cpu = observe_rows(run(CPU, BENCH, tmp_path, "cpu"), "#observe")
    gpu = observe_rows(run(GPU, BENCH, tmp_path, "gpu", {"MDACP_GPU_PRECISION": "mixed"}), "#observe")
    assert cpu.shape == gpu.shape
    np.testing.assert_allclose(gpu[:, 0], cpu[:, 0])
    # FP32 pair forces: the trajectory separates slowly from the FP64 one
    np.testing.assert_allclose(gpu[:, 1:], cpu[:, 1:], rtol=2e-3, atol=2e-3)


NOSEHOOVER = """Mode=Benchmark
ThermalizeLoop=10
SystemSize=20
TimeStep=0.002
TotalLoop=300
ObserveLoop=100
InitialVelocity=1.0
ControlTemperature=yes
AimedTemperature=0.9
HeatbathType=NoseHoover
"""


def test_nose_hoover_benchmark_matches_reference_program(tmp_path):
    """Benchmark project with the Nose-Hoover heat bath (two temperature reductions a step)."""
    cpu = observe_rows(run(CPU, NOSEHOOVER, tmp_path, "cpu"), "#observe")
    gpu = observe_rows(run(GPU, NOSEHOOVER, tmp_path, "gpu"), "#observe")
    assert cpu.shape == gpu.shape == (3, 4)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


LANGEVIN = """Mode=Langevin
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
"""


def test_langevin_project_thermalises(tmp_path):
    """Langevin project: the noise stream differs from the reference's (Philox on the
    device, std::mt19937 on the host), so the runs agree in distribution only: same
    first observation, and both relax towards T = 0.9."""
    cpu = observe_rows(run(CPU, LANGEVIN, tmp_path, "cpu"), None)
    gpu = observe_rows(run(GPU, LANGEVIN, tmp_path, "gpu"), None)
    assert cpu.shape == gpu.shape == (30, 4)
    np.testing.assert_allclose(gpu[:, 0], cpu[:, 0])
    assert abs(gpu[0, 1] - cpu[0, 1]) < 0.02 * cpu[0, 1]
    t_gpu, t_cpu = gpu[15:, 1].mean(), cpu[15:, 1].mean()
    assert abs(t_gpu - 0.9) < 0.06 and abs(t_cpu - 0.9) < 0.06, (t_gpu, t_cpu)
