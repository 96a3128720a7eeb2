"""The reference's own program (main.cc, MDManager, ProjectManager and its
projects, unmodified) running on the MI355X facade (mdacp_amd/facade ->
libmdgpu.so), against what the same program printed and dumped on the CPU.

oracle/_ref/mdacp_gpu is built from the reference sources by
`make -C oracle programs`; the CPU program's output is a fixture written in the
build container by tests/golden/make_program_fixtures.py (tests/golden/
programs.npz), so the reference never runs here.  The programs print observables
with 6 significant digits (std::ostream defaults), which bounds the comparison;
cdview dumps (conf*.cd: "0 type x y z px py pz" per particle, mdunit.cc:52-67)
are matched particle by particle by position, since the facade stores particles
in its own order.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU = os.path.join(ROOT, "oracle", "_ref", "mdacp_gpu")
GOLDEN = os.path.join(ROOT, "tests", "golden")

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not os.path.exists(GPU), reason="oracle/_ref/mdacp_gpu not built (make -C oracle programs)"),
]


def fixture(name):
    with open(os.path.join(GOLDEN, "programs.json")) as f:
        meta = json.load(f)[name]
    arrs = np.load(os.path.join(GOLDEN, "programs.npz"))
    rows = arrs[f"{name}/rows"]
    dumps = {fn: arrs[f"{name}/{fn}"] for fn in meta["dumps"]}
    return meta, rows, dumps


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


def run_gpu(name, tmp_path, env_extra=None, timeout=240):
    meta, rows, dumps = fixture(name)
    d = tmp_path / name
    d.mkdir()
    cfg = d / "input.cfg"
    cfg.write_text(meta["cfg"])
    env = dict(os.environ, OMP_NUM_THREADS=str(meta.get("threads", 1)))
    env.update(env_extra or {})
    print(f"running mdacp_gpu {name} threads={env['OMP_NUM_THREADS']} {env_extra or ''}", flush=True)
    r = subprocess.run([GPU, str(cfg)], cwd=d, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, f"mdacp_gpu failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    got = {fn: np.loadtxt(d / fn, ndmin=2) for fn in sorted(os.listdir(d)) if fn.endswith(".cd")}
    return observe_rows(r.stdout, meta["marker"]), rows, got, dumps


def match_dump(got, want, pos_tol, mom_tol):
    """Pairs every particle of one dump with the nearest particle of the other by position
    (a bijection is required), then bounds the position and momentum differences."""
    from scipy.spatial import cKDTree

    assert got.shape == want.shape, (got.shape, want.shape)
    assert np.all(got[:, 0] == 0.0) and np.all(want[:, 0] == 0.0)
    dist, idx = cKDTree(want[:, 2:5]).query(got[:, 2:5])
    assert len(np.unique(idx)) == len(idx), "positions do not pair one to one"
    w = want[idx]
    assert np.array_equal(got[:, 1], w[:, 1])  # particle types
    dq = np.abs(got[:, 2:5] - w[:, 2:5]).max()
    dp = np.abs(got[:, 5:8] - w[:, 5:8]).max()
    assert dq <= pos_tol and dp <= mom_tol, (dq, dp)
    return dq, dp


def test_benchmark_project_matches_reference_program(tmp_path):
    """Benchmark project (NVE, vacancy FCC, |v0| = 1): t, T, P, E every 100 steps to the
    last printed digit (relative 1e-5 of each value's scale)."""
    gpu, cpu, _, _ = run_gpu("benchmark", tmp_path)
    assert cpu.shape == gpu.shape == (4, 4), (cpu, gpu)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


def test_benchmark_project_mixed_precision(tmp_path):
    """The same with FP32 pair forces: the trajectory leaves the FP64 one slowly, and T, P,
    E agree to within 1e-4 of their scale after 400 steps."""
    gpu, cpu, _, _ = run_gpu("benchmark", tmp_path, {"MDACP_GPU_PRECISION": "mixed"})
    assert cpu.shape == gpu.shape
    np.testing.assert_allclose(gpu[:, 0], cpu[:, 0])
    np.testing.assert_allclose(gpu[:, 1:], cpu[:, 1:], rtol=1e-4, atol=1e-4)


def test_nose_hoover_benchmark_matches_reference_program(tmp_path):
    """Benchmark project with the Nose-Hoover heat bath (two temperature reductions a step)."""
    gpu, cpu, _, _ = run_gpu("nosehoover", tmp_path)
    assert cpu.shape == gpu.shape == (3, 4)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


def test_langevin_project_thermalises(tmp_path):
    """Langevin project: the noise stream differs from the reference's (Philox on the
    device, std::mt19937 on the host), so the runs agree in distribution only: same
    first observation, and both relax towards T = 0.9."""
    gpu, cpu, _, _ = run_gpu("langevin", tmp_path)
    assert cpu.shape == gpu.shape == (30, 4)
    np.testing.assert_allclose(gpu[:, 0], cpu[:, 0])
    assert abs(gpu[0, 1] - cpu[0, 1]) < 0.02 * cpu[0, 1]
    t_gpu, t_cpu = gpu[15:, 1].mean(), cpu[15:, 1].mean()
    assert abs(t_gpu - 0.9) < 0.06 and abs(t_cpu - 0.9) < 0.06, (t_gpu, t_cpu)


def test_collision_project_and_cdview_dumps(tmp_path):
    """Collision project (collision.cc:53-79): two droplets at +-1 along x, cdview dumps at
    the start and at every observation.  The initial dump is the same particles to the printed
    digit; later dumps pair one to one within the printed precision."""
    gpu, cpu, got, want = run_gpu("collision", tmp_path)
    assert cpu.shape == gpu.shape == (4, 2)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)
    assert sorted(got) == sorted(want) == [f"conf{i:04d}.cd" for i in range(5)]
    for fn in sorted(want):
        dq, dp = match_dump(got[fn], want[fn], pos_tol=1.5e-4, mom_tol=1.5e-5)
        print(fn, "max |dq|", dq, "max |dp|", dp, flush=True)


def test_burst_project_and_cdview_dumps(tmp_path):
    """Burst project (burst.cc:103-142): a droplet thermalised with the Nose-Hoover bath, then
    NVE: t, T, P, E rows and the cdview dumps of both phases."""
    gpu, cpu, got, want = run_gpu("burst", tmp_path)
    assert cpu.shape == gpu.shape == (3, 4)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)
    assert sorted(got) == sorted(want)
    for fn in sorted(want):
        dq, dp = match_dump(got[fn], want[fn], pos_tol=1.5e-4, mom_tol=1.5e-5)
        print(fn, "max |dq|", dq, "max |dp|", dp, flush=True)


def test_benchmark_project_with_eight_units(tmp_path):
    """OMP_NUM_THREADS=8: MDManager makes 2x2x2 units in the process (ParaInfo), each a device
    unit of one group on the GPU; particles migrate and ghosts come device to device.  The
    Benchmark project's configuration and initial velocities are per unit
    (ExtractConfigurationMaker seeds each unit's lattice vacancies with its id,
    confmaker.cc:174-230), so this run is its own trajectory: the reference's own 8-thread run
    prints the fixture's rows, to the last printed digit."""
    gpu, cpu, _, _ = run_gpu("benchmark_omp8", tmp_path)
    assert cpu.shape == gpu.shape == (4, 4), (cpu, gpu)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


def test_nose_hoover_with_eight_units(tmp_path):
    """The same with the Nose-Hoover bath: zeta from the temperature summed over the units."""
    gpu, cpu, _, _ = run_gpu("nosehoover_omp8", tmp_path)
    assert cpu.shape == gpu.shape == (3, 4)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)


def test_collision_with_four_units(tmp_path):
    """Collision project on four units: the droplets cross unit faces; rows and cdview dumps
    (written unit after unit, matched by position)."""
    gpu, cpu, got, want = run_gpu("collision_omp4", tmp_path)
    assert cpu.shape == gpu.shape == (4, 2)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)
    assert sorted(got) == sorted(want)
    for fn in sorted(want):
        dq, dp = match_dump(got[fn], want[fn], pos_tol=1.5e-4, mom_tol=1.5e-5)
        print(fn, "max |dq|", dq, "max |dp|", dp, flush=True)


def test_burst_with_eight_units(tmp_path):
    """Burst project on eight units: Nose-Hoover thermalisation, then NVE, rows and dumps."""
    gpu, cpu, got, want = run_gpu("burst_omp8", tmp_path)
    assert cpu.shape == gpu.shape == (3, 4)
    np.testing.assert_allclose(gpu, cpu, rtol=2e-5, atol=2e-5)
    assert sorted(got) == sorted(want)
    for fn in sorted(want):
        dq, dp = match_dump(got[fn], want[fn], pos_tol=1.5e-4, mom_tol=1.5e-5)
        print(fn, "max |dq|", dq, "max |dp|", dp, flush=True)


def test_facade_binary_dump_matches_reference_fixture():
    """MDManager::SaveAsDumpFile (mdmanager.cc:474-504), compiled unmodified against the
    facade, writes the device state through the facade's host mirrors: after 100 steps of the
    fcc2048_nve case, the same particles as the reference's own dump
    (tests/golden/fcc2048_nve_dump100.npy), within 1e-9 (the two programs sum their kicks in
    different orders)."""
    import sys

    from scipy.spatial import cKDTree

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import load_golden
    from mdacp_amd.lattice import fcc, initial_velocity
    from oracle import oracle as O

    if not os.path.exists(O.FACADE_DRIVER_BIN):
        pytest.skip("oracle/_ref/mdacp_gpu_driver not built (make -C oracle facade_driver)")
    g = load_golden("fcc2048_nve")
    x = fcc(g["L"], g["density"])
    n = len(x)
    print("running mdacp_gpu_driver fcc2048_nve, binary dump after 100 steps", flush=True)
    _, got = O.run_reference(g["L"], x, initial_velocity(n, g["v0"], 0), np.arange(n, dtype=np.int32),
                             steps=100, dump_every=100, dt=g["dt"], binary=O.FACADE_DRIVER_BIN, kick=False,
                             timeout=240, dump_file_step=100)
    want = np.load(os.path.join(ROOT, "tests", "golden", "fcc2048_nve_dump100.npy"))
    assert got.shape == want.shape == (n, 6)
    dist, idx = cKDTree(want[:, :3]).query(got[:, :3])
    assert len(np.unique(idx)) == n, "positions do not pair one to one"
    dq = np.abs(got[:, :3] - want[idx, :3]).max()
    dp = np.abs(got[:, 3:] - want[idx, 3:]).max()
    print("max |dq|", dq, "max |dp|", dp, flush=True)
    assert dq <= 1e-9 and dp <= 1e-9, (dq, dp)


def test_facade_list_accessors_match_reference_fixture():
    """The reference's MDManager on the facade, driven by oracle/ref_driver.cc linked against
    it (oracle/_ref/mdacp_gpu_driver): the pair lists read through the facade's COO accessors
    (GetKeyParticles / GetPartnerParticles) and its CSR accessors (GetPartnerNumber /
    GetKeyPointer / GetSortedList, meshlist.h:46-53) both equal the reference's own pair sets
    (SHA-256 of the canonical set, tests/golden/fcc2048_nve.json) at every dump, with the
    reference's rebuild steps and T, P, E within 1e-9."""
    import math
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import load_golden, pair_hash
    from mdacp_amd.lattice import fcc, initial_velocity
    from oracle import oracle as O

    if not os.path.exists(O.FACADE_DRIVER_BIN):
        pytest.skip("oracle/_ref/mdacp_gpu_driver not built (make -C oracle facade_driver)")
    g = load_golden("fcc2048_nve")
    x = fcc(g["L"], g["density"])
    n = len(x)
    print("running mdacp_gpu_driver fcc2048_nve", flush=True)
    recs = O.run_reference(g["L"], x, initial_velocity(n, g["v0"], 0), np.arange(n, dtype=np.int32),
                           steps=g["steps"], dump_every=50, dt=g["dt"], binary=O.FACADE_DRIVER_BIN, kick=False,
                           timeout=240)
    dumps = {d["step"]: d for d in g["dumps"]}
    assert [r["step"] for r in recs if r["rebuilt"]] == g["rebuild_steps"]
    seen = 0
    for r in recs:
        if not r["dump"]:
            continue
        want = dumps[r["step"]]
        assert pair_hash(r["pairs"]) == (want["pair_sha256"], want["pair_count"]), r["step"]
        assert pair_hash(r["pairs_csr"]) == (want["pair_sha256"], want["pair_count"]), r["step"]
        for k in ("T", "P", "E"):
            assert r[k] == pytest.approx(want[k], rel=1e-9), (r["step"], k)
        seen += 1
    assert seen == len(g["dumps"])
