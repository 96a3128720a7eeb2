"""GPU parity at the benchmark configurations' stated parameters.

A  131,072 particles (FCC 4*32^3, rho 0.5, L 64), |v0| = sqrt(3), rc 2.5, margin 0.5,
   dt 0.001, 1000 NVE steps: T, P, E and the raw observer sums every 100 steps within
   1e-9 relative of the oracle's trajectory (tests/golden/configA_131072.json), the same
   rebuild steps, and the pair set bit-exact (SHA-256 of the canonical set) after every
   construction.
B  4,000,000 particles (L 200), otherwise A: the trace to step 1000 within 1e-9, the same
   rebuild steps, the pair set at steps 0 and 1000 equal to the oracle's along its own
   trajectory (order-independent checksum, tests/golden/configB_4M.json).  At step 0 (the
   perfect crystal) PE and virial per particle equal the closed-form lattice sums to
   1e-12; the oracle's own sums there are only good to about 1e-9: it adds the ~10^8 list
   terms in one running sum in list order, as the reference's observers do
   (observer.cc:18-79), and on a crystal every term of a shell rounds the same way, so
   the roundings add up instead of averaging out.  Against the oracle step 0 is held to
   1e-8.
C  mixed precision on the 4,000,000-particle system: the per-particle force RMS error
   sqrt(sum |F32 - F64|^2 / sum |F64|^2) against the FP64 oracle's kicks at 10 snapshots
   (the FP64 trajectory at steps 1000 k, k = 1..10) below 2e-6 with cell-relative FP32
   coordinates and below 4e-6 with brick-centred ones, and the FP64 GPU kick against the
   same oracle kicks below 1e-13 in the same measure.
D  the 256 replicas of tests/golden/ensemble_replicas.json (4.45M particles, Langevin,
   gamma 1, dt 0.005), FP64 and mixed: at the start (perfect FCC crystals) every
   replica's PE and virial per particle within 1e-11 of the closed-form lattice sums
   (except where the reference's cell grid misses a ghost layer, below),
   KE within 1e-13 and PE and virial within 1e-9 of the oracle (whose running sum in
   list order carries rounding errors of that size on a crystal, where every pair of a
   shell rounds alike), the pair sets of 32 replicas bit-exact; after 5000 steps every
   replica near its bath temperature and the pair sets of 8 replicas bit-exact with the
   oracle rebuilt from the GPU state.  MeshList::ChangeScale (meshlist.cc:30-55) sizes
   the grid as int(L / (L / ms)) + 2 cells per axis, which rounds to ms + 1 for some
   boxes (replica 165: L = 47.064..., ms = 15): the +L images then wrap into the lower
   ghost layer (meshlist.cc:167-172) and the pairs across that face are not listed.  The
   GPU reproduces this; such replicas are checked by their exact pair sets instead of the
   lattice sums.
E  16,384,000 particles (FCC 4*160^3 at rho 0.75), MDManager::ChangeScale to rho 0.6:
   the pair set equal to the oracle's (tests/golden/configE_16M.json), PE and virial
   per particle equal to the closed-form lattice sums at the rescaled spacing (1e-12; the
   oracle's running sums over 6.5e8 crystal terms are good to about 1e-8 and are held to
   that), then 200 mixed-precision NVE steps with |dE/N| below 1e-4.
"""
import json
import math
import os
import sys

import numpy as np
import pytest

from helpers import GOLDEN, fcc_lattice_sums, load_golden, pair_checksum, pair_hash, wrap_like_reference
from mdacp_amd import mdgpu
from mdacp_amd.ensemble import box_for, build
from mdacp_amd.lattice import fcc, initial_velocity
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RC, MARGIN, DT = 2.5, 0.5, 0.001


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nve_domain(g, precision=mdgpu.FP64):
    x = fcc(g["L"], g["density"])
    n = len(x)
    assert n == g["n"]
    p = initial_velocity(n, g["v0"], g["seed"])
    d = mdgpu.Domain(g["L"], dt=g["dt"], rc=g["rc"], margin=g["margin"], precision=precision, capacity=n)
    d.add_particles(x, p)
    d.make_pair_list()
    return d, x, p


def check_dump(ob, want, step, tol=1e-9):
    for k in ("ke", "pe", "vir", "T", "P", "E"):
        assert ob[k] == pytest.approx(want[k], rel=tol), (step, k, ob[k], want[k])


def test_config_A_131k_trace_rebuilds_and_pairs():
    say("case: config A, 131072 particles, 1000 steps")
    g = load_golden("configA_131072")
    d, _, _ = nve_domain(g)
    assert pair_hash(d.pairs()) == tuple(g["pairs"]["0"])
    dumps = {dd["step"]: dd for dd in g["dumps"]}
    check_dump(d.observables(), dumps[0], 0)
    rebuilt = [0]
    worst = 0.0
    for step in range(1, g["steps"] + 1):
        if d.calculate(1):
            rebuilt.append(step)
            assert pair_hash(d.pairs()) == tuple(g["pairs"][str(step)]), step
        if step in dumps:
            ob = d.observables()
            check_dump(ob, dumps[step], step)
            worst = max(worst, max(abs(ob[k] / dumps[step][k] - 1.0) for k in ("T", "P", "E")))
    assert rebuilt == g["rebuild_steps"]
    say("config A rebuilds", len(rebuilt), "worst T/P/E relative deviation", worst)
    d.close()


def test_config_B_4M_trace_and_pairs():
    say("case: config B, 4M particles, 1000 steps")
    g = load_golden("configB_4M")
    d, _, _ = nve_domain(g)
    assert pair_checksum(d.pairs()) == g["pairs"]["0"]
    dumps = {dd["step"]: dd for dd in g["dumps"]}
    ob0 = d.observables()
    pe_cf, vir_cf, _ = fcc_lattice_sums((4.0 / g["density"]) ** (1.0 / 3.0), g["rc"], g["rc"] + g["margin"])
    assert ob0["pe"] / g["n"] == pytest.approx(pe_cf, rel=1e-12)
    assert ob0["vir"] / g["n"] == pytest.approx(vir_cf, rel=1e-12)
    check_dump(ob0, dumps[0], 0, tol=1e-8)
    rebuilt = [0]
    worst = 0.0
    for step in range(1, g["steps"] + 1):
        if d.calculate(1):
            rebuilt.append(step)
        if step in dumps:
            ob = d.observables()
            check_dump(ob, dumps[step], step)
            worst = max(worst, max(abs(ob[k] / dumps[step][k] - 1.0) for k in ("T", "P", "E")))
    assert rebuilt == g["rebuild_steps"]
    say("config B rebuilds", rebuilt, "worst T/P/E relative deviation", worst)
    assert pair_checksum(d.pairs()) == g["pairs"]["1000"]
    d.close()


def rms_rel(a, b):
    return math.sqrt(float(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_config_C_mixed_force_error_against_fp64_oracle_4M():
    say("case: config C, mixed precision at 4M against the FP64 oracle, 10 snapshots")
    L = 200.0
    x = fcc(L, 0.5)
    n = len(x)
    d64 = mdgpu.Domain(L, dt=DT, rc=RC, margin=MARGIN, capacity=n)
    d64.add_particles(x, initial_velocity(n, math.sqrt(3.0), 0))
    d64.make_pair_list()
    errs = []
    for k in range(1, 11):
        d64.calculate(1000)
        q, p, t = d64.particles()
        q = wrap_like_reference(q, L)
        o = O.OracleSystem(L, DT, RC, MARGIN)
        o.add(q, p, t)
        o.make_pair_list()
        ko = o.kick()[np.argsort(o.state()[2], kind="stable")]
        del o
        d32 = mdgpu.Domain(L, dt=DT, rc=RC, margin=MARGIN, precision=mdgpu.MIXED, capacity=n)
        d32.add_particles(q, p, t)
        d32.make_pair_list()
        order = np.argsort(d32.particles()[2], kind="stable")
        d32.set_option(mdgpu.OPT_MIXED_FRAME, 0)
        k32 = d32.kick()[order]
        d32.set_option(mdgpu.OPT_MIXED_FRAME, 1)
        k32b = d32.kick()[order]
        d32.set_precision(mdgpu.FP64)
        k64 = d32.kick()[order]
        d32.close()
        e32, e32b, e64 = rms_rel(k32, ko), rms_rel(k32b, ko), rms_rel(k64, ko)
        errs.append((e32, e32b))
        say(f"snapshot step {1000 * k}: mixed RMS {e32:.3e} (cell frame), {e32b:.3e} (brick frame), "
            f"FP64 RMS {e64:.3e}")
        assert e32 < 2e-6, (k, e32)
        assert e32b < 4e-6, (k, e32b)
        assert e64 < 1e-13, (k, e64)
    errs = np.array(errs)
    say("config C mixed force RMS relative error, 10 snapshots: max", errs.max(axis=0), "mean", errs.mean(axis=0))
    d64.close()


def replica_rows():
    with open(os.path.join(GOLDEN, "ensemble_replicas.json")) as f:
        return json.load(f)["rows"]


def split(a, counts):
    return np.split(a, np.cumsum(counts)[:-1])


@pytest.mark.parametrize("precision", [mdgpu.FP64, mdgpu.MIXED])
def test_config_D_256_replicas(precision):
    say("case: config D, 256 replicas, precision", precision)
    rows = replica_rows()
    boxes, counts, q, p, ids = build(rows)
    assert len(rows) == 256 and int(counts.sum()) == 4451972
    e = mdgpu.Ensemble(boxes, counts, q, p, ids, T=[r["T"] for r in rows], gamma=1.0, seed=[r["seed"] for r in rows],
                       thermostat=mdgpu.LANGEVIN, dt=0.005, rc=RC, margin=MARGIN, precision=precision)
    ke, pe, vir = e.observe_replicas_raw()
    short_grid = set()
    for r, row in enumerate(rows):
        L = boxes[r, 0]
        ms = int(L / (RC + MARGIN))
        if int(L / (L / ms)) + 2 != ms + 2:
            short_grid.add(r)
            continue
        pe_cf, vir_cf, _ = fcc_lattice_sums((4.0 / row["rho"]) ** (1.0 / 3.0), RC, RC + MARGIN)
        assert pe[r] / counts[r] == pytest.approx(pe_cf, rel=1e-11), r
        assert vir[r] / counts[r] == pytest.approx(vir_cf, rel=1e-11), r
    assert short_grid == {165}
    # 32 replicas spread over size and density, and the short-grid one, for the exact pair sets
    order = np.lexsort((np.array([r["rho"] for r in rows]), counts))
    exact = set(int(r) for r in order[np.linspace(0, 255, 32).astype(int)]) | short_grid
    for r, (qr, pr, ir) in enumerate(zip(split(q, counts), split(p, counts), split(ids, counts))):
        o = O.OracleSystem(boxes[r, 0], 0.005, RC, MARGIN)
        o.add(qr, pr, ir)
        o.make_pair_list()
        ob = o.observables()
        assert ke[r] == pytest.approx(ob["ke"], rel=1e-13), r
        assert pe[r] == pytest.approx(ob["pe"], rel=1e-9), r
        assert vir[r] == pytest.approx(ob["vir"], rel=1e-9), r
        if r in exact:
            assert np.array_equal(O.canonical_pairs(e.replica_pairs(r)), O.canonical_pairs(o.pairs())), r
    say("config D step 0: observers of 256 replicas match the lattice sums and the oracle; 32 pair sets exact")
    rebuilds = sum(e.calculate(500) for _ in range(10))
    say("config D 5000 steps, constructions", rebuilds)
    T = e.observables()["T"]
    Tt = np.array([r["T"] for r in rows])
    # Langevin map p += (-gamma p + noise) dt: stationary <p^2> = T / (1 - gamma dt / 2); the
    # instantaneous temperature of N particles scatters by sqrt(2 / (3 N)) relative
    dev = np.abs(T / (Tt / (1.0 - 0.5 * 0.005)) - 1.0)
    bound = 6.0 * np.sqrt(2.0 / (3.0 * counts)) + 0.01
    assert np.all(dev < bound), (np.argmax(dev / bound), float(np.max(dev / bound)))
    say("config D temperatures: max |T/T_bath - 1| / bound", float(np.max(dev / bound)))
    eq, ep, et = e.particles()
    sl = e.replica_slices()
    for r in sorted(exact)[::4]:
        o = O.OracleSystem(boxes[r, 0], 0.005, RC, MARGIN)
        o.add(wrap_like_reference(eq[sl[r]], boxes[r, 0]), ep[sl[r]], et[sl[r]])
        o.make_pair_list()
        e_single = mdgpu.Domain(boxes[r, 0], dt=0.005, rc=RC, margin=MARGIN, precision=precision)
        e_single.add_particles(wrap_like_reference(eq[sl[r]], boxes[r, 0]), ep[sl[r]], et[sl[r]])
        e_single.make_pair_list()
        assert np.array_equal(O.canonical_pairs(e_single.pairs()), O.canonical_pairs(o.pairs())), r
        e_single.close()
    e.close()


def test_config_E_16M_change_scale_pairs_and_closed_form():
    say("case: config E, 16.4M particles, ChangeScale rho 0.75 -> 0.6")
    g = load_golden("configE_16M")
    k = 160
    L = box_for(k, g["density"])
    assert L == g["L"]
    x = fcc(L, g["density"])
    n = len(x)
    assert n == g["n"] == 4 * k ** 3
    p = initial_velocity(n, g["v0"], g["seed"])
    d = mdgpu.Domain(L, dt=DT, rc=RC, margin=MARGIN, precision=mdgpu.MIXED, capacity=n)
    d.add_particles(x, p)
    del x, p
    d.make_pair_list()
    d.change_scale(g["alpha"])
    pr = d.pairs()
    assert pair_checksum(pr) == g["pairs"]
    own = pr[:, 2] == 13
    half_pairs = int(own.sum()) + int((~own).sum()) // 2
    del pr, own
    a = (4.0 / g["density"]) ** (1.0 / 3.0) * g["alpha"]  # the rescaled lattice's cubic cell
    pe_cf, vir_cf, half_cf = fcc_lattice_sums(a, RC, RC + MARGIN)
    assert half_pairs == half_cf * n
    ke, pe, vir = d.observe_raw()
    assert pe / n == pytest.approx(pe_cf, rel=1e-12)
    assert vir / n == pytest.approx(vir_cf, rel=1e-12)
    assert ke == pytest.approx(g["observe"]["ke"], rel=1e-12)
    for key, v in (("pe", pe), ("vir", vir)):
        assert v == pytest.approx(g["observe"][key], rel=1e-8), key
    e0 = d.observables()["E"]
    d.calculate(200)
    e1 = d.observables()["E"]
    say("config E after ChangeScale: PE/N", pe / n, "half pairs/N", half_cf, "dE/N over 200 mixed steps", e1 - e0)
    assert abs(e1 - e0) < 1e-4
    d.close()
