"""GPU parity: the HIP path (libmdgpu.so, through its C ABI) against the oracle
and the reference's golden fixtures.

Bars: pair sets bit-exact as sets of canonical (i, j, image) pairs; FP64 kicks
within 1e-12 of the oracle per particle, relative to the sum of the pair-kick
magnitudes on that particle; observables within 1e-12 relative on the same
state; trajectories (chaotic, different summation order) within 1e-9 relative
on T, P, E after 200 steps and 1e-7 after 1000, with identical rebuild steps.
"""
import math
import sys

import numpy as np
import pytest

from helpers import lj_C, load_golden, pair_force_sums, pair_hash, wrap_like_reference
from mdacp_amd import mdgpu
from mdacp_amd.lattice import fcc, initial_velocity
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(L, x, p, ids, dt=0.001, rc=2.5, margin=0.5, precision=mdgpu.FP64):
    d = mdgpu.Domain(L, dt=dt, rc=rc, margin=margin, precision=precision, capacity=len(x))
    d.add_particles(x, p, ids)
    o = O.OracleSystem(L, dt, rc, margin)
    o.add(x, p, ids)
    return d, o


def melted_state(L, rho, steps=300, seed=0):
    """A fluid configuration from the oracle (dt 0.005 from a hot lattice)."""
    x = fcc(L, rho)
    n = len(x)
    rng = np.random.default_rng(seed)
    o = O.OracleSystem(L, 0.005, 2.5, 0.5)
    o.add(x, rng.normal(0, 1.5, (n, 3)), rng.permutation(n).astype(np.int32))
    o.make_pair_list()
    for _ in range(steps):
        o.step()
    q, p, t = o.state()
    return np.mod(q, L), p, t


def compare_state(d, o, L, rc, dt, label):
    """Pairs exact, ghosts exact, kicks 1e-12, observables 1e-12 on the same built state."""
    gp, op = d.pairs(), o.pairs()
    assert np.array_equal(O.canonical_pairs(gp), O.canonical_pairs(op)), label
    gs, gc, gq = d.ghosts()
    os_, oc, oq = o.ghosts()
    ka = np.lexsort((gc, gs)); kb = np.lexsort((oc, os_))
    assert np.array_equal(gs[ka], os_[kb]) and np.array_equal(gc[ka], oc[kb]), label
    assert np.array_equal(gq[ka], oq[kb]), label
    q, p, t = d.particles()
    oq_, opp, ot = o.state()
    ia, ib = np.argsort(t), np.argsort(ot)
    assert np.array_equal(q[ia], oq_[ib]) and np.array_equal(p[ia], opp[ib]), label
    kg = d.kick()[ia]
    ko = o.kick()[ib]
    C0, C2 = lj_C(rc)
    S = pair_force_sums(oq_[ib], op.astype(np.int64), L, rc, dt, C2)
    err = np.abs(kg - ko).max(axis=1)
    assert np.all(err <= 1e-12 * np.maximum(S, 1e-300) + 1e-300), (label, float((err / np.maximum(S, 1e-300)).max()))
    og, oo = d.observables(), o.observables()
    assert og["ke"] == pytest.approx(oo["ke"], rel=1e-13), label
    assert og["pe"] == pytest.approx(oo["pe"], rel=1e-12), label
    assert og["vir"] == pytest.approx(oo["vir"], rel=1e-12), label
    say(label, "pairs", len(gp), "ghosts", len(gs), "max kick err/S", float((err / np.maximum(S, 1e-300)).max()))


@pytest.mark.parametrize("rc,margin", [(3.0, 0.3), (2.5, 0.5)])
def test_built_state_parity_lattice_and_fluid(rc, margin):
    say("case: built-state parity rc", rc, "margin", margin)
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    for label, (xx, pp, ids) in {
        "fcc2048": (x, initial_velocity(n, math.sqrt(3.0), 0), np.arange(n, dtype=np.int32)),
        "fluid2048": melted_state(L, 0.5),
    }.items():
        d, o = make_pair(L, xx, pp, ids, rc=rc, margin=margin)
        d.make_pair_list()
        o.make_pair_list()
        compare_state(d, o, L, rc, 0.001, f"{label} rc={rc}")
        d.close()


def test_built_state_parity_uneven_box_and_edges():
    say("case: uneven box, particles on faces")
    L = (13.7, 19.0, 11.05)
    rng = np.random.default_rng(11)
    n = 900
    x = rng.uniform(0, 1, (n, 3)) * np.array(L)
    x[:20, 0] = 0.0  # on the low face: no image at +L (strict test)
    x[20:40, 1] = np.nextafter(L[1], 0.0)
    ids = rng.permutation(n).astype(np.int32)
    d, o = make_pair(L, x, None, ids, rc=2.5, margin=0.5)
    d.make_pair_list()
    o.make_pair_list()
    compare_state(d, o, np.array(L), 2.5, 0.001, "uneven")


@pytest.mark.parametrize("cells,stage", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_search_kernels_agree_with_oracle(cells, stage):
    """Both neighbour-search kernels, with and without LDS staging, list the oracle's pairs on a
    fluid with a dense cluster (a cell of more than 64 rows, masks longer than the first guess)
    and particles on the upper faces (binned into the upper ghost layer)."""
    say("case: search kernel cells", cells, "stage", stage)
    L = 16.0
    q, p, ids = melted_state(L, 0.5)
    rng = np.random.default_rng(5)
    cl = 7.3 + rng.uniform(0.0, 0.6, (100, 3))
    q = np.concatenate([q, cl])
    q[:6, 2] = L
    q[6:9, 0] = L
    p = np.concatenate([p, np.zeros((100, 3))])
    ids = np.concatenate([ids, np.arange(len(ids), len(ids) + 100, dtype=np.int32)])
    d, o = make_pair(L, q, p, ids)
    d.set_option(mdgpu.OPT_SEARCH_CELLS, cells)
    d.set_option(mdgpu.OPT_SEARCH_STAGE, stage)
    d.make_pair_list()
    o.make_pair_list()
    gp, op = d.pairs(), o.pairs()
    assert np.array_equal(O.canonical_pairs(gp), O.canonical_pairs(op))
    say("pairs", len(gp))
    d.close()


def test_layout_options_keep_pairs_and_kicks():
    """Row order (MDGPU_OPT_ROW_SORT), per-axis brick edges (MDGPU_OPT_BRICK) and the brick LDS
    budget (MDGPU_OPT_BRICK_KB) change where rows and partners sit, never the pairs: the oracle's
    pair set every time; kicks bit-identical across row orders (a row sums its partners in the
    same order wherever it sits) and within 1e-12 across bricks (the partners' order changes)."""
    say("case: layout options")
    L = 21.0
    q, p, ids = melted_state(L, 0.6, steps=150, seed=4)
    d, o = make_pair(L, q, p, ids)
    o.make_pair_list()
    op = O.canonical_pairs(o.pairs())
    ko = o.kick()[np.argsort(o.state()[2])]
    C0, C2 = lj_C(2.5)
    qo, _, to = o.state()
    S = pair_force_sums(qo[np.argsort(to)], o.pairs().astype(np.int64), L, 2.5, 0.001, C2)
    kicks = {}
    for name, opt, value in [("sort8", mdgpu.OPT_ROW_SORT, 8), ("sort1", mdgpu.OPT_ROW_SORT, 1),
                             ("sort0", mdgpu.OPT_ROW_SORT, 0), ("sort16", mdgpu.OPT_ROW_SORT, 16),
                             ("brick345", mdgpu.OPT_BRICK, 3 | 4 << 8 | 5 << 16),
                             ("brick2", mdgpu.OPT_BRICK, 2 | 2 << 8 | 2 << 16),
                             ("kb40", mdgpu.OPT_BRICK_KB, 40), ("kb128", mdgpu.OPT_BRICK_KB, 128)]:
        d.set_option(opt, value)
        d.make_pair_list()
        assert np.array_equal(O.canonical_pairs(d.pairs()), op), name
        t = d.particles()[2]
        kicks[name] = d.kick()[np.argsort(t)]
        err = np.abs(kicks[name] - ko).max(axis=1)
        assert np.all(err <= 1e-12 * np.maximum(S, 1e-300) + 1e-300), name
        if name.startswith("sort"):
            d.set_option(mdgpu.OPT_ROW_SORT, 8)
    for name in ("sort1", "sort0", "sort16"):
        assert np.array_equal(kicks[name], kicks["sort8"]), name
    say("pairs", len(op), "bricks", d.get_option(mdgpu.OPT_BRICK) & 0xff)
    d.close()


def test_golden_pairs_and_kicks_fcc2048(golden_dir):
    say("case: golden fixture fcc2048 step 0")
    g = load_golden("fcc2048_nve")
    x = fcc(g["L"], g["density"])
    n = len(x)
    d = mdgpu.Domain(g["L"], dt=g["dt"], rc=g["rc"], margin=g["margin"])
    d.add_particles(x, initial_velocity(n, g["v0"], 0))
    d.make_pair_list()
    assert pair_hash(d.pairs()) == (g["dumps"][0]["pair_sha256"], g["dumps"][0]["pair_count"])
    _, _, t = d.particles()
    kick = d.kick()[np.argsort(t)]
    ref = np.load(f"{golden_dir}/fcc2048_nve_kick0.npy")
    assert np.abs(kick - ref).max() < 1e-15  # perfect lattice: all kicks ~0
    ob = d.observables()
    assert ob["pe"] == pytest.approx(g["dumps"][0]["pe"], rel=1e-12)
    # E = KE/N + PE/N cancels: bound it by the scale of its terms
    scale = (abs(ob["ke"]) + abs(ob["pe"])) / n
    assert ob["E"] == pytest.approx(g["dumps"][0]["E"], abs=1e-12 * scale)


def run_gpu_case(g, steps=None, precision=mdgpu.FP64):
    steps = g["steps"] if steps is None else steps
    x = fcc(g["L"], g["density"])
    n = len(x)
    p = initial_velocity(n, g["v0"], 0)
    ids = np.arange(n, dtype=np.int32)
    if g["reverse_ids"]:
        ids = ids[::-1].copy()
    d = mdgpu.Domain(g["L"], dt=g["dt"], rc=g["rc"], margin=g["margin"], precision=precision)
    d.add_particles(x, p, ids)
    kind = {"nve": mdgpu.NVE, "langevin": mdgpu.LANGEVIN, "nosehoover": mdgpu.NOSEHOOVER}[g["thermostat"]]
    d.set_thermostat(kind, g["T"], g["gamma"], g["tau"])
    d.make_pair_list()
    dumps = {dd["step"]: dd for dd in g["dumps"]}
    rebuilt = [0]
    results = {}
    for step in range(1, steps + 1):
        if d.calculate(1):
            rebuilt.append(step)
        if step in dumps:
            results[step] = d.observables()
    return d, rebuilt, results, dumps


@pytest.mark.parametrize("name,tol", [("fcc2048_nve", 1e-9), ("fcc2048_nosehoover", 1e-8)])
def test_trajectory_against_reference_fixture(name, tol):
    say("case: trajectory", name)
    g = load_golden(name)
    d, rebuilt, res, dumps = run_gpu_case(g)
    assert rebuilt == g["rebuild_steps"]
    for step, ob in res.items():
        for k in ("T", "P", "E"):
            assert ob[k] == pytest.approx(dumps[step][k], rel=tol), (step, k)
    say(name, "rebuilds", rebuilt)


def test_trajectory_131k_1000_steps_against_reference_fixture():
    say("case: trajectory fcc131072 1000 steps")
    g = load_golden("fcc131072_nve")
    d, rebuilt, res, dumps = run_gpu_case(g)
    assert rebuilt == g["rebuild_steps"]
    worst = 0.0
    for step, ob in res.items():
        for k in ("T", "P", "E"):
            rel = abs(ob[k] - dumps[step][k]) / abs(dumps[step][k])
            worst = max(worst, rel)
            assert rel < 1e-7, (step, k, ob[k], dumps[step][k])
    say("131k worst relative T/P/E deviation", worst)
    # pair list rebuilt from the GPU's final state is bit-exact with the oracle's on the same state
    q, p, t = d.particles()
    o = O.OracleSystem(g["L"], g["dt"], g["rc"], g["margin"])
    o.add(wrap_like_reference(q, g["L"]), p, t)
    d.make_pair_list()
    o.make_pair_list()
    assert np.array_equal(O.canonical_pairs(d.pairs()), O.canonical_pairs(o.pairs()))


def test_langevin_philox_against_oracle():
    say("case: Langevin with counter-based noise")
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    p = initial_velocity(n, 1.0, 0)
    ids = np.arange(n, dtype=np.int32)[::-1].copy()
    d, o = make_pair(L, x, p, ids, dt=0.005, rc=2.5, margin=0.5)
    d.set_thermostat(mdgpu.LANGEVIN, 0.9, 0.5, 0.1)
    d.set_seed(7)
    o.set_thermostat(O.LANGEVIN, 0.9, 0.5, 0.1)
    o.set_noise(O.NOISE_PHILOX, 7)
    d.make_pair_list()
    o.make_pair_list()
    for step in range(1, 121):
        assert bool(d.calculate(1)) == o.step(), step
        if step % 40 == 0:
            og, oo = d.observables(), o.observables()
            for k in ("T", "P", "E"):
                assert og[k] == pytest.approx(oo[k], rel=1e-9), (step, k)


def test_piecewise_api_equals_fused_step():
    say("case: piecewise vs fused")
    L = 16.0
    xx, pp, ids = melted_state(L, 0.5)
    a = mdgpu.Domain(L, dt=0.002)
    b = mdgpu.Domain(L, dt=0.002)
    for d in (a, b):
        d.add_particles(xx, pp, ids)
        d.make_pair_list()
    for _ in range(60):
        a.calculate(1)
        # MDManager::Calculate, piece by piece
        if b.is_pair_list_expired():
            b.make_pair_list()
        b.update_position_half()
        b.send_border_particles()
        b.calculate_force()
        b.update_position_half()
    qa, pa, ta = a.particles_by_id()
    qb, pb, tb = b.particles_by_id()
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)


@pytest.mark.parametrize("thermostat,precision,frame", [("nve", mdgpu.FP64, 0), ("langevin", mdgpu.FP64, 0),
                                                        ("nve", mdgpu.MIXED, 0), ("nve", mdgpu.MIXED, 1),
                                                        ("langevin", mdgpu.MIXED, 1)])
def test_runs_ahead_match_steps_one_at_a_time(thermostat, precision, frame):
    """Steps queued ahead of the host (the device-side validity gate, MDGPU_OPT_RUN_AHEAD) give
    the same bits, time and constructions as steps that each wait for the host: the gate does
    the budget arithmetic the host does and stops the run where the host must take over."""
    say("case: run ahead", thermostat, precision, frame)
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0 * 1.5), 3)
    ds = []
    for ahead in (64, 0):
        d = mdgpu.Domain(L, dt=0.005, precision=precision, capacity=n)
        d.set_option(mdgpu.OPT_MIXED_FRAME, frame)
        d.add_particles(x, p)
        if thermostat == "langevin":
            d.set_thermostat(mdgpu.LANGEVIN, 1.2, 1.0, 0.1)
            d.set_seed(11)
        d.set_option(mdgpu.OPT_RUN_AHEAD, ahead)
        ds.append(d)
    a, b = ds
    assert a.get_option(mdgpu.OPT_RUN_AHEAD) == 64 and b.get_option(mdgpu.OPT_RUN_AHEAD) == 0
    ra = rb = 0
    for chunk in (1, 37, 100, 2, 60):
        ra += a.calculate(chunk)
        rb += b.calculate(chunk)
        assert ra == rb, (chunk, ra, rb)
    assert ra >= 5, ra
    qa, pa, ta = a.particles()
    qb, pb, tb = b.particles()
    assert np.array_equal(ta, tb)
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    assert a.time() == b.time()
    say("constructions", ra)
    a.close()
    b.close()


def test_switch_to_cell_frame_codes_the_built_list():
    """Only the cell-frame mixed kick reads the list's cell-offset codes, so a construction in
    another frame (or in FP64) writes none.  Switching a built list to the cell frame writes
    them: its kicks and steps are then bit-identical to those of a list built in that frame."""
    say("case: switch to the cell frame after construction")
    L = 16.0
    xx, pp, ids = melted_state(L, 0.5, steps=100, seed=6)
    a = mdgpu.Domain(L, dt=0.002, precision=mdgpu.MIXED)
    a.set_option(mdgpu.OPT_MIXED_FRAME, 0)
    b = mdgpu.Domain(L, dt=0.002, precision=mdgpu.MIXED)  # brick frame
    c = mdgpu.Domain(L, dt=0.002)  # FP64
    for d in (a, b, c):
        d.add_particles(xx, pp, ids)
        d.make_pair_list()
    b.set_option(mdgpu.OPT_MIXED_FRAME, 0)
    c.set_precision(mdgpu.MIXED)
    c.set_option(mdgpu.OPT_MIXED_FRAME, 0)
    ka = a.kick()
    assert np.array_equal(ka, b.kick()) and np.array_equal(ka, c.kick())
    for d in (a, b, c):
        d.calculate(30)
    qa, pa, _ = a.particles()
    for d in (b, c):
        q, p, _ = d.particles()
        assert np.array_equal(qa, q) and np.array_equal(pa, p)
    for d in (a, b, c):
        d.close()


@pytest.mark.parametrize("thermostat,precision", [("nve", mdgpu.FP64), ("langevin", mdgpu.FP64),
                                                  ("langevin", mdgpu.MIXED)])
def test_runs_launched_directly_match_graph_runs(thermostat, precision):
    """A run's steps launched one kernel at a time (MDGPU_OPT_GRAPH 0) give the same bits, time
    and constructions as the same run replayed from captured graphs."""
    say("case: direct runs", thermostat, precision)
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0 * 1.5), 4)
    ds = []
    for graph in (1, 0):
        d = mdgpu.Domain(L, dt=0.005, precision=precision, capacity=n)
        d.add_particles(x, p)
        if thermostat == "langevin":
            d.set_thermostat(mdgpu.LANGEVIN, 1.2, 1.0, 0.1)
            d.set_seed(5)
        d.set_option(mdgpu.OPT_GRAPH, graph)
        ds.append(d)
    a, b = ds
    assert a.get_option(mdgpu.OPT_GRAPH) == 1 and b.get_option(mdgpu.OPT_GRAPH) == 0
    ra = rb = 0
    for chunk in (1, 64, 3, 90):
        ra += a.calculate(chunk)
        rb += b.calculate(chunk)
        assert ra == rb, (chunk, ra, rb)
    assert ra >= 4, ra
    qa, pa, ta = a.particles()
    qb, pb, tb = b.particles()
    assert np.array_equal(ta, tb)
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    assert a.time() == b.time()
    a.close()
    b.close()


@pytest.mark.parametrize("thermostat", ["nve", "langevin"])
def test_timed_steps_match_runs(thermostat):
    """Timed steps (HIP events around each kick, launched one at a time, final positions written
    by a kernel of their own after the kick as a run does) give the same bits, constructions and
    time as steps in runs, and the events time every kick of the call."""
    say("case: timed steps", thermostat)
    L = 16.0
    x = fcc(L, 0.5)
    p = initial_velocity(len(x), math.sqrt(3.0 * 1.5), 5)
    ds = []
    for timed in (True, False):
        d = mdgpu.Domain(L, dt=0.005, capacity=len(x))
        d.add_particles(x, p)
        if thermostat == "langevin":
            d.set_thermostat(mdgpu.LANGEVIN, 1.2, 1.0, 0.1)
            d.set_seed(7)
        d.enable_timing(timed)
        ds.append(d)
    a, b = ds
    ra = rb = 0
    for chunk in (1, 50, 99):
        ra += a.calculate(chunk)
        rb += b.calculate(chunk)
        assert ra == rb, (chunk, ra, rb)
        qa, pa, ta = a.particles()
        qb, pb, tb = b.particles()
        assert np.array_equal(ta, tb)
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    assert a.time() == b.time()
    tm = a.timing()  # of the last call
    assert tm["force_launches"] == 99 and tm["force_ms"] > 0.0, tm
    a.close()
    b.close()


def test_runs_are_bitwise_deterministic():
    say("case: determinism")
    L = 20.0
    xx, pp, ids = melted_state(L, 0.6, steps=100, seed=2)
    outs = []
    for _ in range(2):
        d = mdgpu.Domain(L, dt=0.002)
        d.add_particles(xx, pp, ids)
        d.make_pair_list()
        d.calculate(150)
        outs.append(d.particles_by_id())
        d.close()
    assert all(np.array_equal(a, b) for a, b in zip(outs[0], outs[1]))


@pytest.mark.parametrize("frame,bound", [(0, 2e-6), (1, 4e-6)])
def test_mixed_precision_force_error_is_bounded(frame, bound):
    """Mixed-precision kicks against FP64 on a fluid: RMS relative error below 2e-6 in the
    cell-relative frame and below 4e-6 in the brick-centred frame, whose coordinates are up to
    about half a neighbourhood wide (rounding 2^-21..2^-20 instead of 2^-23..2^-22)."""
    say("case: mixed precision force error, frame", frame)
    L = 20.0
    xx, pp, ids = melted_state(L, 0.5, steps=200, seed=4)
    d64 = mdgpu.Domain(L, dt=0.001)
    d32 = mdgpu.Domain(L, dt=0.001, precision=mdgpu.MIXED)
    d32.set_option(mdgpu.OPT_MIXED_FRAME, frame)
    for d in (d64, d32):
        d.add_particles(xx, pp, ids)
        d.make_pair_list()
    k64 = d64.kick()
    k32 = d32.kick()
    rms = math.sqrt(float(np.sum((k32 - k64) ** 2) / np.sum(k64 ** 2)))
    say("mixed force RMS relative error", rms)
    assert rms < bound
    assert np.array_equal(O.canonical_pairs(d64.pairs()), O.canonical_pairs(d32.pairs()))


def test_change_scale_against_oracle():
    say("case: ChangeScale")
    L = 16.0
    xx, pp, ids = melted_state(L, 0.75, steps=50)
    d, o = make_pair(L, xx, pp, ids, rc=2.5, margin=0.5)
    d.make_pair_list()
    o.make_pair_list()
    alpha = (0.75 / 0.6) ** (1.0 / 3.0)
    d.change_scale(alpha)
    o.change_scale(alpha)
    compare_state(d, o, L * alpha, 2.5, 0.001, "change_scale")


def test_change_scale_picks_bricks_for_the_new_density():
    """Bricks a domain picked from its density are picked again after ChangeScale, as a fresh
    domain at the new density picks them; bricks the caller set are kept."""
    say("case: ChangeScale bricks")
    L = 40.0
    x = fcc(L, 0.75)
    alpha = (0.75 / 0.45) ** (1.0 / 3.0)
    d = mdgpu.Domain(L, dt=0.001, capacity=len(x))
    d.add_particles(x)
    d.make_pair_list()
    dense = d.get_option(mdgpu.OPT_BRICK)
    d.change_scale(alpha)
    f = mdgpu.Domain(L * alpha, dt=0.001, capacity=len(x))
    f.add_particles(x * alpha)
    f.make_pair_list()
    say("bricks", hex(dense), hex(d.get_option(mdgpu.OPT_BRICK)), hex(f.get_option(mdgpu.OPT_BRICK)))
    assert d.get_option(mdgpu.OPT_BRICK) == f.get_option(mdgpu.OPT_BRICK) != dense
    u = mdgpu.Domain(L, dt=0.001, capacity=len(x))
    u.add_particles(x)
    u.set_option(mdgpu.OPT_BRICK, dense)
    u.make_pair_list()
    u.change_scale(alpha)
    assert u.get_option(mdgpu.OPT_BRICK) == dense
    for g in (d, f, u):
        g.close()


def test_error_paths_raise():
    say("case: error paths")
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    d = mdgpu.Domain(L, dt=0.001)
    p = np.zeros((n, 3))
    p[5] = [400.0, 0.0, 0.0]  # 2 * |p| * dt > margin
    d.add_particles(x, p)
    d.make_pair_list()
    with pytest.raises(mdgpu.MDGPUError) as e:
        d.calculate(1)
    assert e.value.code == mdgpu.ERR_TOO_FAST
    d2 = mdgpu.Domain(L, dt=0.001)
    x2 = x.copy()
    d2.add_particles(x2)
    d2.make_pair_list()
    q, pp, t = d2.particles()
    pp[0] = [1.0e5, 0, 0]  # drifts to ~ 3L in one half step: wrap cannot bring it back
    d2.set_momenta(pp)
    d2.update_position_half()
    with pytest.raises(mdgpu.MDGPUError) as e:
        d2.make_pair_list()
    assert e.value.code == mdgpu.ERR_INVALID_POSITION



# smoothed LJ constants at rc = 2.5 from C2 = 6 s^14 - 3 s^8, C0 = -s^12 + s^6 - C2 rc^2, s = 1/rc
# (variables.cc:16-28), evaluated in exact arithmetic
C2_25, C0_25 = -0.0019499738726400002, 0.016266559488


def test_empty_domain_and_add_filter():
    """No particles: a build, steps and observers run and give zeros.  Particles outside
    [0, L) are not added (MDRect::IsInside), as in the oracle."""
    say("case: empty domain, add filter")
    L = 16.0
    d = mdgpu.Domain(L, dt=0.001)
    d.make_pair_list()
    assert d.calculate(5) >= 0
    assert d.counts() == (0, 0)
    ob = d.observables()
    assert ob["ke"] == 0.0 and ob["pe"] == 0.0 and ob["vir"] == 0.0
    assert np.isnan(ob["T"]) and np.isnan(ob["E"])  # 0 / 0, as the reference's double division gives
    q, p, t = d.particles()
    assert q.shape == (0, 3) and p.shape == (0, 3) and t.shape == (0,)
    x = np.array([[0.0, 1.0, 1.0], [L, 1.0, 1.0], [np.nextafter(L, 0.0), 2.0, 2.0], [-1e-12, 3.0, 3.0],
                  [5.0, 5.0, L], [8.0, 8.0, 8.0]])
    d.add_particles(x, None, np.arange(len(x), dtype=np.int32))
    o = O.OracleSystem(L, 0.001, 2.5, 0.5)
    o.add(x, None, np.arange(len(x), dtype=np.int32))
    pn, _ = d.counts()
    assert pn == 3 == len(o.state()[2])
    assert np.array_equal(np.sort(d.particles()[2]), np.array([0, 2, 5]))
    d.close()


@pytest.mark.parametrize("across", [False, True])
def test_dimer_closed_form(across):
    """Two particles 1.0 apart, inside the box or through the periodic boundary: kick
    (24 - 8 C2) dt along the separation, PE 4 (C2 + C0), virial (-24 + 8 C2) / 3; an NVE run
    follows the oracle, energy included (the pair flies apart from r = 1, where the repulsion is
    steep: position Verlet at dt = 0.001 ends 1.7e-3 above the initial energy, in both).  Inside the box the two rows see exactly
    opposite separations, so the kicks are exact negatives and the total momentum stays exactly
    zero; through the boundary each row pairs with the other particle's image, whose position is
    rounded differently (15.3 - 16 against 0.3 + 16), and the kicks are opposite to rounding."""
    say("case: dimer", "across the boundary" if across else "inside the box")
    L, dt = 16.0, 0.001
    x = np.array([[0.3, 8.0, 8.0], [L - 0.7, 8.0, 8.0]]) if across else np.array([[8.0, 8.0, 8.0], [9.0, 8.0, 8.0]])
    sign = 1.0 if across else -1.0  # particle 0's partner lies at -x (its image) or at +x
    d = mdgpu.Domain(L, dt=dt)
    d.add_particles(x, None, np.array([0, 1], dtype=np.int32))
    d.make_pair_list()
    q, p, t = d.particles()
    k = d.kick()[np.argsort(t)]
    want = sign * (24.0 - 8.0 * C2_25) * dt
    assert k[0] == pytest.approx([want, 0.0, 0.0], rel=1e-13, abs=1e-300)
    if across:
        assert k[1] == pytest.approx(-k[0], rel=1e-13, abs=1e-300)
    else:
        assert np.array_equal(k[1], -k[0])
    ob = d.observables()
    assert ob["pe"] == pytest.approx(4.0 * (C2_25 + C0_25), rel=1e-12)
    assert ob["vir"] == pytest.approx((-24.0 + 8.0 * C2_25) / 3.0, rel=1e-12)
    d.calculate(2000)
    q, p, t = d.particles()
    if across:
        assert np.abs(p.sum(axis=0)).max() <= 1e-12 * np.abs(p).max()
    else:
        assert np.array_equal(p.sum(axis=0), np.zeros(3))
    ob = d.observables()
    o = O.OracleSystem(L, dt, 2.5, 0.5)
    o.add(x, None, np.array([0, 1], dtype=np.int32))
    o.make_pair_list()
    for _ in range(2000):
        o.step()
    oq, op, ot = o.state()
    ia, ib = np.argsort(t), np.argsort(ot)
    assert np.abs(q[ia] - oq[ib]).max() < 1e-10 and np.abs(p[ia] - op[ib]).max() < 1e-10
    assert ob["E"] == pytest.approx(o.observables()["E"], rel=1e-10)
    d.close()

def test_full_size_4M_pairs_exact_and_closed_form():
    """4,000,000 particles, rho 0.5, L 200, rc 2.5, margin 0.5 (perfect FCC):
    27 half pairs per particle, bit-exact with the oracle; again after 1000 steps."""
    say("case: 4M full size")
    L = 200.0
    x = fcc(L, 0.5)
    n = len(x)
    assert n == 4_000_000
    p = initial_velocity(n, math.sqrt(3.0), 0)
    d = mdgpu.Domain(L, dt=0.001, capacity=n)
    d.add_particles(x, p)
    d.make_pair_list()
    gp = d.pairs()
    own = gp[:, 2] == 13
    assert int(own.sum()) + int((~own).sum()) // 2 == 27 * n
    ob = d.observables()
    assert ob["pe"] / n == pytest.approx(-2.413162698239619, rel=1e-12)
    o = O.OracleSystem(L, 0.001, 2.5, 0.5)
    o.add(x, p)
    o.make_pair_list()
    assert np.array_equal(O.canonical_pairs(gp), O.canonical_pairs(o.pairs()))
    del o
    say("4M step-0 pairs exact; running 1000 steps")
    nreb = d.calculate(1000)
    say("4M rebuilds in 1000 steps:", nreb)
    q, pp, t = d.particles()
    o = O.OracleSystem(L, 0.001, 2.5, 0.5)
    o.add(wrap_like_reference(q, L), pp, t)
    d.make_pair_list()
    o.make_pair_list()
    assert np.array_equal(O.canonical_pairs(d.pairs()), O.canonical_pairs(o.pairs()))


@pytest.mark.parametrize("delta", [50, 150])
def test_pruned_inner_list_is_bitwise_neutral(delta):
    """Dynamic pruning (MDGPU_OPT_PRUNE): the kick reads only pairs within rc + delta, pruned whenever
    the particles could have moved delta; the same pairs are summed in the same order, so FP64 runs
    (and mixed precision with plain FP32 row sums) match the full list bit for bit, across rebuilds."""
    say("case: pruned inner list, delta", delta / 1000)
    L = 16.0
    xx, pp, ids = melted_state(L, 0.5)
    for prec, ksum in ((mdgpu.FP64, None), (mdgpu.MIXED, 0)):
        a = mdgpu.Domain(L, dt=0.002, precision=prec)
        b = mdgpu.Domain(L, dt=0.002, precision=prec)
        for d in (a, b):
            if ksum is not None:
                d.set_option(mdgpu.OPT_KAHAN, ksum)
            d.add_particles(xx, pp, ids)
            d.make_pair_list()
        b.set_option(mdgpu.OPT_PRUNE, delta)
        assert b.get_option(mdgpu.OPT_PRUNE) == delta
        ra = sum(a.calculate(50) for _ in range(6))
        rb = sum(b.calculate(50) for _ in range(6))
        assert ra == rb and ra >= 2, (ra, rb)
        qa, pa, ta = a.particles()
        qb, pb, tb = b.particles()
        assert np.array_equal(ta, tb)
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb), prec
        # mixed precision with FP64 group sums regroups the terms: equal to rounding
        if prec == mdgpu.MIXED:
            for d in (a, b):
                d.set_option(mdgpu.OPT_KAHAN, 2)
            a.calculate(40)
            b.calculate(40)
            qa, pa, _ = a.particles_by_id()
            qb, pb, _ = b.particles_by_id()
            assert np.abs(qa - qb).max() < 1e-6 and np.abs(pa - pb).max() < 1e-5


def test_pruned_inner_list_langevin_and_nose_hoover():
    """Langevin (one half drift per step) prunes on the same budget; Nose-Hoover steps the full list."""
    say("case: pruning with thermostats")
    L = 16.0
    xx, pp, ids = melted_state(L, 0.5)
    for kind in (mdgpu.LANGEVIN, mdgpu.NOSEHOOVER):
        a = mdgpu.Domain(L, dt=0.004)
        b = mdgpu.Domain(L, dt=0.004)
        for d in (a, b):
            d.add_particles(xx, pp, ids)
            d.set_thermostat(kind, 1.2, 0.5, 0.2)
            d.set_seed(3)
            d.make_pair_list()
        b.set_option(mdgpu.OPT_PRUNE, 100)
        assert a.calculate(150) == b.calculate(150)
        qa, pa, ta = a.particles()
        qb, pb, tb = b.particles()
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb), kind


@pytest.mark.parametrize("lds_kb", [0, 24])
def test_pipelined_kick_is_bitwise_identical(lds_kb):
    """The pipelined FP64 kick (MDGPU_OPT_PIPELINE: persistent workgroups whose producer
    wavefronts stage the next brick with LDS-direct loads) sums every row's partners in the
    same order as the one-workgroup-per-brick kick: NVE and Langevin runs with rebuilds give
    the same bits, also where bricks overflow the LDS buffers (a 24 KiB budget: partners read
    from global memory) and with the pruned inner list."""
    say("case: pipelined kick, lds_kb", lds_kb)
    L = 30.0
    xx, pp, ids = melted_state(L, 0.55, steps=150, seed=8)
    for kind, prune in ((mdgpu.NVE, 0), (mdgpu.LANGEVIN, 0), (mdgpu.NVE, 150)):
        a = mdgpu.Domain(L, dt=0.002)
        b = mdgpu.Domain(L, dt=0.002)
        for d in (a, b):
            d.add_particles(xx, pp, ids)
            d.set_thermostat(kind, 1.1, 0.5, 0.1)
            d.set_seed(5)
            if lds_kb:
                d.set_option(mdgpu.OPT_LDS_KB, lds_kb)
            if prune:
                d.set_option(mdgpu.OPT_PRUNE, prune)
            d.make_pair_list()
        b.set_option(mdgpu.OPT_PIPELINE, 1)
        assert b.get_option(mdgpu.OPT_PIPELINE) == 1
        ra = sum(a.calculate(60) for _ in range(5))
        rb = sum(b.calculate(60) for _ in range(5))
        assert ra == rb and ra >= 2, (ra, rb)
        qa, pa, _ = a.particles_by_id()
        qb, pb, _ = b.particles_by_id()
        assert np.array_equal(qa, qb) and np.array_equal(pa, pb), (kind, prune)
        assert np.array_equal(a.kick(), b.kick())
        a.close()
        b.close()


def test_rebuild_cadence_without_displacement_check():
    """MDGPU_OPT_DISPLACEMENT_CHECK 0: the list expires as soon as the Verlet budget is spent,
    without the exact displacement check.  That rebuilds at least as often and lists the same
    pairs within rc, so the trajectory differs only by the summation order of the kicks, and
    a list built at any step is the oracle's."""
    say("case: rebuild cadence without displacement check")
    L = 20.0
    xx, pp, ids = melted_state(L, 0.5, steps=200, seed=6)
    a = mdgpu.Domain(L, dt=0.002)
    b = mdgpu.Domain(L, dt=0.002)
    for d in (a, b):
        d.add_particles(xx, pp, ids)
        d.make_pair_list()
    b.set_option(mdgpu.OPT_DISPLACEMENT_CHECK, 0)
    assert b.get_option(mdgpu.OPT_DISPLACEMENT_CHECK) == 0
    ra = rb = 0
    for _ in range(4):
        ra += a.calculate(50)
        rb += b.calculate(50)
    assert rb > ra >= 1, (ra, rb)
    qa, pa, _ = a.particles_by_id()
    qb, pb, _ = b.particles_by_id()
    # positions are wrapped at constructions, which fall on different steps here
    dq = np.abs(np.remainder(qa - qb + 0.5 * L, L) - 0.5 * L)
    assert dq.max() < 1e-9 and np.abs(pa - pb).max() < 1e-8
    q, p, t = b.particles()
    o = O.OracleSystem(L, 0.002, 2.5, 0.5)
    o.add(wrap_like_reference(q, L), p, t)
    b.make_pair_list()
    o.make_pair_list()
    assert np.array_equal(O.canonical_pairs(b.pairs()), O.canonical_pairs(o.pairs()))
    say("constructions with the check", ra, "without", rb)
