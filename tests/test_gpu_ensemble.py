"""GPU parity of the ensemble path (mdgpu_create_ensemble): independent replicas
advanced by one pair-force launch per step (launched, or replayed from a captured graph).

Bars:
- each replica's pair set bit-exact with the oracle's for that replica alone;
- each replica's kick bit-identical to a single domain's over the same input
  with the same brick shape (same list, same summation order), and within
  1e-12 of the oracle per particle (relative to the sum of pair-kick
  magnitudes);
- trajectories bit-identical to single domains that rebuild their lists at
  the same steps (a construction orders particles by cell and id only), and
  graph replay bit-identical to launched steps;
- lists stay valid along a run with rebuilds (kick with the running list ==
  kick with a fresh list, to 1e-12);
- per-replica observers within 1e-12 of single domains; Langevin replicas
  thermalise to their own temperatures.
"""
import math
import sys

import numpy as np
import pytest

from helpers import lj_C, pair_force_sums
from mdacp_amd import mdgpu
from mdacp_amd.ensemble import box_for
from mdacp_amd.lattice import fcc, initial_velocity
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RC, MARGIN = 2.5, 0.5

# (k, rho, T, seed): uneven sizes and densities, boxes from 6.6 to 16.6
SPEC = [(4, 0.9, 0.8, 11), (5, 0.3, 1.4, 12), (6, 0.6, 1.0, 13), (4, 0.75, 1.2, 14), (5, 0.5, 0.7, 15),
        (7, 0.4, 1.1, 16)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replicas(spec, melt_seed=None):
    """boxes, counts, q, p, ids for the rows of spec; melted (random displacements) when melt_seed is set."""
    boxes, counts, qs, ps, ids = [], [], [], [], []
    rng = np.random.default_rng(melt_seed) if melt_seed is not None else None
    for k, rho, T, seed in spec:
        L = box_for(k, rho)
        x = fcc(L, rho)
        n = len(x)
        if rng is not None:
            x = np.mod(x + rng.uniform(-0.15, 0.15, x.shape), L)
            x = np.where(x >= L, 0.0, x)
        boxes.append((L, L, L))
        counts.append(n)
        qs.append(x)
        ps.append(initial_velocity(n, math.sqrt(3.0 * T), seed))
        ids.append(rng.permutation(n).astype(np.int32) if rng is not None else np.arange(n, dtype=np.int32))
    return np.array(boxes), np.array(counts, dtype=np.int32), np.concatenate(qs), np.concatenate(ps), \
        np.concatenate(ids)


def split(a, counts):
    return np.split(a, np.cumsum(counts)[:-1])


def ensemble(spec, melt_seed=None, dt=0.005, thermostat=mdgpu.LANGEVIN, precision=mdgpu.FP64, brick=None,
             gamma=1.0):
    boxes, counts, q, p, ids = replicas(spec, melt_seed)
    e = mdgpu.Ensemble(boxes, counts, q, p, ids, T=[s[2] for s in spec], gamma=gamma,
                       seed=[s[3] for s in spec], thermostat=thermostat, dt=dt, rc=RC, margin=MARGIN,
                       precision=precision)
    if brick is not None:
        e.set_option(mdgpu.OPT_BRICK, brick | (brick << 8) | (brick << 16))
        e.make_pair_list()
    return e, (boxes, counts, q, p, ids)


def single(L, q, p, ids, T, seed, dt=0.005, thermostat=mdgpu.LANGEVIN, precision=mdgpu.FP64, brick=None,
           gamma=1.0):
    d = mdgpu.Domain(L, dt=dt, rc=RC, margin=MARGIN, precision=precision, capacity=len(q))
    if brick is not None:
        d.set_option(mdgpu.OPT_BRICK, brick | (brick << 8) | (brick << 16))
    d.add_particles(q, p, ids)
    d.set_thermostat(thermostat, T, gamma, 0.1)
    d.set_seed(seed)
    d.make_pair_list()
    return d


def by_id(q, p, t):
    o = np.argsort(t, kind="stable")
    return q[o], p[o], t[o]


@pytest.mark.parametrize("brick", [4, 2])
def test_replica_lists_and_kicks(brick):
    say(f"ensemble lists and kicks, brick {brick}")
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=1, brick=brick)
    assert list(e.replica_counts()) == list(counts)
    kick = e.kick()
    eq, ep, et = e.particles()
    sl = e.replica_slices()
    C0, C2 = lj_C(RC)
    for r, (qr, pr, ir) in enumerate(zip(split(q, counts), split(p, counts), split(ids, counts))):
        L = boxes[r, 0]
        o = O.OracleSystem(L, 0.005, RC, MARGIN)
        o.add(qr, pr, ir)
        o.make_pair_list()
        gp = e.replica_pairs(r)
        assert np.array_equal(O.canonical_pairs(gp), O.canonical_pairs(o.pairs())), f"replica {r} pairs"
        d = single(L, qr, pr, ir, SPEC[r][2], SPEC[r][3], brick=brick)
        dq, dp, dt_ = d.particles()
        # same list in the same order: the replica's kick is bit-identical to the single domain's
        assert np.array_equal(et[sl[r]], dt_), f"replica {r} particle order"
        assert np.array_equal(kick[sl[r]], d.kick()), f"replica {r} kick"
        # and within 1e-12 of the oracle
        oq, _, ot = o.state()
        ib = np.argsort(ot, kind="stable")
        ko = o.kick()[ib]
        S = pair_force_sums(oq[ib], o.pairs().astype(np.int64), L, RC, 0.005, C2)
        k_sorted = kick[sl[r]][np.argsort(et[sl[r]], kind="stable")]
        err = np.abs(k_sorted - ko).max(axis=1)
        assert np.all(err <= 1e-12 * np.maximum(S, 1e-300) + 1e-300), f"replica {r} kick vs oracle"
        d.close()


def test_auto_bricks_same_pairs():
    say("ensemble with per-replica brick shapes")
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=2)
    for r, (qr, pr, ir) in enumerate(zip(split(q, counts), split(p, counts), split(ids, counts))):
        o = O.OracleSystem(boxes[r, 0], 0.005, RC, MARGIN)
        o.add(qr, pr, ir)
        o.make_pair_list()
        assert np.array_equal(O.canonical_pairs(e.replica_pairs(r)), O.canonical_pairs(o.pairs())), r


def test_trajectory_bitwise_without_rebuild():
    """dt small enough that no list expires: every replica follows its single-domain twin bit for bit."""
    say("ensemble vs single domains, 30 Langevin steps without rebuild")
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=3, dt=0.001, brick=4)
    nreb = e.calculate(30)
    assert nreb == 0
    eq, ep, et = e.particles()
    sl = e.replica_slices()
    for r, (qr, pr, ir) in enumerate(zip(split(q, counts), split(p, counts), split(ids, counts))):
        d = single(boxes[r, 0], qr, pr, ir, SPEC[r][2], SPEC[r][3], dt=0.001, brick=4)
        assert d.calculate(30) == 0
        dq, dp, dt_ = d.particles()
        assert np.array_equal(et[sl[r]], dt_), r
        assert np.array_equal(eq[sl[r]], dq), f"replica {r} positions"
        assert np.array_equal(ep[sl[r]], dp), f"replica {r} momenta"
        d.close()


@pytest.mark.parametrize("precision", [mdgpu.FP64, mdgpu.MIXED])
def test_trajectory_bitwise_with_mirrored_rebuilds(precision):
    """Langevin run with rebuilds; each single domain rebuilds where the ensemble did and
    follows its replica bit for bit."""
    say(f"ensemble vs single domains with mirrored rebuilds, precision {precision}")
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=7, dt=0.005, precision=precision, brick=4)
    ds = [single(boxes[r, 0], qr, pr, ir, SPEC[r][2], SPEC[r][3], dt=0.005, precision=precision, brick=4)
          for r, (qr, pr, ir) in enumerate(zip(split(q, counts), split(p, counts), split(ids, counts)))]
    nreb = 0
    for step in range(70):
        rebuilt = e.calculate(1)
        nreb += rebuilt
        for d in ds:
            if rebuilt:
                d.make_pair_list()
            assert d.calculate(1) == 0, step  # never expires before the ensemble rebuilds
    assert nreb >= 2
    eq, ep, et = e.particles()
    sl = e.replica_slices()
    for r, d in enumerate(ds):
        dq, dp, dt_ = d.particles()
        assert np.array_equal(et[sl[r]], dt_), r
        assert np.array_equal(eq[sl[r]], dq), f"replica {r} positions"
        assert np.array_equal(ep[sl[r]], dp), f"replica {r} momenta"
        assert d.time()[1] == e.time()[1]
        d.close()


@pytest.mark.parametrize("precision", [mdgpu.FP64, mdgpu.MIXED])
def test_graph_replay_matches_launches(precision):
    """The captured-and-replayed step gives the same bits as launching its kernels, across rebuilds."""
    say(f"graph replay vs launches, precision {precision}")
    a, _ = ensemble(SPEC, melt_seed=4, precision=precision)
    b, _ = ensemble(SPEC, melt_seed=4, precision=precision)
    a.set_option(mdgpu.OPT_GRAPH, 1)
    assert a.get_option(mdgpu.OPT_GRAPH) == 1 and b.get_option(mdgpu.OPT_GRAPH) == 0
    ra = sum(a.calculate(10) for _ in range(8))
    rb = sum(b.calculate(10) for _ in range(8))
    assert ra == rb and ra >= 2, (ra, rb)
    qa, pa, ta = a.particles()
    qb, pb, tb = b.particles()
    assert np.array_equal(ta, tb)
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)
    assert a.time() == b.time()


def unordered(pairs3):
    a = np.minimum(pairs3[:, 0], pairs3[:, 1]).astype(np.int64)
    b = np.maximum(pairs3[:, 0], pairs3[:, 1]).astype(np.int64)
    return np.unique((a << 32) | b)


def test_lists_stay_valid_across_rebuilds():
    """Along a run with rebuilds (the replicas' own budgets and displacement checks), every pair
    inside rc is in each replica's running list, and the running kick equals that of a list built
    afresh from the same state (a second ensemble)."""
    say("list validity along an ensemble run")
    e, (boxes, counts, _, _, _) = ensemble(SPEC, melt_seed=5, dt=0.005)
    total = 0
    for chunk in range(10):
        total += e.calculate(7)
        e.send_border_particles()  # ghosts at the current positions, as a fresh build has them
        k_run = e.kick()
        q, p, t = e.particles()
        run = [unordered(e.replica_pairs(r)) for r in range(e.R)]
        rep_of = np.repeat(np.arange(e.R), e.replica_counts())
        Lp = boxes[rep_of]
        qw = np.where(q < 0.0, q + Lp, np.where(q >= Lp, q - Lp, q))
        f = mdgpu.Ensemble(boxes, counts, qw, p, t, T=[x[2] for x in SPEC], gamma=1.0,
                           seed=[x[3] for x in SPEC], thermostat=mdgpu.LANGEVIN, dt=0.005, rc=RC, margin=MARGIN)
        fq, fp, ft = f.particles()
        sl = f.replica_slices()
        for r in range(e.R):
            fresh = f.replica_pairs(r)
            qr = fq[sl[r]][np.argsort(ft[sl[r]], kind="stable")]
            code = fresh[:, 2]
            sh = np.stack([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1], axis=-1) * boxes[r, 0]
            dvec = qr[fresh[:, 1]] + sh - qr[fresh[:, 0]]
            inside = fresh[np.einsum("ij,ij->i", dvec, dvec) <= RC * RC]
            assert np.all(np.isin(unordered(inside), run[r])), (chunk, r)
        k_new = f.kick()
        ka = k_run[np.lexsort((t, rep_of))]
        kb = k_new[np.lexsort((ft, rep_of))]
        assert np.abs(ka - kb).max() <= 1e-11 * max(np.abs(kb).max(), 1.0), chunk
        f.close()
    assert total >= 2


def test_observers_per_replica():
    say("per-replica observers vs single domains")
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=6, thermostat=mdgpu.NVE)
    e.calculate(20)
    ke, pe, vir = e.observe_replicas_raw()
    eq, ep, et = e.particles()
    sl = e.replica_slices()
    for r in range(e.R):
        # kinetic energy of replica r's current momenta, summed on the host
        pr = ep[sl[r]]
        k1 = float(np.sum(0.5 * pr * pr))
        assert abs(ke[r] - k1) <= 1e-12 * abs(k1), r
    # PE / virial on a freshly built state, against the oracle
    e.make_pair_list()
    ke, pe, vir = e.observe_replicas_raw()
    eq, ep, et = e.particles()
    for r in range(e.R):
        o = O.OracleSystem(boxes[r, 0], 0.005, RC, MARGIN)
        o.add(eq[sl[r]], ep[sl[r]], et[sl[r]])
        o.make_pair_list()
        ob = o.observables()
        ko, po, vo = ob["ke"], ob["pe"], ob["vir"]
        assert abs(ke[r] - ko) <= 1e-12 * abs(ko), r
        assert abs(pe[r] - po) <= 1e-12 * abs(po), r
        assert abs(vir[r] - vo) <= 1e-12 * abs(vo), r


def test_langevin_replicas_follow_their_temperatures():
    """Replicas of one density relax towards their own bath temperatures (gamma = 1, dt = 0.005):
    the kinetic temperatures keep the order of the targets and stay near them."""
    say("Langevin replicas at different temperatures")
    spec = [(5, 0.7, 0.8, 21), (5, 0.7, 1.1, 22), (5, 0.7, 1.4, 23), (5, 0.7, 1.7, 24)]
    e, _ = ensemble(spec, dt=0.005)
    e.calculate(400)
    Ts = []
    for _ in range(40):
        e.calculate(10)
        Ts.append(e.observables()["T"])
    Tm = np.mean(Ts, axis=0)
    target = np.array([s[2] for s in spec])
    say("mean T", Tm, "targets", target)
    assert np.all(np.diff(Tm) > 0), Tm
    assert np.all(np.abs(Tm / target - 1.0) < 0.25), (Tm, target)


def test_autotune_keeps_results():
    """Besides the list layout and the rebuild cadence, which decide when the list is rebuilt,
    the autotuner changes launch shapes only: an ensemble with the tuned launch shapes follows
    one with the default shapes bit for bit when both use the tuned layout and cadence."""
    say("autotuned ensemble vs default launch shapes")
    from mdacp_amd.autotune import autotune, CADENCE, FORCE, LAYOUT, SEARCH

    a, _ = ensemble(SPEC, melt_seed=8)
    b, _ = ensemble(SPEC, melt_seed=8)
    res = autotune(a, steps=5)
    ch = res["choice"]
    assert (ch["search_waves"], ch["search_stage"]) in SEARCH
    assert (ch["waves"], ch["lds_kb"], ch["unroll"]) in FORCE
    assert ch["row_sort"] in LAYOUT and ch["displacement_check"] in CADENCE
    assert a.get_option(mdgpu.OPT_WAVES) == ch["waves"] and a.get_option(mdgpu.OPT_UNROLL) == ch["unroll"]
    assert a.get_option(mdgpu.OPT_DISPLACEMENT_CHECK) == ch["displacement_check"]
    say("cadence timings", res["cadence"], "layout timings", res["layout"])
    if b.get_option(mdgpu.OPT_ROW_SORT) != ch["row_sort"]:
        b.set_option(mdgpu.OPT_ROW_SORT, ch["row_sort"])
    b.set_option(mdgpu.OPT_DISPLACEMENT_CHECK, ch["displacement_check"])
    ra, rb = a.calculate(60), b.calculate(60)
    assert ra == rb
    qa, pa, ta = a.particles()
    qb, pb, tb = b.particles()
    assert np.array_equal(ta, tb) and np.array_equal(qa, qb) and np.array_equal(pa, pb)


def test_pruned_ensemble_is_bitwise_neutral():
    """Dynamic pruning in an ensemble (budget from the largest |p| of all replicas), replayed from graphs."""
    say("pruned ensemble vs full list")
    a, _ = ensemble(SPEC, melt_seed=9)
    b, _ = ensemble(SPEC, melt_seed=9)
    b.set_option(mdgpu.OPT_PRUNE, 200)
    ra = sum(a.calculate(20) for _ in range(4))
    rb = sum(b.calculate(20) for _ in range(4))
    assert ra == rb and ra >= 2
    qa, pa, ta = a.particles()
    qb, pb, tb = b.particles()
    assert np.array_equal(ta, tb) and np.array_equal(qa, qb) and np.array_equal(pa, pb)


def test_nose_hoover_replicas_follow_single_domains():
    """Nose-Hoover in every replica (MDManager::CalculateNoseHoover per system): each replica
    with its own temperature, target and zeta follows a single domain run with the same
    thermostat for 300 steps (1e-9 in positions and momenta: the two rebuild at different
    steps and sum the kinetic energy in another order), its zeta to 1e-9 relative."""
    spec = SPEC[:4]
    e, (boxes, counts, q, p, ids) = ensemble(spec, melt_seed=3, dt=0.002, thermostat=mdgpu.NOSEHOOVER)
    e.set_thermostat(mdgpu.NOSEHOOVER, T=1.0, tau=0.2)
    for r, (k, rho, T, seed) in enumerate(spec):
        e.set_replica_thermostat(r, T, 1.0, seed)
    e.make_pair_list()
    e.calculate(300)
    zs = e.replica_zetas()
    qe, pe_, te = e.particles()
    for r, (k, rho, T, seed) in enumerate(spec):
        L = boxes[r][0]
        d = mdgpu.Domain(L, dt=0.002, rc=RC, margin=MARGIN, capacity=int(counts[r]))
        d.add_particles(split(q, counts)[r], split(p, counts)[r], split(ids, counts)[r])
        d.set_thermostat(mdgpu.NOSEHOOVER, T=T, tau=0.2)
        d.make_pair_list()
        d.calculate(300)
        q1, p1, t1 = d.particles_by_id()
        sl = e.replica_slices()[r]
        o = np.argsort(te[sl], kind="stable")
        q2, p2, t2 = qe[sl][o], pe_[sl][o], te[sl][o]
        assert np.array_equal(t1, t2)
        dq = np.remainder(q1 - q2 + L / 2, L) - L / 2
        say(f"replica {r}: T {T}, zeta {zs[r]:.6e} vs {d.zeta():.6e}, max |dq| {np.abs(dq).max():.2e}")
        assert np.abs(dq).max() < 1e-9 and np.abs(p1 - p2).max() < 1e-9
        assert zs[r] == pytest.approx(d.zeta(), rel=1e-9, abs=1e-12)
        d.close()
    e.close()


def test_change_scale_of_every_replica():
    """MDManager::ChangeScale(alpha) in every replica: positions become fl(q * alpha) of the
    wrapped positions, boxes alpha L, and each replica's pair set equals a single domain's
    after its own ChangeScale; the run goes on."""
    alpha = 1.05
    e, (boxes, counts, q, p, ids) = ensemble(SPEC, melt_seed=5)
    e.calculate(20)
    q0, _, t0 = e.particles()
    e.change_scale(alpha)
    q1, _, t1 = e.particles()
    for r, sl in enumerate(e.replica_slices()):
        L = boxes[r][0]
        a = {int(t): x for t, x in zip(t0[sl], q0[sl])}
        for t, x in zip(t1[sl], q1[sl]):
            w = np.where(a[int(t)] < 0.0, a[int(t)] + L, np.where(a[int(t)] > L, a[int(t)] - L, a[int(t)]))
            assert np.array_equal(x, w * alpha), (r, int(t))
        d = mdgpu.Domain(L, dt=0.005, rc=RC, margin=MARGIN, capacity=int(counts[r]))
        d.add_particles(np.mod(q0[sl], L), None, t0[sl])
        d.make_pair_list()
        d.change_scale(alpha)
        assert np.array_equal(O.canonical_pairs(e.replica_pairs(r)), O.canonical_pairs(d.pairs())), r
        d.close()
    e.calculate(50)
    assert np.all(np.isfinite(e.observables()["E"]))
    e.close()
