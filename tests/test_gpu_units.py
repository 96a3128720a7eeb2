"""Several units of a process grid in one process on one GPU (mdgpu_create_unit):
MDManager's particle migration and halo between its MDUnits (mdmanager.cc:194-348,
mdunit.cc:78-191) done device to device, checked against the restated relay of
oracle/units.py and against a single periodic domain holding the same particles
(itself checked against the oracle and the reference's fixtures by the other GPU
tests).

Bars: a grid of one unit places exactly the periodic domain's ghosts, in the same
order, so its kicks and observables are bit-identical; on larger grids every unit's
pair set equals the restated relay's (oracle/units.py, pinned in test_oracle_units.py)
exactly, with kicks within 1e-13 of each particle's sum of pair-kick magnitudes and
potential energy and virial within 1e-12; trajectories of 300 steps with migration
agree with the periodic domain's to 1e-9 in positions and momenta, every particle
staying inside its unit's box.
"""
import sys

import numpy as np
import pytest

from mdacp_amd import mdgpu
from mdacp_amd.lattice import fcc, initial_velocity
from oracle import oracle as O
from oracle import units as OU

pytestmark = pytest.mark.gpu

RC, MARGIN = 2.5, 0.5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def undirected(pairs3):
    """Pairs as a sorted set of undirected keys: (min id, max id, image of the larger
    id seen from the smaller one)."""
    p = np.asarray(pairs3, dtype=np.int64)
    if len(p) == 0:
        return np.zeros(0, dtype=np.int64)
    a, b, c = p[:, 0], p[:, 1], p[:, 2]
    flip = a > b
    lo, hi, code = np.where(flip, b, a), np.where(flip, a, b), np.where(flip, 26 - c, c)
    return np.unique((lo << 36) | (hi << 5) | code)


def make_units(L, grid, q, p, ids, dt=0.005):
    units = mdgpu.unit_grid(L, grid, dt=dt, rc=RC, margin=MARGIN, capacity=len(q))
    for u in units:
        u.add_particles(q, p, ids)  # each unit keeps what lies in its box
    assert sum(u.counts()[0] for u in units) == len(q)
    return units


def make_domain(L, q, p, ids, dt=0.005):
    d = mdgpu.Domain(L, dt=dt, rc=RC, margin=MARGIN, capacity=len(q))
    d.add_particles(q, p, ids)
    return d


def kicks_by_id(doms):
    ids, dp = [], []
    for d in doms:
        _, _, t = d.particles()
        ids.append(t)
        dp.append(d.kick())
    ids, dp = np.concatenate(ids), np.concatenate(dp)
    o = np.argsort(ids)
    return ids[o], dp[o]


def state_by_id(doms):
    q, p, t = zip(*(d.particles() for d in doms))
    q, p, t = np.concatenate(q), np.concatenate(p), np.concatenate(t)
    o = np.argsort(t)
    return q[o], p[o], t[o]


def observe_sum(doms):
    return np.sum([d.observe_raw() for d in doms], axis=0)


def check_boxes(units, n):
    """Every owned particle inside its unit's box (the upper face of the whole box
    included, as after Variables::AdjustPeriodicBoundary), each id once."""
    seen = []
    for u in units:
        q, _, t = u.particles()
        seen.append(t)
        assert np.all(q >= u.s) and np.all((q < u.e) | (q == u.L)), (u.pos, q.min(0), q.max(0))
    seen = np.concatenate(seen)
    assert len(seen) == n and len(np.unique(seen)) == n


def lattice_state(L, rho, v0, seed, jitter=0.0):
    x = fcc(L, rho)
    n = len(x)
    if jitter:
        x = np.mod(x + np.random.default_rng(seed).uniform(-jitter, jitter, x.shape), L)
    p = initial_velocity(n, v0, seed)
    ids = np.random.default_rng(seed + 1).permutation(n).astype(np.int32)
    return x, p, ids


def test_one_unit_places_the_periodic_domains_ghosts():
    """A 1x1x1 grid: the unit is its own neighbour on every face, and its ghosts, pair
    list, kicks and observables are those of the periodic domain, bit for bit, with rows
    sorted by list length (the default) and in cell order."""
    L = 16.0
    q, p, ids = lattice_state(L, 0.8, 1.0, 3, jitter=0.2)
    d = make_domain(L, q, p, ids)
    (u,) = make_units(L, (1, 1, 1), q, p, ids)
    say("one unit vs periodic domain, n =", len(q))
    d.make_pair_list()
    u.group_make_pair_list()
    gs, gc, gq = d.ghosts()
    us, uc, uq = u.ghosts()
    assert np.array_equal(gs, us) and np.array_equal(gc, uc) and np.array_equal(gq, uq)
    assert np.array_equal(undirected(d.pairs()), undirected(u.pairs()))
    assert np.array_equal(d.particles()[0], u.particles()[0])
    assert np.array_equal(d.kick(), u.kick())
    assert d.observe_raw() == u.observe_raw()
    for x in (d, u):
        x.set_option(mdgpu.OPT_ROW_SORT, 0)
    d.make_pair_list()
    u.group_make_pair_list()
    assert np.array_equal(d.particles()[0], u.particles()[0])
    assert np.array_equal(d.ghosts()[2], u.ghosts()[2])
    assert np.array_equal(d.kick(), u.kick())


@pytest.mark.parametrize("L,grid", [
    ((26.0, 20.0, 14.0), (2, 2, 2)),
    ((26.0, 20.0, 14.0), (3, 2, 1)),
    ((26.0, 20.0, 14.0), (4, 1, 2)),
    ((20.0, 20.0, 20.0), (2, 1, 1)),
])
def test_unit_lists_kicks_and_observers_match_the_relay(L, grid):
    """Grids of one, two and more units per axis (the relay's three cases), with particles
    placed exactly on unit faces and on the edges of the border strips: every unit's pair
    set equals the restated relay's (oracle/units.py) exactly, its kicks agree to 1e-13 of
    each particle's sum of pair-kick magnitudes and its potential energy and virial to
    1e-12.  A particle exactly on a unit's low face is not sent down (MDRect::IsInsideEdge is
    strict), so these lists are not the periodic domain's: the relay fixes them."""
    rng = np.random.default_rng(7)
    q, p, ids = lattice_state(L, 0.75, 1.0, 11, jitter=0.3)
    extra = []
    for a in range(3):
        for k in range(grid[a] + 1):
            f = L[a] / grid[a] * k
            for off in (0.0, RC + MARGIN, -(RC + MARGIN)):
                v = rng.uniform(0, 1, 3) * np.array(L)
                v[a] = f + off
                if 0.0 <= v[a] < L[a]:
                    extra.append(v)
    extra = np.array(extra)
    # keep the extras clear of the lattice: no pair closer than 0.8
    from scipy.spatial import cKDTree
    near = cKDTree(extra).query_ball_point(q, 0.8)
    keep = np.array([len(n) == 0 for n in near])
    q, p, ids = q[keep], p[keep], ids[keep]
    q = np.concatenate([q, extra])
    p = np.concatenate([p, rng.normal(0, 1, extra.shape)])
    ids = np.concatenate([ids, np.arange(ids.max() + 1, ids.max() + 1 + len(extra), dtype=np.int32)])
    units = make_units(L, grid, q, p, ids)
    say("grid", grid, "L", L, "n", len(q), "units", [u.counts()[0] for u in units])
    units[0].group_make_pair_list()
    check_boxes(units, len(q))
    ru = OU.distribute(L, grid, q, ids.astype(np.int64))
    rg = OU.relay(ru, grid, L, RC + MARGIN)
    for u, r, g in zip(units, ru, rg):
        assert tuple(u.pos) == tuple(r["pos"])
        rp = OU.unit_pairs(r, g, RC + MARGIN)
        want = O.canonical_pairs(OU.canonical_unit_pairs(r, g, rp))
        got = O.canonical_pairs(u.pairs())
        assert np.array_equal(got, want), (u.pos, len(got), len(want))
        dp, mag, pe, vir = OU.unit_kicks_and_observers(r, g, rp, RC, 0.005)
        _, _, t = u.particles()
        o = {int(v): k for k, v in enumerate(r["ids"])}
        idx = np.array([o[int(v)] for v in t])
        k = u.kick()
        err = np.abs(k - dp[idx]).max(axis=1)
        assert np.all(err <= 1e-13 * np.maximum(mag[idx], 1e-300)), (u.pos, err.max())
        _, gpe, gvir = u.observe_raw()
        assert gpe == pytest.approx(pe, rel=1e-12) and gvir == pytest.approx(vir, rel=1e-12), u.pos


def run_sequence(doms, group, steps, observe_every, on_rebuild=None):
    """MDManager::Calculate (mdmanager.cc:116-146): expiry of every unit, the group's
    construction if any expired, half drift, halo, force, half drift."""
    rows, rebuilds = [], []
    for s in range(steps):
        expired = [d.is_pair_list_expired() for d in doms]
        if any(expired):
            if group:
                doms[0].group_make_pair_list()
            else:
                doms[0].make_pair_list()
            rebuilds.append(s)
            if on_rebuild:
                on_rebuild()
        for d in doms:
            d.update_position_half()
        for d in doms:
            d.send_border_particles()
        for d in doms:
            d.calculate_force()
            d.update_position_half()
        if (s + 1) % observe_every == 0:
            rows.append(observe_sum(doms))
    return np.array(rows), rebuilds


@pytest.mark.parametrize("grid", [(2, 2, 2), (3, 1, 2)])
def test_units_follow_the_periodic_domain(grid):
    """A hot lattice melting for 300 steps at dt 0.005: particles migrate between units at
    every construction; positions (modulo L) and momenta by id stay within 1e-9 of the
    periodic domain's and the kinetic energy within 1e-10 relative every 50 steps.  The
    observers read ghosts as the last halo exchange left them (after the first half drift,
    as the reference's do), so potential energy and virial depend on which pairs cross a
    unit face; with every halo refreshed after the run they agree to 1e-9."""
    L = 20.0
    q, p, ids = lattice_state(L, 0.8, 2.0, 5)
    n = len(q)
    d = make_domain(L, q, p, ids)
    units = make_units(L, grid, q, p, ids)
    owner0 = {}
    for k, u in enumerate(units):
        for t in u.particles()[2]:
            owner0[int(t)] = k
    say("trajectory grid", grid, "n", n)
    ref_rows, ref_rb = run_sequence([d], False, 300, 50)
    rows, rb = run_sequence(units, True, 300, 50, on_rebuild=lambda: check_boxes(units, n))
    say("rebuilds: domain", ref_rb, "units", rb)
    assert len(rb) >= 3
    moved = sum(1 for k, u in enumerate(units) for t in u.particles()[2] if owner0[int(t)] != k)
    say("particles in another unit than at the start:", moved)
    assert moved > 0
    np.testing.assert_allclose(rows[:, 0], ref_rows[:, 0], rtol=1e-10, atol=0)
    for x in [d] + units:
        x.send_border_particles()
    np.testing.assert_allclose(observe_sum(units), d.observe_raw(), rtol=1e-9, atol=0)
    q1, p1, t1 = state_by_id([d])
    q2, p2, t2 = state_by_id(units)
    assert np.array_equal(t1, t2)
    dq = np.remainder(q1 - q2 + L / 2, L) - L / 2
    assert np.abs(dq).max() < 1e-9 and np.abs(p1 - p2).max() < 1e-9, (np.abs(dq).max(), np.abs(p1 - p2).max())


def test_unit_refusals():
    """What a unit refuses: the fused step, rescaling, a second unit at one grid position, a
    unit of another grid; a pair list before its group's construction."""
    L = 20.0
    units = mdgpu.unit_grid(L, (2, 1, 1), rc=RC, margin=MARGIN)
    q, p, ids = lattice_state(L, 0.5, 1.0, 2)
    for u in units:
        u.add_particles(q, p, ids)
    with pytest.raises(mdgpu.MDGPUError):
        units[0].make_pair_list()
    units[0].group_make_pair_list()
    with pytest.raises(mdgpu.MDGPUError):
        units[0].calculate(1)
    with pytest.raises(mdgpu.MDGPUError):
        units[0].change_scale(1.01)
    with pytest.raises(mdgpu.MDGPUError):
        mdgpu.Unit(L, (2, 1, 1), (0, 0, 0), (0, 0, 0), (10, 20, 20), join=units[0], rc=RC, margin=MARGIN)
    with pytest.raises(mdgpu.MDGPUError):
        mdgpu.Unit(L, (1, 2, 1), (0, 1, 0), (0, 10, 0), (20, 20, 20), join=units[0], rc=RC, margin=MARGIN)
    # the group still steps
    run_sequence(units, True, 5, 5)


@pytest.mark.parametrize("thermostat,precision", [(mdgpu.NVE, mdgpu.FP64), (mdgpu.LANGEVIN, mdgpu.FP64),
                                                   (mdgpu.NVE, mdgpu.MIXED)])
def test_group_calculate_is_the_sequence(thermostat, precision):
    """mdgpu_group_calculate (the whole of MDManager::Calculate for the group in one call)
    gives the trajectory of the same steps taken through the per-unit calls, bit for bit:
    the same per-brick work in the same order (the units' kicks in one launch of all their
    bricks instead of one launch each), the expiry tests batched over the group."""
    L = 20.0
    q, p, ids = lattice_state(L, 0.8, 1.5, 13)
    a = make_units(L, (2, 2, 2), q, p, ids)
    b = make_units(L, (2, 2, 2), q, p, ids)
    for u in a + b:
        u.set_thermostat(thermostat, T=1.2, gamma=0.5)
        u.set_seed(7)
        u.set_precision(precision)
    say("group_calculate vs per-unit sequence, thermostat", thermostat, "precision", precision)
    if thermostat == mdgpu.LANGEVIN:
        rows, rb = [], []
        for s in range(200):
            if any([u.is_pair_list_expired() for u in a]):
                a[0].group_make_pair_list()
                rb.append(s)
            for u in a:
                u.update_position_half()
            for u in a:
                u.send_border_particles()
            for u in a:
                u.calculate_force()
                u.langevin()
    else:
        _, rb = run_sequence(a, True, 200, 50)
    nb = b[0].group_calculate(200)
    assert nb == len(rb), (nb, rb)
    qa, pa, ta = state_by_id(a)
    qb, pb, tb = state_by_id(b)
    assert np.array_equal(ta, tb)
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)


def test_group_nose_hoover_follows_the_periodic_domain():
    """The Nose-Hoover step of the group (MDManager::CalculateNoseHoover, the group's
    temperature at both half steps) against the periodic domain's over 200 steps."""
    L = 20.0
    q, p, ids = lattice_state(L, 0.8, 1.5, 17)
    d = make_domain(L, q, p, ids)
    units = make_units(L, (2, 2, 2), q, p, ids)
    for x in [d] + units:
        x.set_thermostat(mdgpu.NOSEHOOVER, T=0.9, tau=0.1)
    d.make_pair_list()
    units[0].group_make_pair_list()
    d.calculate(200)
    units[0].group_calculate(200)
    q1, p1, t1 = state_by_id([d])
    q2, p2, t2 = state_by_id(units)
    dq = np.remainder(q1 - q2 + L / 2, L) - L / 2
    assert np.abs(dq).max() < 1e-9 and np.abs(p1 - p2).max() < 1e-9, (np.abs(dq).max(), np.abs(p1 - p2).max())
    for u in units:
        assert u.zeta() == pytest.approx(d.zeta(), rel=1e-9, abs=1e-12)
