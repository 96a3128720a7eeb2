"""Pins the CPU oracle (oracle/liboracle.so) before the GPU tests trust it.

* against golden fixtures written by the reference itself (tests/golden);
* against the compiled reference directly, when oracle/_ref is built here;
* against what the mathematics fixes: brute-force pair sets, FCC lattice sums,
  the dimer force, momentum conservation, central differences of the energy,
  and published known-answer vectors of the random-number generators.
"""
import hashlib
import math

import numpy as np
import pytest

from helpers import fcc_lattice_sums, lj_C, load_golden, pair_checksum, pair_hash, state_hash_by_id
from mdacp_amd.lattice import fcc, initial_velocity
from oracle import oracle as O


def run_oracle_case(g, steps=None):
    """Replays a reference fixture on the oracle and compares every dump bitwise.
    Fixtures with ref_init_v0 take their momenta from the oracle's SetInitialVelocity
    (variables.cc:82-108); with change_scale = (alpha, step) the oracle's ChangeScale
    (mdmanager.cc:432-443) runs after that step and is compared with the record the
    reference wrote right after its own."""
    steps = g["steps"] if steps is None else steps
    x = fcc(g["L"], g["density"])
    n = len(x)
    ids = np.arange(n, dtype=np.int32)
    if g["reverse_ids"]:
        ids = ids[::-1].copy()
    s = O.OracleSystem(g["L"], g["dt"], g["rc"], g["margin"])
    if g.get("ref_init_v0"):
        s.add(x, None, ids)
        s.set_initial_velocity(g["v0"], 0)
    else:
        s.add(x, initial_velocity(n, g["v0"], 0), ids)
    kind = {"nve": O.NVE, "langevin": O.LANGEVIN, "nosehoover": O.NOSEHOOVER}[g["thermostat"]]
    s.set_thermostat(kind, g["T"], g["gamma"], g["tau"])
    s.set_seed(g["seed"])
    s.make_pair_list()
    dumps = {(d["step"], d.get("scaled", False)): d for d in g["dumps"]}
    scale = g.get("change_scale")
    rebuilt = [0]
    checked = 0

    def compare(d, step):
        q, pp, t = s.state()
        ob = s.observables()
        assert pair_hash(s.pairs()) == (d["pair_sha256"], d["pair_count"]), step
        assert state_hash_by_id(q, pp, t) == d["state_sha256"], step
        for k in ("ke", "pe", "vir", "T", "P", "E", "time"):
            assert ob[k] == d[k], (step, k)

    for step in range(0, steps + 1):
        if step > 0 and s.step():
            rebuilt.append(step)
        if (step, False) in dumps:
            compare(dumps[(step, False)], step)
            checked += 1
        if scale is not None and step == scale[1]:
            s.change_scale(scale[0])
            if (step, True) in dumps:
                compare(dumps[(step, True)], step)
                checked += 1
    assert rebuilt == [r for r in g["rebuild_steps"] if r <= steps]
    return checked


@pytest.mark.parametrize("name", ["fcc2048_nve", "fcc2048_langevin", "fcc2048_nosehoover", "fcc2048_quench"])
def test_oracle_matches_reference_fixture_bitwise(name):
    g = load_golden(name)
    assert run_oracle_case(g) == len(g["dumps"])


def test_oracle_matches_reference_fixture_131k_first_200_steps():
    g = load_golden("fcc131072_nve")
    assert run_oracle_case(g, steps=200) == 3


def test_oracle_kick_and_pairs_match_reference_fixture(golden_dir):
    g = load_golden("fcc2048_nve")
    x = fcc(g["L"], g["density"])
    n = len(x)
    s = O.OracleSystem(g["L"], g["dt"], g["rc"], g["margin"])
    s.add(x, initial_velocity(n, g["v0"], 0))
    s.make_pair_list()
    kick_ref = np.load(f"{golden_dir}/fcc2048_nve_kick0.npy")
    q, p, t = s.state()
    kick = s.kick()[np.argsort(t)]
    assert np.array_equal(kick, kick_ref)
    assert np.array_equal(O.canonical_pairs(s.pairs()), np.load(f"{golden_dir}/fcc2048_nve_pairs0.npy"))


def sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


def test_binary_dump_fixture_is_the_oracle_state(golden_dir):
    """The reference's binary dump (MDManager::SaveAsDumpFile, mdmanager.cc:474-504) after
    100 steps of the fcc2048_nve case holds, up to row order, the oracle's q and p after the
    same steps, bit for bit: 6 doubles per particle (x, y, z, px, py, pz)."""
    g = load_golden("fcc2048_nve")
    x = fcc(g["L"], g["density"])
    n = len(x)
    s = O.OracleSystem(g["L"], g["dt"], g["rc"], g["margin"])
    s.add(x, initial_velocity(n, g["v0"], 0))
    s.make_pair_list()
    for _ in range(100):
        s.step()
    q, p, _ = s.state()
    want = np.load(f"{golden_dir}/fcc2048_nve_dump100.npy")
    assert want.shape == (n, 6)
    assert np.array_equal(sorted_rows(np.concatenate([q, p], axis=1)), sorted_rows(want))


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
def test_oracle_matches_reference_live_shuffled_ids():
    """Random ids, non-lattice start: owned order, migration and rebuilds all bitwise."""
    rng = np.random.default_rng(3)
    L = 17.5
    x = fcc(L, 0.45)
    x = x + rng.uniform(-0.05, 0.05, x.shape)
    x = np.mod(x, L)
    n = len(x)
    ids = rng.permutation(n).astype(np.int32)
    p = rng.normal(0, 1.2, (n, 3))
    recs = O.run_reference(L, x, p, ids, steps=150, dump_every=30, dt=0.002)
    s = O.OracleSystem(L, 0.002, 3.0, 0.3)
    s.add(x, p, ids)
    s.make_pair_list()
    for r in recs[1:]:
        rb = s.step()
        assert rb == r["rebuilt"], r["step"]
        if r["dump"]:
            q, pp, t = s.state()
            assert np.array_equal(q, r["q"]) and np.array_equal(pp, r["p"]) and np.array_equal(t, r["type"])
            assert np.array_equal(O.canonical_pairs(s.pairs()), O.canonical_pairs(r["pairs"]))
            assert np.array_equal(s.kick(), r["kick"])


# ---------------------------------------------------------------- mathematics
def brute_force_pairs(x, ids, L, SL):
    """All (owned, owned-or-image) pairs within SL with the reference's image rule:
    an image at q + s*L exists per axis iff 0 < q < SL (s=+1) or L-SL < q < L (s=-1)."""
    n = len(x)
    S2 = SL * SL
    out = []
    lo = (x > 0.0) & (x < 0.0 + SL)
    hi = (x > L - SL) & (x < L)
    for code in range(27):
        sh = np.array([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1])
        ok = np.ones(n, dtype=bool)
        for d in range(3):
            if sh[d] == 1:
                ok &= lo[:, d]
            elif sh[d] == -1:
                ok &= hi[:, d]
        img = x + np.where(sh > 0, L, np.where(sh < 0, -L, 0.0))
        js = np.nonzero(ok)[0]
        for i in range(n):
            d = x[i] - img[js]
            r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            sel = js[r2 <= S2]
            for j in sel:
                if code == 13:
                    if ids[j] > ids[i]:
                        out.append((ids[i], ids[j], 13))
                else:
                    out.append((ids[i], ids[j], code))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


@pytest.mark.parametrize("rc,margin,seed", [(2.5, 0.5, 0), (3.0, 0.3, 1)])
def test_oracle_pairs_equal_brute_force(rc, margin, seed):
    rng = np.random.default_rng(seed)
    L = 9.3
    n = 400
    x = rng.uniform(0, L, (n, 3))
    ids = rng.permutation(n).astype(np.int32)
    s = O.OracleSystem(L, 0.001, rc, margin)
    s.add(x, None, ids)
    s.make_pair_list()
    bf = brute_force_pairs(x, ids, L, rc + margin)
    assert np.array_equal(O.canonical_pairs(s.pairs()), O.canonical_pairs(bf))


@pytest.mark.parametrize("rc,margin,pe_ref,vir_ref,pairs_ref", [
    (2.5, 0.5, -2.413162698239619, 4.804417156711153, 27),
    (3.0, 0.3, -2.749029052037014, 5.221892975936016, 39),
])
def test_oracle_fcc_closed_forms(rc, margin, pe_ref, vir_ref, pairs_ref):
    L = 16.0
    x = fcc(L, 0.5)
    n = len(x)
    pe_sum, vir_sum, half = fcc_lattice_sums(2.0, rc, rc + margin)
    assert half == pairs_ref
    assert pe_sum == pytest.approx(pe_ref, rel=1e-13)
    assert vir_sum == pytest.approx(vir_ref, rel=1e-13)
    s = O.OracleSystem(L, 0.001, rc, margin)
    s.add(x)
    s.make_pair_list()
    ob = s.observables()
    assert ob["pe"] / n == pytest.approx(pe_ref, rel=1e-12)
    assert ob["vir"] / n == pytest.approx(vir_ref, rel=1e-12)
    pr = s.pairs()
    own = pr[:, 2] == 13
    assert int(own.sum()) + int((~own).sum()) // 2 == pairs_ref * n
    kick = s.kick()
    assert np.max(np.abs(kick)) < 1e-12  # perfect lattice: zero net force
    C0, C2 = lj_C(rc)
    if rc == 2.5:
        assert C2 == -0.0019499738726400002 and C0 == 0.016266559488


def test_oracle_dimer_kick_closed_form():
    L = 20.0
    rc, dt = 2.5, 0.001
    C0, C2 = lj_C(rc)
    for r in [0.95, 1.12, 1.5, 2.2, 2.49]:
        x = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        s = O.OracleSystem(L, dt, rc, 0.5)
        s.add(x)
        s.make_pair_list()
        kick = s.kick()
        # F_x on particle 0 = -dV/dr * (x0 - x1)/r with V = 4(r^-12 - r^-6 + C2 r^2 + C0)
        dVdr = 4.0 * (-12.0 * r ** -13 + 6.0 * r ** -7 + 2.0 * C2 * r)
        f0 = dVdr  # (x1 - x0)/r = +1
        assert kick[0, 0] == pytest.approx(f0 * dt, rel=1e-12)
        assert kick[1, 0] == pytest.approx(-f0 * dt, rel=1e-12)
        assert np.all(kick[:, 1:] == 0.0)
        pe = s.observables()["pe"]
        assert pe == pytest.approx(4.0 * (r ** -12 - r ** -6 + C2 * r * r + C0), rel=1e-12)


def test_oracle_force_is_minus_gradient_and_conserves_momentum():
    rng = np.random.default_rng(5)
    L = 10.0
    x = fcc(L, 0.6) + rng.uniform(-0.1, 0.1, (len(fcc(L, 0.6)), 3))
    x = np.mod(x, L)
    n = len(x)
    dt = 0.001

    def pe_of(xx):
        s = O.OracleSystem(L, dt, 2.5, 0.5)
        s.add(xx)
        s.make_pair_list()
        return s.observables()["pe"]

    s = O.OracleSystem(L, dt, 2.5, 0.5)
    s.add(x)
    s.make_pair_list()
    q, _, t = s.state()
    kick = s.kick()[np.argsort(t)]
    assert np.abs(kick.sum(axis=0)).max() < 1e-13 * np.abs(kick).sum()
    h = 1e-6
    for i in [0, 7, n // 2]:
        for d in range(3):
            xp = x.copy(); xp[i, d] += h
            xm = x.copy(); xm[i, d] -= h
            g = (pe_of(xp) - pe_of(xm)) / (2 * h)
            assert kick[i, d] / dt == pytest.approx(-g, rel=1e-5, abs=1e-6)


def test_mt19937_known_answers():
    # C++ [rand.predef]: the 10000th output of a default-seeded mt19937 is 4123659995
    out = np.zeros(10000, dtype=np.uint32)
    O.lib().ora_mt_raw(5489, 10000, out.ctypes.data_as(O.ctypes.c_void_p))
    assert int(out[-1]) == 4123659995
    assert int(out[0]) == 3499211612
    from mdacp_amd.lattice import _mt19937_raw
    assert np.array_equal(_mt19937_raw(5489, 10000).astype(np.uint32), out)


def test_philox4x32_10_known_answers():
    # Random123 kat_vectors for philox4x32 with 10 rounds
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, exp in cases:
        c = np.array(ctr, dtype=np.uint32); k = np.array(key, dtype=np.uint32); o = np.zeros(4, dtype=np.uint32)
        O.lib().ora_philox4x32(c.ctypes.data_as(O.ctypes.c_void_p), k.ctypes.data_as(O.ctypes.c_void_p),
                               o.ctypes.data_as(O.ctypes.c_void_p))
        assert tuple(int(v) for v in o) == exp


def test_normal_stream_statistics():
    out = np.zeros(200000)
    O.lib().ora_mt_normal(12345, 2.0, len(out), out.ctypes.data_as(O.ctypes.c_void_p))
    assert abs(out.mean()) < 0.02 and abs(out.std() - 2.0) < 0.02
    phil = np.array([[0.0, 0.0, 0.0]] * 20000)
    for i in range(20000):
        buf = np.zeros(3)
        O.lib().ora_philox_normal3(99, i, 7, buf.ctypes.data_as(O.ctypes.c_void_p))
        phil[i] = buf
    assert abs(phil.mean()) < 0.02 and abs(phil.std() - 1.0) < 0.02


def test_host_initial_conditions_match_oracle():
    for L, rho in [(16.0, 0.5), (20.0, 0.7)]:
        x = fcc(L, rho)
        assert np.array_equal(x, O.fcc(L, rho))
        s = O.OracleSystem(L, 0.001, 2.5, 0.5)
        s.add(x)
        s.set_initial_velocity(math.sqrt(3.0), 0)
        assert np.array_equal(s.state()[1], initial_velocity(len(x), math.sqrt(3.0), 0))


def test_langevin_stationary_variance_closed_form():
    """With F = 0 the Langevin map p <- p + (-g p + n*D) dt has <p^2> = T/(1 - g dt/2)."""
    L = 40.0
    x = np.array([[5.0, 5.0, 5.0], [25.0, 25.0, 25.0]])  # beyond rc: no force
    s = O.OracleSystem(L, 0.005, 2.5, 0.5)
    s.add(x)
    s.set_thermostat(O.LANGEVIN, 1.3, 1.0, 0.1)
    s.set_noise(O.NOISE_PHILOX, 11)
    s.make_pair_list()
    acc = []
    for k in range(20000):
        s.step()
        if k > 2000:
            acc.append(s.state()[1].ravel() ** 2)
    m = np.mean(acc)
    assert m == pytest.approx(1.3 / (1 - 1.0 * 0.005 / 2), rel=0.05)


# ------------------------------------------------- ChangeScale, SetInitialVelocity
def test_quench_fixture_covers_initial_velocity_and_change_scale():
    """The fixture that pins ora_set_initial_velocity and ora_change_scale holds what
    it is meant to: the reference's own SetInitialVelocity momenta at step 0, one
    ChangeScale record, and rebuilds after it."""
    g = load_golden("fcc2048_quench")
    assert g["ref_init_v0"] and g["change_scale"] is not None
    alpha, at = g["change_scale"]
    scaled = [d for d in g["dumps"] if d["scaled"]]
    assert len(scaled) == 1 and scaled[0]["step"] == at
    assert any(r > at for r in g["rebuild_steps"])
    # the host's initial momenta (mdacp_amd.lattice, used by the GPU tests and the
    # bench) equal the reference's SetInitialVelocity bit for bit
    x = fcc(g["L"], g["density"])
    n = len(x)
    ids = np.arange(n, dtype=np.int32)[::-1].copy()
    d0 = [d for d in g["dumps"] if d["step"] == 0][0]
    assert state_hash_by_id(x, initial_velocity(n, g["v0"], 0), ids) == d0["state_sha256"]


def test_initial_velocity_mathematics():
    """Every particle gets speed V0 in a random direction and then the mean is
    removed: so there is one vector c (the removed mean) with |p_i + c| = V0 for all
    i, found here by least squares from 2 p_i.c + |c|^2 = V0^2 - |p_i|^2; the result
    has zero mean, and the directions are uniform on the sphere (<z> = 0, <z^2> = 1/3)."""
    n, v0 = 20000, 1.7
    s = O.OracleSystem(40.0, 0.001, 2.5, 0.5)
    s.add(np.random.default_rng(0).uniform(0, 40.0, (n, 3)))
    s.set_initial_velocity(v0, 3)
    p = s.state()[1]
    assert np.abs(p.mean(axis=0)).max() < 1e-15
    A = np.concatenate([2.0 * p, np.ones((n, 1))], axis=1)
    b = v0 * v0 - np.einsum("ij,ij->i", p, p)
    sol = np.linalg.lstsq(A, b, rcond=None)[0]
    c = sol[:3]
    assert sol[3] == pytest.approx(c @ c, rel=1e-6, abs=1e-12)
    speed = np.linalg.norm(p + c, axis=1)
    assert np.abs(speed - v0).max() < 1e-12
    u = (p + c) / v0
    assert abs(np.mean(u[:, 2])) < 0.02 and abs(np.mean(u[:, 2] ** 2) - 1.0 / 3.0) < 0.02
    # a plausible slip -- a direction drawn from z = u instead of z = 2u - 1 -- would
    # put every z >= 0: the test above rejects it; so does the reference fixture


def test_change_scale_mathematics():
    """ora_change_scale: positions become fl(q * alpha) bit for bit (by id), momenta do
    not change, the box becomes alpha L, the pair set equals brute force at the new
    geometry, and a perfect FCC crystal at spacing a becomes one at spacing a alpha
    with the closed-form lattice sums."""
    rng = np.random.default_rng(11)
    L, alpha = 9.3, 1.0772173450159417
    n = 300
    x = rng.uniform(0, L, (n, 3))
    ids = rng.permutation(n).astype(np.int32)
    p = rng.normal(0, 1, (n, 3))
    s = O.OracleSystem(L, 0.001, 2.5, 0.5)
    s.add(x, p, ids)
    s.make_pair_list()
    q0, p0, t0 = s.state()
    s.change_scale(alpha)
    q1, p1, t1 = s.state()
    o0, o1 = np.argsort(t0), np.argsort(t1)
    assert np.array_equal(q1[o1], q0[o0] * alpha)
    assert np.array_equal(p1[o1], p0[o0])
    assert np.array_equal(s.L, np.full(3, L * alpha))
    bf = brute_force_pairs(q0[o0] * alpha, np.sort(t0), L * alpha, 3.0)
    assert np.array_equal(O.canonical_pairs(s.pairs()), O.canonical_pairs(bf))
    # rescaled perfect crystal: rho 0.5 (a = 2) to a = 2 alpha
    Lf = 16.0
    xf = fcc(Lf, 0.5)
    nf = len(xf)
    for rc, margin in [(2.5, 0.5), (3.0, 0.3)]:
        f = O.OracleSystem(Lf, 0.001, rc, margin)
        f.add(xf)
        f.change_scale(alpha)
        pe_sum, vir_sum, _ = fcc_lattice_sums(2.0 * alpha, rc, rc + margin)
        ob = f.observables()
        assert ob["pe"] / nf == pytest.approx(pe_sum, rel=1e-12)
        assert ob["vir"] / nf == pytest.approx(vir_sum, rel=1e-12)
        assert np.max(np.abs(f.kick())) < 1e-12


def test_pair_checksum_detects_changes():
    """The order-independent checksum used for the 4M and 16.4M pair sets: invariant under
    reordering; a dropped, added or altered pair (partner id or image code) changes it."""
    rng = np.random.default_rng(9)
    a = rng.integers(0, 1 << 24, 5000)
    b = rng.integers(0, 1 << 24, 5000)
    c = rng.integers(0, 27, 5000)
    p = np.stack([np.minimum(a, b), np.maximum(a, b), c], 1).astype(np.int32)
    ref = pair_checksum(p)
    assert pair_checksum(p[rng.permutation(len(p))]) == ref
    assert pair_checksum(p, chunk=77) == ref
    assert pair_checksum(p[1:]) != ref
    q = p.copy(); q[10, 1] += 1
    assert pair_checksum(q) != ref
    q = p.copy(); q[10, 2] = (q[10, 2] + 1) % 27
    assert pair_checksum(q) != ref
    q = np.concatenate([p, p[:1]])
    assert pair_checksum(q) != ref
