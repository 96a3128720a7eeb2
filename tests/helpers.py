"""Shared helpers for the parity tests (numpy restatements used as checks)."""
import hashlib
import json
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


def pair_hash(pairs3):
    keys = O.canonical_pairs(pairs3)
    return hashlib.sha256(keys.tobytes()).hexdigest(), int(len(keys))


def _mix64(x):
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def pair_checksum(pairs3, chunk=1 << 24):
    """Order-independent checksum of a canonical pair set, for sets too large to sort
    quickly (4M and 16.4M particles): the count, and the sum (mod 2^64) and the xor of
    splitmix64 of each canonical key.  Two sets with equal checksums are equal but for
    a 2^-64-scale accident; a missing, extra or changed pair changes the sum."""
    n = len(pairs3)
    tot = np.uint64(0)
    xor = np.uint64(0)
    for a in range(0, n, chunk):
        p = np.asarray(pairs3[a:a + chunk], dtype=np.int64)
        key = ((p[:, 0] << 36) | (p[:, 1] << 5) | p[:, 2]).astype(np.uint64)
        h = _mix64(key)
        with np.errstate(over="ignore"):
            tot = tot + np.sum(h, dtype=np.uint64)
        xor ^= np.bitwise_xor.reduce(h)
    return dict(count=int(n), sum=f"{int(tot):016x}", xor=f"{int(xor):016x}")


def state_hash_by_id(q, p, ids):
    o = np.argsort(ids, kind="stable")
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(q[o]).tobytes())
    h.update(np.ascontiguousarray(p[o]).tobytes())
    return h.hexdigest()


def shift_vec(code, L):
    code = np.asarray(code)
    s = np.stack([code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1], axis=-1).astype(np.float64)
    return s * np.asarray(L)


def image_positions(q_by_id, b, code, L):
    """fl(q_b + shift*L) per component, as the reference's halo computes it."""
    s = shift_vec(code, L)
    return q_by_id[b] + s  # elementwise double add: q + (+-L or 0.0)


def pair_force_sums(q_by_id, pairs3, L, rc, dt, C2):
    """Per-particle sum of |f_ij| * dt over listed pairs within rc (tolerance scale)."""
    a, b, code = pairs3[:, 0], pairs3[:, 1], pairs3[:, 2]
    xb = image_positions(q_by_id, b, code, L)
    d = xb - q_by_id[a]
    r2 = np.einsum("ij,ij->i", d, d)
    inside = r2 <= rc * rc
    r6 = r2 ** 3
    df = ((24.0 * r6 - 48.0) / (r6 * r6 * r2) + 8.0 * C2) * dt
    mag = np.abs(df) * np.sqrt(r2) * inside
    n = len(q_by_id)
    S = np.bincount(a, weights=mag, minlength=n)
    own = code == 13
    S += np.bincount(b[own], weights=mag[own], minlength=n)
    return S


def lj_C(rc):
    s2 = 1.0 / (rc * rc)
    s6 = s2 * s2 * s2
    s8 = s6 * s2
    s12 = s6 * s6
    s14 = s12 * s2
    C2 = 6.0 * s14 - 3.0 * s8
    C0 = -s12 + s6 - C2 / s2
    return C0, C2


def wrap_like_reference(q, L):
    """Variables::AdjustPeriodicBoundary (variables.cc:59-79): one +-L per component."""
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), q.shape)
    return np.where(q < 0.0, q + L, np.where(q > L, q - L, q))


def oracle_system(L, x, p, ids, dt, rc, margin):
    s = O.OracleSystem(L, dt, rc, margin)
    s.add(x, p, ids)
    return s


def fcc_lattice_sums(a, rc, SL):
    """Exact lattice sums of the smoothed LJ potential on an infinite FCC crystal of cubic
    cell a: PE per particle, virial per particle (VirialObserver / N), and the number of
    half pairs per particle within SL."""
    import math
    C0, C2 = lj_C(rc)
    m = int(math.ceil(SL / a)) + 2
    r = np.arange(-m, m + 1)
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    base = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(float) * a
    pts = np.concatenate([base + np.array(b) * a for b in [(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0)]])
    r2 = np.einsum("ij,ij->i", pts, pts)
    r2 = r2[r2 > 0]
    inside = r2[r2 <= rc * rc]
    pe = 0.5 * np.sum(4.0 * (inside ** -6 - inside ** -3 + C2 * inside + C0))
    r6 = inside ** 3
    vir = 0.5 * np.sum(((24.0 * r6 - 48.0) / (r6 * r6 * inside) + C2 * 8.0) * inside) / 3.0
    half_pairs_sl = int(np.sum(r2 <= SL * SL)) // 2
    return pe, vir, half_pairs_sl
