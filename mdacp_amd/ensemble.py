"""Ensembles of independent Langevin replicas (host-side set-up).

replica_table()   the batched-replica workload: per replica an FCC lattice of
                  N = 4 k^3 particles (N drawn log-uniform, k rounded), a
                  density, a Langevin temperature and a key, drawn in that order
                  from one seeded generator.
box_for(k, rho)   the cubic box edge for which the FCC maker
                  (ConfigurationMaker, confmaker.cc:8-44) places exactly k cells
                  per axis.
build()           positions and momenta (Variables::SetInitialVelocity,
                  variables.cc:82-108, |v0| = sqrt(3 T), key = replica seed) of
                  every replica, replica after replica.
"""
import math

import numpy as np

from .lattice import fcc, initial_velocity


def box_for(k, rho):
    s = 1.0 / math.pow(rho * 0.25, 1.0 / 3.0)
    L = k * s
    while int(L / s) < k:
        L = np.nextafter(L, np.inf)
    return float(L)


def replica_table(nrep=256, nmin=2048, nmax=65536, rho=(0.3, 0.9), T=(0.7, 1.5), seed=20261202):
    """Rows (k, N, rho, T, seed) of the batched workload, drawn in that order."""
    g = np.random.Generator(np.random.MT19937(seed))
    rows = []
    for r in range(nrep):
        n = math.exp(g.uniform(math.log(nmin), math.log(nmax)))
        k = int(round((n / 4.0) ** (1.0 / 3.0)))
        d = float(g.uniform(*rho))
        t = float(g.uniform(*T))
        key = int(g.integers(0, 2 ** 63))
        rows.append(dict(k=k, N=4 * k ** 3, rho=d, T=t, seed=key))
    return rows


def build(rows):
    """boxes (R, 3), counts (R,), q, p (sum N, 3), ids (per replica 0..N-1)."""
    boxes, counts, qs, ps, ids = [], [], [], [], []
    for row in rows:
        L = box_for(row["k"], row["rho"])
        x = fcc(L, row["rho"])
        n = len(x)
        assert n == row["N"], (n, row)
        boxes.append((L, L, L))
        counts.append(n)
        qs.append(x)
        ps.append(initial_velocity(n, math.sqrt(3.0 * row["T"]), row["seed"] & 0xFFFFFFFF))
        ids.append(np.arange(n, dtype=np.int32))
    return (np.array(boxes), np.array(counts, dtype=np.int32), np.concatenate(qs), np.concatenate(ps),
            np.concatenate(ids))
