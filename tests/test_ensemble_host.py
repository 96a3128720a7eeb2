"""CPU checks of the batched workload's host set-up (mdacp_amd.ensemble)."""
import json
import os

import numpy as np

from mdacp_amd.ensemble import box_for, build, replica_table
from mdacp_amd.lattice import fcc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ensemble_replicas.json")


def test_table_matches_fixture():
    want = json.load(open(GOLDEN))["rows"]
    got = replica_table()
    assert len(got) == len(want) == 256
    for g, w in zip(got, want):
        for key in ("k", "N", "rho", "T", "seed"):
            assert g[key] == w[key], key
        assert box_for(g["k"], g["rho"]) == w["L"]


def test_table_ranges():
    rows = replica_table()
    N = np.array([r["N"] for r in rows])
    assert N.min() >= 2048 and N.max() <= 4 * 41 ** 3
    assert all(0.3 <= r["rho"] <= 0.9 and 0.7 <= r["T"] <= 1.5 for r in rows)
    assert len({r["seed"] for r in rows}) == len(rows)


def test_box_holds_exactly_k_cells():
    """ConfigurationMaker places int(L/a) cells per axis (confmaker.cc:8-44): the box edge must give k."""
    for k, rho in [(8, 0.9), (13, 0.3024066561245918), (25, 0.3), (11, 0.8409558583069494)]:
        L = box_for(k, rho)
        a = 1.0 / (rho * 0.25) ** (1.0 / 3.0)
        assert len(fcc(L, rho)) == 4 * k ** 3
        assert int(L / a) == k and abs(L - k * a) <= 1e-12 * L


def test_build_layout():
    rows = replica_table()[:5]
    boxes, counts, q, p, ids = build(rows)
    assert list(counts) == [r["N"] for r in rows]
    off = 0
    for r, n in enumerate(counts):
        assert np.all(q[off:off + n] >= 0.0) and np.all(q[off:off + n] < boxes[r])
        assert np.array_equal(ids[off:off + n], np.arange(n))
        # SetInitialVelocity: |p| = sqrt(3T) before the mean is removed, so the mean speed stays near it
        sp = np.sqrt((p[off:off + n] ** 2).sum(axis=1))
        assert abs(sp.mean() - np.sqrt(3 * rows[r]["T"])) < 0.05 * np.sqrt(3 * rows[r]["T"])
        off += n
