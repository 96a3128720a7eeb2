"""Fixtures for the benchmark configurations at their stated parameters, written
from the oracle (oracle/liboracle.so, the C restatement of the reference step;
pinned bitwise against the reference compiled here at its built-in rc 3.0 /
margin 0.3, and by closed forms at rc 2.5 / margin 0.5).  The GPU tests in
tests/test_gpu_configs.py compare the HIP path with these.

  configA_131072  N = 4*32^3 FCC, rho 0.5, L 64, |v0| = sqrt(3) (SetInitialVelocity,
                  seed 0), rc 2.5, margin 0.5, dt 0.001, 1000 NVE steps: T/P/E and the
                  raw observer sums every 100 steps, the rebuild steps, and the SHA-256
                  of the canonical pair set after every construction
  configB_4M      N = 4*100^3, L 200, otherwise the same: the same trace to step 1000
                  and order-independent checksums of the pair set at steps 0 and 1000
  configE_16M     N = 4*160^3 FCC at rho 0.75 (L 279.5457), MDManager::ChangeScale to
                  rho 0.6 (alpha = (0.75/0.6)^(1/3)) at step 0: pair-set checksum and
                  observer sums right after the rescale

    python tests/golden/make_config_fixtures.py [A] [B] [E]
"""
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import pair_checksum  # noqa: E402
from mdacp_amd.ensemble import box_for  # noqa: E402
from mdacp_amd.lattice import fcc, initial_velocity  # noqa: E402
from oracle import oracle as O  # noqa: E402

RC, MARGIN, DT = 2.5, 0.5, 0.001


def sha_pairs(pairs):
    keys = O.canonical_pairs(pairs)
    return hashlib.sha256(keys.tobytes()).hexdigest(), int(len(keys))


def obs_record(o, step):
    ob = o.observables()
    return dict(step=step, **{k: ob[k] for k in ("ke", "pe", "vir", "T", "P", "E", "time")})


def nve_trace(name, k, steps, every, pair_sig, sig_steps=None):
    L = box_for(k, 0.5)
    x = fcc(L, 0.5)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0), 0)
    o = O.OracleSystem(L, DT, RC, MARGIN)
    o.add(x, p)
    t0 = time.time()
    o.make_pair_list()
    out = dict(name=name, n=n, L=L, density=0.5, v0=math.sqrt(3.0), seed=0, rc=RC, margin=MARGIN, dt=DT,
               steps=steps, rebuild_steps=[0], dumps=[obs_record(o, 0)], pairs={})
    if sig_steps is None or 0 in sig_steps:
        out["pairs"]["0"] = pair_sig(o.pairs())
    for step in range(1, steps + 1):
        if o.step():
            out["rebuild_steps"].append(step)
            if sig_steps is None:
                out["pairs"][str(step)] = pair_sig(o.pairs())
        if sig_steps is not None and step in sig_steps:
            out["pairs"][str(step)] = pair_sig(o.pairs())
        if step % every == 0:
            out["dumps"].append(obs_record(o, step))
            print(name, step, f"{time.time() - t0:.0f}s", out["dumps"][-1]["E"], flush=True)
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, "rebuilds", out["rebuild_steps"], flush=True)


def config_a():
    nve_trace("configA_131072", 32, 1000, 100, lambda pr: list(sha_pairs(pr)))


def config_b():
    nve_trace("configB_4M", 100, 1000, 100, pair_checksum, sig_steps={0, 1000})


def config_e():
    k, rho0, rho1 = 160, 0.75, 0.6
    L = box_for(k, rho0)
    x = fcc(L, rho0)
    n = len(x)
    p = initial_velocity(n, math.sqrt(3.0 * 0.9), 0)
    alpha = (rho0 / rho1) ** (1.0 / 3.0)
    o = O.OracleSystem(L, DT, RC, MARGIN)
    o.add(x, p)
    o.make_pair_list()
    o.change_scale(alpha)
    ob = o.observables()
    pr = o.pairs()
    out = dict(name="configE_16M", n=n, L=L, density=rho0, density_after=rho1, alpha=alpha, v0=math.sqrt(2.7),
               seed=0, rc=RC, margin=MARGIN, dt=DT, L_after=float(o.L[0]),
               observe={kk: ob[kk] for kk in ("ke", "pe", "vir", "T", "P", "E")}, pairs=pair_checksum(pr))
    del pr
    with open(os.path.join(HERE, "configE_16M.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("configE_16M", out["pairs"], ob["pe"] / n, flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["A", "B", "E"]
    for w in which:
        {"A": config_a, "B": config_b, "E": config_e}[w]()
