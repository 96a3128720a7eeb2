"""Write golden fixtures from the reference itself.

Runs oracle/_ref/mdacp_ref (kaityo256/mdacp compiled unmodified from its own
sources, default flags, rc = 3.0 and margin 0.3 as compiled in) on seeded FCC
inputs and stores small fixtures: hashes of canonical pair sets, per-particle
kicks for a small system, observables and rebuild schedules.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mdacp_amd.lattice import fcc, initial_velocity  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def pair_hash(pairs):
    keys = O.canonical_pairs(pairs)
    return hashlib.sha256(keys.tobytes()).hexdigest(), int(len(keys))


def state_hash(rec):
    o = np.argsort(rec["type"], kind="stable")
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rec["q"][o]).tobytes())
    h.update(np.ascontiguousarray(rec["p"][o]).tobytes())
    return h.hexdigest()


def run_case(name, L, rho, v0, steps, dump_every, dt=0.001, thermostat="nve", T=1.0, gamma=0.1,
             tau=0.1, seed=1, reverse_ids=False, keep_kicks=False):
    x = fcc(L, rho)
    n = len(x)
    p = initial_velocity(n, v0, 0)
    ids = np.arange(n, dtype=np.int32)
    if reverse_ids:
        ids = ids[::-1].copy()
    recs = O.run_reference(L, x, p, ids, steps=steps, dump_every=dump_every, dt=dt, thermostat=thermostat,
                           T=T, gamma=gamma, tau=tau, seed=seed)
    out = dict(name=name, L=L, density=rho, v0=v0, n=n, steps=steps, dt=dt, rc=3.0, margin=0.3,
               thermostat=thermostat, T=T, gamma=gamma, tau=tau, seed=seed, reverse_ids=reverse_ids,
               rebuild_steps=[r["step"] for r in recs if r["rebuilt"]], dumps=[])
    for r in recs:
        if not r["dump"]:
            continue
        ph, pc = pair_hash(r["pairs"])
        d = dict(step=r["step"], pn=r["pn"], tn=r["tn"], ke=r["ke"], pe=r["pe"], vir=r["vir"], T=r["T"],
                 P=r["P"], E=r["E"], time=r["time"], pair_sha256=ph, pair_count=pc, state_sha256=state_hash(r))
        out["dumps"].append(d)
        if keep_kicks and r["step"] == 0:
            o = np.argsort(r["type"], kind="stable")
            np.save(os.path.join(OUT, f"{name}_kick0.npy"), r["kick"][o])
            np.save(os.path.join(OUT, f"{name}_pairs0.npy"), O.canonical_pairs(r["pairs"]))
    with open(os.path.join(OUT, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, n, "rebuilds", out["rebuild_steps"][:12], flush=True)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/mdacp_ref is not built: make -C oracle ref")
    # small NVE system: full pair set and kicks at step 0, state hashes to 200 steps
    run_case("fcc2048_nve", 16.0, 0.5, np.sqrt(3.0), 200, 50, keep_kicks=True)
    # Langevin and Nose-Hoover paths of MDManager::Calculate (dt 0.005)
    run_case("fcc2048_langevin", 16.0, 0.5, 1.0, 120, 40, dt=0.005, thermostat="langevin", T=0.9,
             gamma=0.5, seed=7, reverse_ids=True)
    run_case("fcc2048_nosehoover", 16.0, 0.5, 1.0, 120, 40, dt=0.005, thermostat="nosehoover", T=0.9,
             tau=0.2, reverse_ids=True)
    # 131072 particles, rho 0.5, L 64: 1000 NVE steps, observables every 100
    run_case("fcc131072_nve", 64.0, 0.5, np.sqrt(3.0), 1000, 100)
