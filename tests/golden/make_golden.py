"""Write golden fixtures from the reference itself.

Runs oracle/_ref/mdacp_ref (kaityo256/mdacp compiled unmodified from its own
sources, default flags, rc = 3.0 and margin 0.3 as compiled in) on seeded FCC
inputs and stores small fixtures: hashes of canonical pair sets, per-particle
kicks for a small system, observables and rebuild schedules, and one binary
dump (MDManager::SaveAsDumpFile).

    python tests/golden/make_golden.py [case ...]
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mdacp_amd.ensemble import box_for  # noqa: E402
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
             tau=0.1, seed=1, reverse_ids=False, keep_kicks=False, ref_init_v0=False, change_scale=None):
    """ref_init_v0: the momenta come from the reference's own SetInitialVelocity(v0)
    (MDManager::SetInitialVelocity, variables.cc:82-108) instead of the host's;
    change_scale = (alpha, step): MDManager::ChangeScale(alpha) after that step."""
    x = fcc(L, rho)
    n = len(x)
    p = np.zeros_like(x) if ref_init_v0 else initial_velocity(n, v0, 0)
    ids = np.arange(n, dtype=np.int32)
    if reverse_ids:
        ids = ids[::-1].copy()
    recs = O.run_reference(L, x, p, ids, steps=steps, dump_every=dump_every, dt=dt, thermostat=thermostat,
                           T=T, gamma=gamma, tau=tau, seed=seed, init_v0=v0 if ref_init_v0 else None,
                           change_scale=change_scale)
    out = dict(name=name, L=L, density=rho, v0=v0, n=n, steps=steps, dt=dt, rc=3.0, margin=0.3,
               thermostat=thermostat, T=T, gamma=gamma, tau=tau, seed=seed, reverse_ids=reverse_ids,
               ref_init_v0=ref_init_v0, change_scale=change_scale,
               rebuild_steps=[r["step"] for r in recs if r["rebuilt"] and not r.get("scaled")], dumps=[])
    for r in recs:
        if not r["dump"]:
            continue
        ph, pc = pair_hash(r["pairs"])
        d = dict(step=r["step"], scaled=r["scaled"], pn=r["pn"], tn=r["tn"], ke=r["ke"], pe=r["pe"],
                 vir=r["vir"], T=r["T"], P=r["P"], E=r["E"], time=r["time"], pair_sha256=ph, pair_count=pc,
                 state_sha256=state_hash(r))
        out["dumps"].append(d)
        if keep_kicks and r["step"] == 0:
            o = np.argsort(r["type"], kind="stable")
            np.save(os.path.join(OUT, f"{name}_kick0.npy"), r["kick"][o])
            np.save(os.path.join(OUT, f"{name}_pairs0.npy"), O.canonical_pairs(r["pairs"]))
    with open(os.path.join(OUT, f"{name}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, n, "rebuilds", out["rebuild_steps"][:12], flush=True)


def dump_case(name, L, rho, v0, steps, dt=0.001):
    """The reference's binary dump (MDManager::SaveAsDumpFile, mdmanager.cc:474-504: x, y, z,
    px, py, pz per particle, unit by unit) after `steps` NVE steps, rows in the program's
    own order."""
    x = fcc(L, rho)
    n = len(x)
    _, dump = O.run_reference(L, x, initial_velocity(n, v0, 0), np.arange(n, dtype=np.int32), steps=steps,
                              dump_every=steps, dt=dt, kick=False, dump_file_step=steps)
    np.save(os.path.join(OUT, f"{name}.npy"), dump)
    print(name, dump.shape, flush=True)


CASES = {
    # small NVE system: full pair set and kicks at step 0, state hashes to 200 steps
    "fcc2048_nve": lambda: run_case("fcc2048_nve", 16.0, 0.5, np.sqrt(3.0), 200, 50, keep_kicks=True),
    # Langevin and Nose-Hoover paths of MDManager::Calculate (dt 0.005)
    "fcc2048_langevin": lambda: run_case("fcc2048_langevin", 16.0, 0.5, 1.0, 120, 40, dt=0.005,
                                         thermostat="langevin", T=0.9, gamma=0.5, seed=7, reverse_ids=True),
    "fcc2048_nosehoover": lambda: run_case("fcc2048_nosehoover", 16.0, 0.5, 1.0, 120, 40, dt=0.005,
                                           thermostat="nosehoover", T=0.9, tau=0.2, reverse_ids=True),
    # 131072 particles, rho 0.5, L 64: 1000 NVE steps, observables every 100
    "fcc131072_nve": lambda: run_case("fcc131072_nve", 64.0, 0.5, np.sqrt(3.0), 1000, 100),
    # the binary dump of the small NVE system after 100 steps
    "fcc2048_nve_dump100": lambda: dump_case("fcc2048_nve_dump100", 16.0, 0.5, np.sqrt(3.0), 100),
    # the quench path in small: FCC at rho 0.75 with the reference's own initial
    # momenta (SetInitialVelocity), 50 NVE steps, ChangeScale to rho 0.6, 100 more
    "fcc2048_quench": lambda: run_case("fcc2048_quench", box_for(8, 0.75), 0.75, 1.5, 150, 25, dt=0.002,
                                       reverse_ids=True, ref_init_v0=True,
                                       change_scale=((0.75 / 0.6) ** (1.0 / 3.0), 50)),
}

if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/mdacp_ref is not built: make -C oracle ref")
    # arguments: names of the cases to write (default: all)
    for name in (sys.argv[1:] or CASES):
        CASES[name]()
