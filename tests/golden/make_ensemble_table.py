"""Write ensemble_replicas.json: the 256-row parameter table of the batched
Langevin workload (mdacp_amd.ensemble.replica_table with its defaults) and
per-row checks of the lattice it builds (particle count, box edge).

Run from the repository root:  python tests/golden/make_ensemble_table.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mdacp_amd.ensemble import box_for, replica_table  # noqa: E402

rows = replica_table()
for r in rows:
    r["L"] = box_for(r["k"], r["rho"])
with open(os.path.join(HERE, "ensemble_replicas.json"), "w") as f:
    json.dump({"generator": "numpy Generator(MT19937(20261202)): per replica log-uniform N in [2048, 65536] "
                            "-> k = round((N/4)^(1/3)), rho U[0.3, 0.9], T U[0.7, 1.5], key U[0, 2^63)",
               "rows": rows}, f, indent=0)
print(len(rows), "rows,", sum(r["N"] for r in rows), "particles")
