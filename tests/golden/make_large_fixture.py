"""Oracle reference values for the 700-water cluster (2100 atoms, 4200 AOs).

The reference's own tests hold no system of this size, so the expected values
come from the FP64 oracle (itself pinned against the reference's golden
vectors).  Run on the CPU (a few minutes with MKL on 8 cores):

    python tests/golden/make_large_fixture.py

writes tests/golden/water700_oracle.json.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from xtb_amd import workloads  # noqa: E402


def main():
    z, x = workloads.large_water_cluster(2024)
    t = time.time()
    r = po.singlepoint(z, x, lapack=True, threads=os.cpu_count() or 1)
    g = np.asarray(r["gradient"])
    q = np.asarray(r["q"])
    out = {
        "system": "700 waters, 10x10x7 grid, seed 2024 (xtb_amd.workloads.large_water_cluster)",
        "natoms": int(len(z)), "nao": int(r["nao"]),
        "e_total": r["e_total"], "gnorm": r["gnorm"], "iterations": int(r["iterations"]),
        "converged": bool(r["converged"]), "hl_gap_ev": r["hl_gap_ev"],
        "dipole": np.asarray(r["dipole"]).tolist(),
        "q_first": q[:24].tolist(), "q_abs_sum": float(np.abs(q).sum()),
        "gradient_first": g[:8].tolist(), "gradient_abs_sum": float(np.abs(g).sum()),
        "oracle_seconds": time.time() - t,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water700_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, out["e_total"], out["gnorm"], out["iterations"])


if __name__ == "__main__":
    main()
