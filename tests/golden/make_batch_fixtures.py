"""Oracle reference values for the batched and served workloads.

The reference's own tests hold no workload of this kind, so the expected
values come from the FP64 oracle (itself pinned against the reference's golden
vectors, tests/test_oracle_golden.py), run on the CPU with one molecule per
process:

    python tests/golden/make_batch_fixtures.py [procs]

writes
  tests/golden/batch512_oracle.npz  the 512-molecule batched workload
      (seeded water clusters / all-trans alkanes, 10-200 atoms log-uniform,
      seed 2024: xtb_amd.workloads.mixed_batch(512, 2024))
  tests/golden/serve_oracle.npz     the request mix of the serving test
      (xtb_amd.workloads.mixed_batch(24, 4321, lo=20, hi=400): 20-400 atoms)

Each file holds per molecule the total energy, gradient norm, HOMO-LUMO gap,
SCC iteration count and convergence flag, and the atomic charges and the
gradient of all molecules concatenated (offsets from the atom counts).
"""
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from xtb_amd import workloads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SETS = {
    "batch512_oracle.npz": dict(nmol=512, seed=2024, lo=10, hi=200),
    "serve_oracle.npz": dict(nmol=24, seed=4321, lo=20, hi=400),
}


def _one(args):
    from oracle import pyoracle as po
    z, x = args
    r = po.singlepoint(z, x, lapack=True, threads=1)
    return (r["e_total"], r["gnorm"], r["hl_gap_ev"], r["iterations"], r["converged"],
            np.asarray(r["q"]), np.asarray(r["gradient"]))


def make(name, spec, procs):
    mols = workloads.mixed_batch(spec["nmol"], seed=spec["seed"], lo=spec["lo"], hi=spec["hi"])
    t = time.time()
    # largest first so that the pool's tail is short
    order = sorted(range(len(mols)), key=lambda m: -len(mols[m][0]))
    with mp.get_context("fork").Pool(procs) as pool:
        got = dict(zip(order, pool.map(_one, [mols[m] for m in order], chunksize=1)))
    res = [got[m] for m in range(len(mols))]
    out = dict(
        natoms=np.array([len(z) for z, _ in mols], dtype=np.int32),
        energy=np.array([r[0] for r in res]), gnorm=np.array([r[1] for r in res]),
        hl_gap_ev=np.array([r[2] for r in res]), iterations=np.array([r[3] for r in res], dtype=np.int32),
        converged=np.array([r[4] for r in res], dtype=np.int32),
        charges=np.concatenate([r[5] for r in res]), gradient=np.concatenate([r[6].ravel() for r in res]),
        seed=np.int64(spec["seed"]), lo=np.int64(spec["lo"]), hi=np.int64(spec["hi"]))
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **out)
    print(f"{path}: {len(mols)} molecules, {out['natoms'].sum()} atoms, {time.time() - t:.1f} s", flush=True)


def main():
    procs = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    for name, spec in SETS.items():
        make(name, spec, procs)


if __name__ == "__main__":
    main()
