"""Project autotuner of the batched engine on one MI355X (gfx950).

Times, on the device, the choices the engine leaves open and writes them to a
tuning file that every batch reads at set-up (``tuning_gfx950.txt`` next to
``libxtb_amd.so``; environment variables ``XTBAMD_*`` override it):

* ``eig_blk_min`` and ``eig_blk_iters``: the smallest eigenproblem
  dimension that goes to the blocked eigensolver (Householder
  tridiagonalisation, divide and conquer, back-transformation) instead of
  block-Jacobi rounds, and for how many SCC iterations (from the first; the
  block-Jacobi rounds start from the previous orbitals after that).
  One-molecule single points per size bin (water clusters and alkanes, up to
  the 700-water cluster) are timed with block Jacobi, the blocked solver in
  every iteration and in the first 2 or 4; the cutoff is the smallest bin
  from which a blocked variant wins in every larger bin, the iteration count
  that of the largest bin, and both are kept only if the 512-molecule batch
  is not slower with them.
* ``jac_lanes`` (1..3) and ``jac_cross`` (0/1): the stream lanes and the
  cross-round schedule of the block-Jacobi rounds, timed on the 512-molecule
  workload.
* ``gemm`` (mfma / valu): the batched FP64 product kernel.

    python -m xtb_amd.tune [--out DIR] [--steps K]

writes ``DIR/tuning_gfx950.txt`` and ``DIR/autotune_gfx950.md`` (the
measurements).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from xtb_amd import workloads  # noqa: E402

KEYS = ("XTBAMD_EIG_BLK_MIN", "XTBAMD_EIG_BLK_ITERS", "XTBAMD_JAC_LANES", "XTBAMD_JAC_CROSS", "XTBAMD_GEMM")


def _set(env):
    for k in KEYS:
        os.environ.pop(k, None)
    os.environ["XTBAMD_TUNING"] = "off"
    for k, v in env.items():
        os.environ[k] = str(v)


def _time_batch(mols, env, steps, warmup=1):
    import xtb_amd
    _set(env)
    b = xtb_amd.Batch([(z, x, 0.0, 0) for z, x in mols])
    for _ in range(warmup):
        b.run()
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        b.run()
        ts.append(time.perf_counter() - t)
    e = [r["energy"] for r in b.results()]
    b.close()
    return float(np.median(ts)), e


def tune(steps=3, out=HERE, log=print):
    rows_bins, rows_batch = [], []
    # 1. eigensolver path per size bin (one molecule per batch): block Jacobi,
    # the blocked solver in every SCC iteration, or in the first few only
    bins = [("water", 12), ("alkane", 24), ("water", 22), ("alkane", 44), ("water", 43), ("alkane", 86),
            ("water", 64), ("water", 100), ("water", 170), ("water", 256), ("water", 400), ("water", 700)]
    hybrid = (2, 4)
    best_per_bin = []
    for kind, size in bins:
        if kind == "water" and size == 700:
            z, x = workloads.large_water_cluster(2024)
        else:
            z, x = workloads.water_cluster(size, 7) if kind == "water" else workloads.alkane(size, 7)
        nao = workloads.ao_count(z)
        tj, ej = _time_batch([(z, x)], {"XTBAMD_EIG_BLK_MIN": 1 << 30}, steps)
        tb, eb = _time_batch([(z, x)], {"XTBAMD_EIG_BLK_MIN": 65}, steps)
        th = []
        for it in hybrid:
            t, _ = _time_batch([(z, x)], {"XTBAMD_EIG_BLK_MIN": 65, "XTBAMD_EIG_BLK_ITERS": it}, steps)
            th.append(t)
        dE = abs(ej[0] - eb[0])
        rows_bins.append((kind, size, nao, tj, tb, th, dE))
        opts = {1 << 30: tb}
        opts.update({it: t for it, t in zip(hybrid, th)})
        it_best = min(opts, key=lambda k: opts[k])
        best_per_bin.append((nao, opts[it_best] < tj, it_best, tj, opts))
        log(f"tune eig: {kind} {size} nao={nao} jacobi {1e3 * tj:.1f} ms blocked {1e3 * tb:.1f} ms "
            + " ".join(f"first-{it} {1e3 * t:.1f} ms" for it, t in zip(hybrid, th)) + f" dE={dE:.1e}")
    best_per_bin.sort(key=lambda r: r[0])
    # one setting for every size from the cut on: the iteration count that
    # is best for the largest molecule, and the smallest cut from which that
    # count beats block Jacobi at every size measured
    cut, iters = 1 << 30, 1 << 30
    if best_per_bin and best_per_bin[-1][1]:
        it_l = best_per_bin[-1][2]
        for i in range(len(best_per_bin)):
            if all(opts[it_l] < tj for _, _, _, tj, opts in best_per_bin[i:]):
                cut, iters = best_per_bin[i][0], it_l
                break
    # 2. batch schedule on the 512-molecule workload
    mols = workloads.mixed_batch(512, seed=2024)
    best = None
    for lanes in (1, 2, 3):
        for cross in (0, 1):
            t, _ = _time_batch(mols, {"XTBAMD_JAC_LANES": lanes, "XTBAMD_JAC_CROSS": cross,
                                      "XTBAMD_EIG_BLK_MIN": 1 << 30}, steps)
            rows_batch.append((f"jac_lanes={lanes} jac_cross={cross}", t))
            log(f"tune batch: lanes {lanes} cross {cross}: {1e3 * t:.1f} ms")
            if best is None or t < best[0]:
                best = (t, lanes, cross)
    _, lanes, cross = best
    base = {"XTBAMD_JAC_LANES": lanes, "XTBAMD_JAC_CROSS": cross}
    tv, _ = _time_batch(mols, dict(base, XTBAMD_GEMM="valu", XTBAMD_EIG_BLK_MIN=1 << 30), steps)
    tm, _ = _time_batch(mols, dict(base, XTBAMD_GEMM="mfma", XTBAMD_EIG_BLK_MIN=1 << 30), steps)
    rows_batch += [("gemm=valu", tv), ("gemm=mfma", tm)]
    gemm = "valu" if tv < tm else "mfma"
    # the cutoff must not slow the batch down (its largest problems are below
    # it unless the blocked path wins there too)
    if cut < 1 << 30:
        tc, _ = _time_batch(mols, dict(base, XTBAMD_GEMM=gemm, XTBAMD_EIG_BLK_MIN=cut, XTBAMD_EIG_BLK_ITERS=iters),
                            steps)
        rows_batch.append((f"eig_blk_min={cut}", tc))
        if tc > min(tv, tm):
            rows_batch.append((f"eig_blk_min={cut} rejected for the batch", tc))
            cut = max(cut, 1 + max(workloads.ao_count(z) for z, _ in mols))
    _set({})
    os.environ.pop("XTBAMD_TUNING", None)
    stamp = time.strftime("%Y-%m-%d %H:%M:%S")
    with open(os.path.join(out, "tuning_gfx950.txt"), "w") as f:
        f.write(f"# written by python -m xtb_amd.tune ({stamp}); environment variables XTBAMD_* override\n")
        f.write(f"eig_blk_min {cut}\neig_blk_iters {iters}\njac_lanes {lanes}\njac_cross {cross}\ngemm {gemm}\n")
    with open(os.path.join(out, "autotune_gfx950.md"), "w") as f:
        f.write("# Autotuner measurements (one MI355X)\n\n")
        f.write("Command: `python -m xtb_amd.tune --out <dir>` (median of "
                f"{steps} timed runs after one warm-up per configuration).\n\n")
        f.write("## Eigensolver path per size (one molecule per batch, whole single point)\n\n")
        f.write("| molecule | AOs | block Jacobi ms | blocked, every iteration ms | "
                + " | ".join(f"blocked, first {it} iterations ms" for it in hybrid)
                + " | dE Jacobi vs blocked (Eh) |\n|---|---:|---:|---:|" + "---:|" * len(hybrid) + "---:|\n")
        for kind, size, nao, tj, tb, th, dE in rows_bins:
            name = f"(H2O){size}" if kind == "water" else f"C{size}H{2 * size + 2}"
            f.write(f"| {name} | {nao} | {1e3 * tj:.1f} | {1e3 * tb:.1f} | "
                    + " | ".join(f"{1e3 * t:.1f}" for t in th) + f" | {dE:.1e} |\n")
        f.write("\n## 512-molecule batch (ms per batched step)\n\n| configuration | ms |\n|---|---:|\n")
        for name, t in rows_batch:
            f.write(f"| {name} | {1e3 * t:.1f} |\n")
        f.write(f"\nChosen: eig_blk_min {cut}, eig_blk_iters {iters}, jac_lanes {lanes}, jac_cross {cross}, "
                f"gemm {gemm}.\n")
    return {"eig_blk_min": cut, "eig_blk_iters": iters, "jac_lanes": lanes, "jac_cross": cross, "gemm": gemm}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--steps", type=int, default=3)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    r = tune(a.steps, a.out, log=lambda s: print(s, flush=True))
    print(r, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
