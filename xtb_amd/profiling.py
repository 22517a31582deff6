"""HBM traffic per kernel launch from two rocprofv3 counter runs.

rocprofv3 collects one counter group per run, so the bytes a kernel moves come
from two runs of the same command: one with ``--pmc FETCH_SIZE`` and one with
``--pmc WRITE_SIZE`` (both in KiB per dispatch).  The raw counters are scaled
by what they read for a plain device copy of known size in the same run
(``bench.py --copy-probe-mb``: one fill kernel and one copy kernel of that many
MiB), so the table states bytes, not counter units.

    python -m xtb_amd.profiling FETCH_DIR WRITE_DIR COPY_MIB > table.md

prints a markdown table; ``traffic_table`` reads such a table back.

    python -m xtb_amd.profiling stats RUN_DIR

prints the per-kernel table of a ``--kernel-trace --stats`` run.

    python -m xtb_amd.profiling single NAME

builds the named one-molecule workload (``water32``, ``alkane_c64h130``,
``water700``), runs it once as a batch of one and prints the wall time, the
energy, the SCC iterations and the Jacobi round counters: the program a
kernel trace of one single point is taken over.

    python -m xtb_amd.profiling single-table RUN_DIR NAO

prints the per-kernel table of such a trace with the algorithmic rates of the
kernels whose work has a closed count for an NAO-orbital molecule, and the
spans of each blocked eigensolve (tridiagonalisation, divide and conquer,
back-transformation).
"""
import csv
import glob
import os
import sys
from collections import defaultdict

COPY_KERNEL = "k_copy"


def kernel_short_name(full):
    """'xtbamd::(anonymous namespace)::k_copy(double const*, ...)' -> 'k_copy'."""
    s = full.replace("(anonymous namespace)", "")
    s = s.split("(")[0].split("<")[0]
    return s.split("::")[-1].strip()


def _load(run_dir, counter):
    paths = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    per = defaultdict(list)
    for p in paths:
        with open(p) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                per[kernel_short_name(r["Kernel_Name"])].append((float(r["Counter_Value"]), dur))
    return per


def traffic(fetch_dir, write_dir, copy_mib):
    """Per-kernel launches, raw KiB per launch and calibrated bytes per launch."""
    fetch = _load(fetch_dir, "FETCH_SIZE")
    write = _load(write_dir, "WRITE_SIZE")
    if COPY_KERNEL not in fetch or COPY_KERNEL not in write:
        raise RuntimeError("the calibration copy kernel is missing from the counter runs")
    copy_kib = copy_mib * 1024.0
    f_scale = copy_kib / fetch[COPY_KERNEL][0][0]
    w_scale = copy_kib / write[COPY_KERNEL][0][0]
    rows = []
    for k in fetch:
        if k not in write:
            continue
        f = sum(v for v, _ in fetch[k]) / len(fetch[k])
        w = sum(v for v, _ in write[k]) / len(write[k])
        t = sum(d for _, d in write[k]) / len(write[k])
        rows.append({"kernel": k, "launches": len(fetch[k]), "fetch_kib": f, "write_kib": w,
                     "bytes": (f * f_scale + w * w_scale) * 1024.0, "avg_us": t / 1e3,
                     "total_ms": t * len(write[k]) / 1e6})
    rows.sort(key=lambda r: -r["total_ms"])
    return rows, f_scale, w_scale


def markdown(rows, f_scale, w_scale, copy_mib):
    out = [f"Calibration: a {copy_mib:g} MiB device copy reads FETCH_SIZE x {f_scale:.4f} and "
           f"WRITE_SIZE x {w_scale:.4f} of the true bytes; the byte column applies these factors.",
           "",
           "| kernel | launches | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | HBM bytes/launch | "
           "avg us (counter run) | total ms (counter run) |",
           "|---|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| {r['kernel']} | {r['launches']} | {r['fetch_kib']:.1f} | {r['write_kib']:.1f} | "
                   f"{r['bytes']:.6g} | {r['avg_us']:.1f} | {r['total_ms']:.2f} |")
    return "\n".join(out)


def traffic_table(path):
    """kernel -> HBM bytes per launch, read from a table written by ``markdown``."""
    table = {}
    if not os.path.exists(path):
        return table
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0].startswith("k_"):
                try:
                    table[cells[0]] = float(cells[4])
                except ValueError:
                    pass
    return table


def gpu_time_shares(path):
    """kernel short name -> % of GPU time, read from a table written by
    ``stats_table`` (instantiations of one kernel template summed)."""
    shares = {}
    if not os.path.exists(path):
        return shares
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0].startswith("`"):
                k = kernel_short_name(cells[0].strip("`"))
                try:
                    shares[k] = shares.get(k, 0.0) + float(cells[4])
                except ValueError:
                    pass
    return shares


def stats_table(run_dir, top=24):
    """Markdown table of a ``rocprofv3 --kernel-trace --stats`` run (kernel_stats.csv)."""
    paths = glob.glob(os.path.join(run_dir, "**", "*kernel_stats.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError(f"no kernel_stats.csv under {run_dir}")
    with open(paths[0]) as f:
        rows = list(csv.DictReader(f))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["| kernel | calls | total ms | avg us | % of GPU time |", "|---|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
        t = float(r["TotalDurationNs"])
        out.append(f"| `{name}` | {r['Calls']} | {t / 1e6:.2f} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{100.0 * t / total:.1f} |")
    out.append(f"| total | | {total / 1e6:.2f} | | 100.0 |")
    return "\n".join(out)


def _trace_rows(run_dir):
    paths = glob.glob(os.path.join(run_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not paths:
        raise FileNotFoundError(f"no kernel_trace.csv under {run_dir}")
    with open(paths[0]) as f:
        rows = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                 r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
                 .replace("xtbamd::", "")) for r in csv.DictReader(f)]
    rows.sort()
    return rows


def _solve_phase(name):
    """Phase of a blocked eigensolve a kernel belongs to (None: not part of one)."""
    if name in ("k_trd_col", "k_trd_row", "k_bgemm<true, false>"):
        return "tridiagonalisation"
    if name.startswith("k_dc_"):
        return "divide and conquer"
    if name.startswith("k_bgemm") or name == "k_trd_tblock":
        return "back-transformation"
    return None


def single_table(run_dir, n, top=22):
    """Kernel table of one single point of an n-orbital molecule with the
    algorithmic rates of the kernels whose work has a closed count:
    the full-size products (2 n^3 flops per full product), the tridiagonalisation's
    column kernel (8 (n-2-j)^2 bytes of trailing rows for column j), its panel
    updates (2 (n-j0-32)^2 * 64 flops per 32-column panel) and the Jacobi tile
    updates (28 n^2 bytes per round if every pair rotates: an upper bound);
    then the span of each phase of every blocked eigensolve."""
    rows = _trace_rows(run_dir)
    agg = {}
    for s, e, name in rows:
        a = agg.setdefault(name, [0, 0])
        a[0] += 1
        a[1] += e - s
    total = sum(v[1] for v in agg.values())
    trd_bytes = sum(8.0 * (n - 2 - j) ** 2 for j in range(n - 1))
    upd_flops = sum(2.0 * (n - j0 - 32) ** 2 * 64 for j0 in range(0, n - 1, 32))
    nsolve = agg.get("k_trd_col", [0, 0])[0] // max(1, n - 1)
    out = ["| kernel | calls | total ms | avg us | % of GPU time | algorithmic rate |",
           "|---|---:|---:|---:|---:|---|"]
    for name, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        sec = t * 1e-9
        rate = ""
        if name == "k_trd_col" and nsolve:
            rate = (f"{nsolve * trd_bytes / sec / 1e12:.2f} TB/s (trailing rows, "
                    f"{trd_bytes / (n - 1) / 1e6:.0f} MB per column on average)")
        elif name == "k_gemm_mfma<false, false, false, false>":
            # full products only: the symmetric ones form half the tiles (about
            # half the time) and launches masked off for every molecule take a
            # few us; a full product is one within 25% of the longest launch
            durs = [e - s for s, e, nm in rows if nm == name]
            full = [d for d in durs if d >= 0.75 * max(durs)]
            r = len(full) * 2.0 * n ** 3 / (sum(full) * 1e-9) / 1e12
            rate = (f"{r:.1f} TFLOP/s FP64 MFMA over its {len(full)} full products (2N^3 each), "
                    f"{100 * r / 78.6:.0f}% of 78.6")
        elif name == "k_bgemm<true, false>" and nsolve:
            r = nsolve * upd_flops / sec / 1e12
            rate = f"{r:.1f} TFLOP/s FP64 MFMA (panel updates), {100 * r / 78.6:.0f}% of 78.6"
        elif name == "k_jac_apply_R":
            per = 28.0 * n * n
            rate = f"<= {c * per / sec / 1e12:.2f} TB/s ({per / 1e6:.0f} MB per round if every pair rotates)"
        out.append(f"| `{name}` | {c} | {t / 1e6:.2f} | {t / c / 1e3:.1f} | {100.0 * t / total:.1f} | {rate} |")
    out.append(f"| total | | {total / 1e6:.2f} | | 100.0 | |")
    # blocked eigensolves: a new one starts where the tridiagonalisation follows another phase
    solves, last = [], None
    for s, e, name in rows:
        ph = _solve_phase(name)
        if ph is None:
            continue
        if not solves or (ph == "tridiagonalisation" and last != ph):
            solves.append({})
        span = solves[-1].setdefault(ph, [s, e])
        span[0], span[1] = min(span[0], s), max(span[1], e)
        last = ph
    out.append("")
    out.append("| blocked eigensolve | total ms | tridiagonalisation ms | divide and conquer ms | back-transformation ms |")
    out.append("|---:|---:|---:|---:|---:|")
    for i, sv in enumerate(solves):
        t0 = min(v[0] for v in sv.values())
        t1 = max(v[1] for v in sv.values())
        trd = sv.get("tridiagonalisation", [0, 0])
        dc = sv.get("divide and conquer", [0, 0])
        # the merges of divide and conquer run products too: the back-transformation is what follows its last kernel
        bt = (t1 - dc[1]) if dc[1] else 0
        out.append(f"| {i} | {(t1 - t0) / 1e6:.1f} | {(trd[1] - trd[0]) / 1e6:.1f} | {(dc[1] - dc[0]) / 1e6:.1f} | "
                   f"{bt / 1e6:.1f} |")
    return "\n".join(out)


SINGLES = {"water32": lambda w: w.water_cluster(32, 0), "alkane_c64h130": lambda w: w.alkane(64, 1),
           "water700": lambda w: w.large_water_cluster(2024)}


def run_single(name):
    """One run of a one-molecule batch of a named workload (the first call in
    the process: it allocates and loads the kernels)."""
    import time

    import xtb_amd
    from xtb_amd import workloads
    z, x = SINGLES[name](workloads)
    b = xtb_amd.Batch([(z, x, 0.0, 0)])
    t = time.perf_counter()
    b.run()
    dt = time.perf_counter() - t
    r = b.result(0)
    print(f"single {name}: {dt:.3f} s energy {r['energy']:.10f} iterations {r['iterations']} "
          f"counters {b.counters()}", flush=True)
    b.close()


def main(argv):
    if len(argv) == 2 and argv[0] == "single" and argv[1] in SINGLES:
        run_single(argv[1])
        return 0
    if len(argv) == 3 and argv[0] == "single-table":
        print(single_table(argv[1], int(argv[2])))
        return 0
    if len(argv) == 2 and argv[0] == "stats":
        print(stats_table(argv[1]))
        return 0
    if len(argv) != 3:
        print(__doc__, file=sys.stderr)
        return 2
    copy_mib = float(argv[2])
    rows, fs, ws = traffic(argv[0], argv[1], copy_mib)
    print(markdown(rows, fs, ws, copy_mib))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
