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


def main(argv):
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
