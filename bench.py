"""Benchmark: GFN2-xTB SCC single points with gradient on one MI355X.

Workload: 512 independent neutral singlet molecules, each a seeded water
cluster or all-trans alkane (probability 0.5), sizes log-uniform in 10..200
atoms.  A step is one batched pass (setup, SCC to convergence, gradient) over
all 512 molecules with inputs resident on the device.  The plain path runs the
same molecules one at a time through the same FP64 engine.

Prints one JSON line (see the fields below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from xtb_amd import workloads  # noqa: E402
from xtb_amd.profiling import gpu_time_shares, traffic_table  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def _cpu_worker(args):
    z, x = args
    from oracle import pyoracle as po
    t = time.perf_counter()
    r = po.singlepoint(z, x, lapack=True, threads=1)
    return r["e_total"], time.perf_counter() - t


def cpu_baseline(mols, budget_s, procs):
    """The FP64 C restatement of the reference path (oracle), one molecule per
    process, over the first molecules of the workload until budget_s elapses."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")   # runs before anything touches the GPU
    t0 = time.perf_counter()
    energies, done, busy = [], 0, 0.0
    with ctx.Pool(procs) as pool:
        it = pool.imap(_cpu_worker, mols, chunksize=1)
        for e, dt in it:
            energies.append(e)
            busy += dt
            done += 1
            if time.perf_counter() - t0 > budget_s and done % procs == 0:
                pool.terminate()
                break
    wall = time.perf_counter() - t0
    return done, wall, energies, busy


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def serve(args, reqs, qps):
    """Serving workload: arrivals of seeded water clusters / alkanes of 20-400
    atoms (log-uniform), continuous batching under an HBM budget; Poisson
    arrivals at qps requests/s (qps = inf: every request at t = 0, the
    saturated capacity); latency = finish - arrival."""
    import xtb_amd
    rng = np.random.default_rng(args.seed + 1)
    if np.isinf(qps):
        arrivals = np.zeros(len(reqs))
    else:
        arrivals = np.cumsum(rng.exponential(1.0 / qps, len(reqs)))
    srv = xtb_amd.Server(hbm_gb=args.serve_hbm_gb, max_atoms=400, max_slots=1024)
    # warm-up: a few requests through the server before the clock starts
    for z, x in reqs[:4]:
        srv.submit(z, x)
    while srv.step() > 0:
        pass
    srv.finished()
    ids = {}
    fin = {}
    t0 = time.perf_counter()
    nsub = 0
    while len(fin) < len(reqs):
        now = time.perf_counter() - t0
        while nsub < len(reqs) and arrivals[nsub] <= now:
            z, x = reqs[nsub]
            ids[srv.submit(z, x)] = nsub
            nsub += 1
        left = srv.step()
        t = time.perf_counter() - t0
        for rid in srv.finished():
            fin[ids[rid]] = t
        if left == 0 and nsub < len(reqs):
            time.sleep(max(0.0, arrivals[nsub] - (time.perf_counter() - t0)))
    lat = np.array([fin[i] - arrivals[i] for i in range(len(reqs))])
    span = max(fin.values()) - arrivals[0]
    srv.close()
    return {"requests": len(reqs), "offered_qps": None if np.isinf(qps) else qps, "value": len(reqs) / span,
            "unit": "single-points/s", "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "hbm_gb": args.serve_hbm_gb}


def _one(xtb_amd, z, x, mixed=False):
    """Second of two runs of a one-molecule batch (the first allocates and
    loads the kernels); mixed: the FP32-MFMA products with FP64 residual."""
    if mixed:
        os.environ["XTBAMD_MIXED"] = "1"
    try:
        b = xtb_amd.Batch([(z, x, 0.0, 0)])
    finally:
        os.environ.pop("XTBAMD_MIXED", None)
    b.run()
    t = time.perf_counter()
    b.run()
    dt = time.perf_counter() - t
    r = b.result(0)
    b.close()
    return dt, r


def single_instances(xtb_amd, c1_oracle=None):
    """One molecule per batch: the 32-water cluster (192 AOs), C64H130 (386
    AOs; FP64 and mixed precision, errors against the FP64 oracle) and the
    700-water cluster (4200 AOs)."""
    cases = [("water32", workloads.water_cluster(32, 0)), ("alkane_c64h130", workloads.alkane(64, 1)),
             ("water700", workloads.large_water_cluster(2024))]
    out = {}
    for name, (z, x) in cases:
        dt, r = _one(xtb_amd, z, x)
        out[name] = {"atoms": int(len(z)), "aos": int(r["nao"]), "ms": 1e3 * dt, "single_points_per_s": 1.0 / dt,
                     "iterations": int(r["iterations"]), "energy": r["energy"], "converged": bool(r["converged"])}
        if name == "alkane_c64h130":
            dtm, rm = _one(xtb_amd, z, x, mixed=True)
            mix = {"precision": "Fock transforms and density products on FP32 matrix cores with FP64 residual; "
                                "eigensolver, mixing and energies FP64",
                   "ms": 1e3 * dtm, "energy": rm["energy"], "iterations": int(rm["iterations"]),
                   "dE_vs_fp64_path": abs(rm["energy"] - r["energy"]),
                   "max_dgrad_vs_fp64_path": float(np.max(np.abs(rm["gradient"] - r["gradient"])))}
            if c1_oracle is not None:
                mix["dE_vs_fp64_oracle"] = abs(rm["energy"] - c1_oracle)
                out[name]["dE_vs_fp64_oracle"] = abs(r["energy"] - c1_oracle)
            out[name]["mixed_precision"] = mix
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nmol", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--plain", type=int, default=512, help="molecules timed on the plain one-at-a-time path")
    ap.add_argument("--plain-alloc", type=int, default=64,
                    help="molecules timed one at a time through a new (allocating) batch per call")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-procs", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-loads", type=float, nargs="*", default=[0.5, 0.9],
                    help="Poisson offered loads as fractions of the saturated serving capacity")
    ap.add_argument("--serve-hbm-gb", type=float, default=64.0)
    ap.add_argument("--no-singles", action="store_true", help="skip the one-molecule timings (C0, C1, C4 sizes)")
    ap.add_argument("--no-solvent", action="store_true", help="skip the batched step in implicit water")
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="run one plain device copy of this size first (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    mols = workloads.mixed_batch(args.nmol, seed=args.seed)
    natoms = np.array([len(z) for z, _ in mols])
    naos = np.array([workloads.ao_count(z) for z, _ in mols])

    cpu = None
    c1_oracle = None
    if not args.no_cpu:
        # the FP64 oracle energy of C64H130 (the precision comparison's reference)
        import multiprocessing as mp
        with mp.get_context("fork").Pool(1) as pool:
            c1_oracle = pool.map(_cpu_worker, [workloads.alkane(64, 1)])[0][0]
        done, wall, ecpu, busy = cpu_baseline(mols, args.cpu_budget, args.cpu_procs)
        cpu = {"value": done / wall, "unit": "single-points/s", "cores": args.cpu_procs, "kind": "port",
               "single_thread_value": done / busy, "cpu_model": cpu_model(),
               "sample": f"first {done} molecules of the workload, one single-threaded process per core, FP64 C "
                         f"oracle with MKL dsygvd/dgemm when loadable, {wall:.1f} s wall; single_thread_value = "
                         f"molecules / summed per-molecule time"}

    import xtb_amd  # first GPU use
    batch_in = [(z, x, 0.0, 0) for z, x in mols]
    if args.copy_probe_mb > 0:
        xtb_amd.library().xtbamd_copy_probe(int(args.copy_probe_mb * 1048576))

    # plain path: one molecule at a time, synchronous, through the
    # single-molecule entry point that allocates nothing (the library's
    # per-device scratch area), so no call pays for device allocation
    lat = []
    eplain = []
    for m in range(min(args.plain, len(mols))):
        t = time.perf_counter()
        r = xtb_amd.singlepoint_noalloc(mols[m][0], mols[m][1])
        lat.append(time.perf_counter() - t)
        eplain.append(r["energy"])
    lat = np.array(lat)
    # the same through a new batch per molecule (device allocation and
    # release on every call), on the first molecules
    lat_alloc = []
    for m in range(min(args.plain_alloc, len(mols))):
        t = time.perf_counter()
        xtb_amd.singlepoint(mols[m][0], mols[m][1])
        lat_alloc.append(time.perf_counter() - t)

    # batched path
    b = xtb_amd.Batch(batch_in)
    for _ in range(args.warmup):
        b.run()
    times = []
    # HIP events around the GEMM launches and a sample of the Jacobi rounds'
    # pivot and tile-update launches, on the streams they run on
    b.profile("gemm", "jacobi", "apply")
    for _ in range(args.steps):
        t = time.perf_counter()
        b.run()
        times.append(time.perf_counter() - t)
    prof = {k: b.profile_result(k) for k in ("gemm", "jacobi", "apply")}
    spans = {"jacobi": b.profile_span("jacobi"), "apply": b.profile_span("apply")}
    res = b.results()
    # every timed step is a complete single point of every molecule: the
    # results of the last step equal the plain path's and iterate from scratch
    steps_run, jac_rounds = b.counters()
    assert steps_run == max(r["iterations"] for r in res) > 1, "a timed step skipped the SCC"

    ebatch = np.array([r["energy"] for r in res])
    conv = all(r["converged"] for r in res)
    iters = np.array([r["iterations"] for r in res])
    ms = 1e3 * float(np.mean(times))

    dE_plain = float(np.max(np.abs(ebatch[:len(eplain)] - np.array(eplain)))) if eplain else None
    dE_cpu = None
    if cpu is not None:
        ec = np.array(ecpu)
        dE_cpu = float(np.max(np.abs(ebatch[:len(ec)] - ec)))

    out = {
        "metric": "GFN2 SCC single-points/s + p50/p99 latency, batched vs plain, dE<=1e-6 Eh",
        "value": len(mols) / (ms * 1e-3),
        "unit": "single-points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(mols)} molecules (seeded water clusters / all-trans alkanes, "
                               f"10-200 atoms log-uniform, seed {args.seed}), energy+gradient, batched",
                   "atoms": int(natoms.sum()), "aos": int(naos.sum()), "max_ao": int(naos.max()),
                   "parallelism": "single GPU"},
        "batched": {"mean_ms": ms, "p50_ms": 1e3 * float(np.percentile(times, 50)),
                    "p99_ms": 1e3 * float(np.percentile(times, 99)), "converged": conv,
                    "scc_iterations_mean": float(iters.mean())},
        "plain": {"molecules": len(lat), "value": len(lat) / float(lat.sum()) if len(lat) else None,
                  "p50_ms": 1e3 * float(np.percentile(lat, 50)) if len(lat) else None,
                  "p99_ms": 1e3 * float(np.percentile(lat, 99)) if len(lat) else None,
                  "path": "one molecule per call, xtbamd_singlepoint_noalloc (no device allocation)"},
        "plain_allocating": {"molecules": len(lat_alloc),
                             "value": len(lat_alloc) / float(np.sum(lat_alloc)) if lat_alloc else None,
                             "path": "one molecule per call, new batch (allocation) per call"},
        "parity": {"max_dE_batched_vs_plain": dE_plain, "max_dE_vs_oracle": dE_cpu},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    ql = os.environ.get("XTBAMD_EIG") == "ql"
    lds = os.environ.get("XTBAMD_PIVOT_KERNEL") == "lds"
    names = {"gemm": "k_gemm_mfma (batched FP64 GEMM on the matrix cores)",
             "jacobi": ql and "k_tred_ql (tridiagonalisation + QL eigensolver, one workgroup per molecule)" or
             lds and "k_jac_pairsP (32x32 pivot problems of the block-Jacobi eigensolver, in LDS)" or
             "k_jac_pairsR (register-resident Jacobi sweep of each 32x32 pivot problem, one wave)",
             "apply": "k_jac_apply_R (A and W tile updates of a block-Jacobi round, FP64 MFMA, one wave per "
                      "32x32 tile)"}
    short = {"gemm": "k_gemm_mfma", "jacobi": ql and "k_tred_ql" or lds and "k_jac_pairsP" or "k_jac_pairsR",
             "apply": "k_jac_apply_R"}
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs of this
    # workload (copy-calibrated) and the GPU-time shares of its kernel-trace
    # run; only valid for the default workload and schedule
    table, shares = {}, {}
    default_run = args.nmol == 512 and args.seed == 2024 and not any(k.startswith("XTBAMD_") for k in os.environ)
    if default_run:
        table = traffic_table(os.path.join(ROOT, "profiles", "batched512_hbm_traffic.md"))
        shares = gpu_time_shares(os.path.join(ROOT, "profiles", "batched512_kernel_stats.md"))
    lines = {}
    for k, (n, tms, work) in prof.items():
        if not n:
            continue
        sec = tms * 1e-3
        events_us = 1e3 * tms / n
        if spans.get(k, 0.0) > 0.0:
            # the Jacobi launches: duration from the sampled launches'
            # execution spans on the device wall clock (first wave's start to
            # last wave's end, as a kernel trace measures it); the HIP-event
            # time beside it also holds the stream's dispatch gap before each
            # launch in the three-lane schedule
            sec = spans[k] * 1e-3
            tms = spans[k]
        if k == "apply":
            ach = work / sec / 1e9
            line = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                    "bytes_per_launch": work / n,
                    "work": "compulsory HBM bytes: each element of A and W a rotation touches read once and "
                            "written once (A through one of its mirrored tiles), counted on the device"}
        else:
            ach = work / sec / 1e12
            line = {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / FP64_PEAK_TFLOPS, "flops_per_launch": work / n}
            if k == "jacobi":
                line["note"] = ("FP64 VALU work (the FP64 vector and matrix peaks are both 78.6 TFLOP/s); bound by "
                                "the latency of its dependent inner rounds")
        if k in spans:
            line["avg_us_events"] = events_us
            line["duration"] = "device wall clock per sampled launch (kernel execution span)"
        line.update({"kernel": names[k], "traffic": table.get(short[k]), "launches": n, "avg_us": 1e3 * tms / n,
                     "share_of_step_events": tms / (args.steps * ms),
                     "share_of_gpu_time_rocprof": shares.get(short[k])})
        lines[k] = line
    if lines:
        # the dominant kernel: largest share of GPU time in the committed
        # kernel-trace run of this workload, else of the event-timed time
        if shares and all(lines[k]["share_of_gpu_time_rocprof"] is not None for k in lines):
            dom = max(lines, key=lambda k: lines[k]["share_of_gpu_time_rocprof"])
        else:
            dom = max(lines, key=lambda k: lines[k]["share_of_step_events"])
        out["roofline"] = lines[dom]
        out["roofline_other"] = [v for k, v in lines.items() if k != dom]
    # the whole step against the FP64 peak: F_SP = it (22/3) N^3 + (13/3) N^3
    # flops per single point (it = its SCC iterations, N = its AOs)
    nao_r = np.array([r["nao"] for r in res], dtype=np.float64)
    fsp = float(np.sum(iters * (22.0 / 3.0) * nao_r ** 3 + (13.0 / 3.0) * nao_r ** 3))
    out["step_roofline"] = {"flops_per_step": fsp, "achieved": fsp / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": fsp / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                            "model": "it * 22/3 N^3 + 13/3 N^3 FP64 flops per single point"}
    out["counters"] = {"scc_steps": steps_run, "jacobi_rounds": jac_rounds}
    if out["plain"]["value"]:
        out["speedup_batched_vs_plain"] = out["value"] / out["plain"]["value"]
    if not args.no_solvent:
        # the same step in implicit water (ALPB: P16 kernel + shape term)
        b.set_solvent("water", "p16", alpb=True)
        b.run()
        t = time.perf_counter()
        b.run()
        dt = time.perf_counter() - t
        rs = b.results()
        out["solvated"] = {"solvent": "water (ALPB)", "ms_per_step": 1e3 * dt, "value": len(mols) / dt,
                           "unit": "single-points/s", "converged": all(r["converged"] for r in rs),
                           "scc_iterations_mean": float(np.mean([r["iterations"] for r in rs])),
                           "e_solv_mean": float(np.mean([r["e_solv"] for r in rs]))}
    b.close()
    if args.serve_requests > 0:
        # C3: saturated capacity of the serving loop on its 20-400-atom mix,
        # then Poisson arrivals at fractions of it
        reqs = workloads.mixed_batch(args.serve_requests, seed=args.seed + 2, lo=20, hi=400)
        sat = serve(args, reqs, float("inf"))
        out["serving"] = {"workload": "seeded water clusters / alkanes, 20-400 atoms log-uniform, energy+gradient, "
                                      "continuous batching under an HBM budget",
                          "saturated": sat,
                          "c2_capacity_single_points_per_s": out["value"]}
        for f in args.serve_loads:
            out["serving"][f"poisson_{f:g}x"] = serve(args, reqs, f * sat["value"])
    if not args.no_singles:
        out["single_instances"] = single_instances(xtb_amd, c1_oracle)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
