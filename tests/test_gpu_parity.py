"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle and
the reference's golden vectors.

Tolerances: total energy 1e-6 Eh, charges 1e-5 e, gradient 1e-5 Eh/bohr
(the path's parity bar); golden values use the reference tests' own
tolerances.  Integer outputs (iteration counts, AO/shell dimensions) must match
exactly.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.test_oracle_golden import SP_CASES, _lookup

pytestmark = pytest.mark.gpu

TOL_E, TOL_Q, TOL_G = 1e-6, 1e-5, 1e-5


def _engine():
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return xtb_amd


def _cmp(r, o, tag):
    assert r["converged"], tag
    assert r["natoms"] == len(o["q"]) and r["nao"] == o["nao"] and r["nshells"] == o["nsh"], tag
    dE = abs(r["energy"] - o["e_total"])
    dq = np.max(np.abs(r["charges"] - o["q"]))
    dg = np.max(np.abs(r["gradient"] - o["gradient"]))
    print(f"{tag}: dE={dE:.2e} dq={dq:.2e} dg={dg:.2e} iters={r['iterations']}/{o['iterations']}", flush=True)
    assert r["iterations"] == o["iterations"], f"{tag} SCC iterations"
    assert dE <= TOL_E, f"{tag} energy {r['energy']} vs {o['e_total']}"
    assert dq <= TOL_Q, f"{tag} charges"
    assert dg <= TOL_G, f"{tag} gradient"
    for k in ("e_es", "e_aes", "e_axc", "e_disp", "e_rep"):
        assert abs(r[k] - o[k]) <= TOL_E, f"{tag} {k}: {r[k]} vs {o[k]}"
    assert abs(r["hl_gap_ev"] - o["hl_gap_ev"]) <= 1e-4, f"{tag} gap"
    assert np.max(np.abs(r["dipole"] - o["dipole"])) <= TOL_Q, f"{tag} dipole"


@pytest.mark.parametrize("case", SP_CASES)
def test_golden_single(golden, case):
    xa = _engine()
    g = golden[case]
    r = xa.singlepoint(g["numbers"], g["positions"], g["charge"], g["uhf"], g["accuracy"], g["etemp"],
                       g["maxiter"], g["guess"])
    o = po.singlepoint(g["numbers"], g["positions"], charge=g["charge"], uhf=g["uhf"], guess=g["guess"],
                       maxiter=g["maxiter"], accuracy=g["accuracy"], etemp=g["etemp"])
    _cmp(r, o, case)
    res = dict(r)
    res["e_total"] = r["energy"]
    res["q"] = r["charges"]
    assert r["ihomo"] == o["ihomo"]
    for key, (ref, tol) in g["expect"].items():
        val = _lookup(res, key)
        assert abs(val - ref) <= max(tol, 1e-7), f"{case} {key}: {val} vs {ref}"


SOLV_CASES = ["gbsa_water_api"] + [f"mindless{k:02d}_solv" for k in range(1, 11)] + ["feco5_gbsa", "feco5_alpb"]


@pytest.mark.parametrize("case", SOLV_CASES)
def test_solvation_golden_single(golden, case):
    """GBSA and GBSA + ALPB shape term in eleven solvents (Born radii, surface,
    H-bond terms on the device) against the oracle and the reference's values."""
    xa = _engine()
    g = golden[case]
    r = xa.singlepoint(g["numbers"], g["positions"], g["charge"], g["uhf"], g["accuracy"], g["etemp"],
                       g["maxiter"], g["guess"], solvent=g["solvent"], solv_kernel=g["solv_kernel"], alpb=g["alpb"])
    o = po.singlepoint(g["numbers"], g["positions"], charge=g["charge"], uhf=g["uhf"], guess=g["guess"],
                       maxiter=g["maxiter"], accuracy=g["accuracy"], etemp=g["etemp"], solvent=g["solvent"],
                       solv_kernel=g["solv_kernel"], alpb=g["alpb"], threads=4)
    _cmp(r, o, case)
    assert abs(r["e_solv"] - o["e_solv"]) <= TOL_E, f"{case} e_solv {r['e_solv']} vs {o['e_solv']}"
    assert r["iterations"] == o["iterations"]
    res = dict(r)
    res["e_total"] = r["energy"]
    for key, (ref, tol) in g["expect"].items():
        val = _lookup(res, key)
        assert abs(val - ref) <= max(tol, 1e-7), f"{case} {key}: {val} vs {ref}"


@pytest.mark.parametrize("kernel,alpb", [("p16", True), ("still", False)])
def test_solvation_clusters_vs_oracle(kernel, alpb):
    """Water clusters and an alkane in water (ALPB / P16 kernel and GBSA / Still
    kernel), one batch: against the oracle, and bitwise the alone results."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(12, 3), w.alkane(10, 4), w.water_cluster(32, 0)]
    b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).set_solvent("water", kernel, alpb).run()
    mix = b.results()
    b.close()
    for k, ((z, x), rm) in enumerate(zip(mols, mix)):
        o = po.singlepoint(z, x, solvent="water", solv_kernel=kernel, alpb=alpb, threads=8, lapack=True)
        _cmp(rm, o, f"{kernel} mol{k} n={len(z)}")
        assert abs(rm["e_solv"] - o["e_solv"]) <= TOL_E
        ra = xa.singlepoint(z, x, solvent="water", solv_kernel=kernel, alpb=alpb)
        _same(ra, rm, f"{kernel} alone mol{k}")


def test_golden_batch_same_alone_and_in_mix(golden):
    """All golden molecules in one batch: bitwise the alone results."""
    xa = _engine()
    cases = [golden[c] for c in SP_CASES if golden[c]["guess"] == "eeq"]
    mols = [(g["numbers"], g["positions"], g["charge"], g["uhf"]) for g in cases]
    b = xa.Batch(mols, maxiter=250).run()
    mix = b.results()
    b.close()
    for g, rm in zip(cases, mix):
        ra = xa.singlepoint(g["numbers"], g["positions"], g["charge"], g["uhf"])
        assert ra["energy"] == rm["energy"]
        assert np.array_equal(ra["gradient"], rm["gradient"])
        assert np.array_equal(ra["charges"], rm["charges"])
        assert ra["iterations"] == rm["iterations"]


def test_water32_cluster():
    """32 H2O on a jittered 4x4x2 grid (96 atoms, 192 AOs): block Jacobi path."""
    xa = _engine()
    from xtb_amd import workloads as w
    z, x = w.water_cluster(32, 0)
    r = xa.singlepoint(z, x)
    o = po.singlepoint(z, x)
    _cmp(r, o, "water32")


def test_alkane_c64():
    """all-trans C64H130 (194 atoms, 386 AOs)."""
    xa = _engine()
    from xtb_amd import workloads as w
    z, x = w.alkane(64, 1)
    r = xa.singlepoint(z, x)
    o = po.singlepoint(z, x, threads=4)
    _cmp(r, o, "alkane64")


def test_water96_device_wide_eeq():
    """96 H2O (288 atoms, 576 AOs): the 289-unknown EEQ system is factorised
    column by column across the device rather than in one workgroup."""
    xa = _engine()
    from xtb_amd import workloads as w
    z, x = w.water_cluster(96, 5)
    r = xa.singlepoint(z, x)
    o = po.singlepoint(z, x, lapack=True, threads=8)
    _cmp(r, o, "water96")
    assert r["iterations"] == o["iterations"]


def test_water700_against_oracle_fixture():
    """The 700-water cluster (2100 atoms, 4200 AOs) against the oracle's values
    committed in tests/golden/water700_oracle.json."""
    import json
    import os
    xa = _engine()
    from xtb_amd import workloads as w
    with open(os.path.join(os.path.dirname(__file__), "golden", "water700_oracle.json")) as f:
        fx = json.load(f)
    z, x = w.large_water_cluster(2024)
    b = xa.Batch([(z, x, 0.0, 0)]).run()
    r = b.result(0)
    b.close()
    g = np.asarray(r["gradient"])
    q = np.asarray(r["charges"])
    dE = abs(r["energy"] - fx["e_total"])
    print(f"water700: dE={dE:.2e} iters={r['iterations']}/{fx['iterations']}", flush=True)
    assert r["converged"] and r["nao"] == fx["nao"] == 4200
    assert r["iterations"] == fx["iterations"]
    assert dE <= TOL_E
    assert abs(np.linalg.norm(g) - fx["gnorm"]) <= TOL_G
    assert np.max(np.abs(q[:24] - fx["q_first"])) <= TOL_Q
    assert np.max(np.abs(g[:8] - np.asarray(fx["gradient_first"]))) <= TOL_G
    assert abs(np.abs(q).sum() - fx["q_abs_sum"]) <= TOL_Q * len(q)
    assert abs(np.abs(g).sum() - fx["gradient_abs_sum"]) <= TOL_G * g.size
    assert np.max(np.abs(np.asarray(r["dipole"]) - fx["dipole"])) <= 1e-4
    assert abs(r["hl_gap_ev"] - fx["hl_gap_ev"]) <= 1e-4


def test_mixed_batch_vs_oracle():
    """16 molecules of the batched workload's distribution, as one batch."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = w.mixed_batch(16, seed=7, lo=10, hi=120)
    b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
    for m, (z, x) in enumerate(mols):
        _cmp(b.result(m), po.singlepoint(z, x), f"mix{m}(n={len(z)})")
    b.close()


def test_c_api_program(tmp_path):
    """The xtb.h conformance program (propyne golden values, error log,
    restart, unsupported calls) against libxtb_amd.so."""
    import os
    import subprocess
    _engine()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "capi_gfn2"
    subprocess.run(["gcc", "-O1", "-std=c11", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c_api", "capi_gfn2.c"), "-L", os.path.join(root, "xtb_amd"),
                    "-lxtb_amd", "-Wl,-rpath," + os.path.join(root, "xtb_amd"), "-lm", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr


def test_restart_guess_reproduces_energy():
    """A second single point seeded with the converged charges and multipoles
    (the reference's results-object reuse) lands on the same energy."""
    xa = _engine()
    from xtb_amd import workloads as w
    z, x = w.water_cluster(8, 3)
    b = xa.Batch([(z, x, 0.0, 0)]).run()
    r1 = b.result(0)
    b.close()
    b2 = xa.Batch([(z, x, 0.0, 0)])
    b2.set_guess(0, r1["qsh"], r1["dipm"], r1["qp"])
    b2.run()
    r2 = b2.result(0)
    b2.close()
    assert r2["iterations"] < r1["iterations"]
    assert abs(r2["energy"] - r1["energy"]) < 1e-7
    assert np.max(np.abs(r2["gradient"] - r1["gradient"])) < 1e-5


def test_rerun_same_batch_repeats_full_single_point():
    """Running a batch again restarts every SCC from the initial guess:
    bitwise identical results and iteration counts."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = w.mixed_batch(6, seed=11, lo=10, hi=80)
    b = xa.Batch([(z, x, 0.0, 0) for z, x in mols])
    b.run()
    first = b.results()
    b.run()
    second = b.results()
    assert b.counters()[0] == max(r["iterations"] for r in second)
    b.close()
    for r1, r2 in zip(first, second):
        assert r1["iterations"] == r2["iterations"] > 1
        assert r1["energy"] == r2["energy"]
        assert np.array_equal(r1["gradient"], r2["gradient"])


def _same(ra, rb, tag):
    assert ra["iterations"] == rb["iterations"], tag
    assert ra["energy"] == rb["energy"], f"{tag}: {ra['energy']!r} vs {rb['energy']!r}"
    assert np.array_equal(ra["gradient"], rb["gradient"]), tag
    assert np.array_equal(ra["charges"], rb["charges"]), tag


def test_block_jacobi_batch_same_alone_and_in_mix():
    """Molecules above 64 AOs (block-Jacobi path) mixed with smaller and
    larger ones: bitwise the alone results."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(12, 5), w.alkane(20, 6), w.water_cluster(3, 7), w.alkane(45, 8), w.water_cluster(30, 9)]
    b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
    mix = b.results()
    b.close()
    for k, ((z, x), rm) in enumerate(zip(mols, mix)):
        _same(xa.singlepoint(z, x), rm, f"mol{k} nao={rm['nao']}")


def test_jacobi_lanes_do_not_change_results(monkeypatch):
    """The block-Jacobi rounds on one, two and three stream lanes (problems
    split between the streams, XCD-ordered work lists): bitwise the same
    results."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(12, 5), w.alkane(20, 6), w.alkane(45, 8), w.water_cluster(30, 9), w.alkane(33, 3)]
    runs = []
    for lanes in ("1", "2", "3"):
        monkeypatch.setenv("XTBAMD_JAC_LANES", lanes)
        b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
        runs.append(b.results())
        b.close()
    for k, rs in enumerate(zip(*runs)):
        for r in rs[1:]:
            _same(rs[0], r, f"mol{k} nao={rs[0]['nao']}")


def test_cross_rounds_agree_with_full_sweeps(monkeypatch):
    """Pivot problems that rotate only the pairs across their two blocks in
    most outer rounds (the default) against a full inner sweep in every round
    (XTBAMD_JAC_CROSS=0): both converge every diagonalisation to the same
    tolerance, so the results agree far inside the parity tolerances; both
    match the oracle."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(12, 5), w.alkane(20, 6), w.alkane(45, 8), w.water_cluster(30, 9), w.alkane(64, 3)]
    runs = []
    for cross in ("1", "0"):
        monkeypatch.setenv("XTBAMD_JAC_CROSS", cross)
        b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
        runs.append(b.results())
        b.close()
    for k, (rx, rf) in enumerate(zip(*runs)):
        tag = f"mol{k} nao={rx['nao']}"
        assert rx["converged"] and rf["converged"], tag
        assert abs(rx["energy"] - rf["energy"]) <= 1e-9, tag
        assert np.abs(rx["gradient"] - rf["gradient"]).max() <= 1e-8, tag
        assert np.abs(rx["charges"] - rf["charges"]).max() <= 1e-8, tag
    z, x = mols[-1]
    _cmp(runs[0][-1], po.singlepoint(z, x, threads=8, lapack=True), "cross rounds, C64H130")


def test_server_results_same_alone_and_in_mix():
    """Continuous batching: requests submitted while others iterate give the
    bitwise alone results; every request finishes."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = w.mixed_batch(14, seed=21, lo=10, hi=150)
    srv = xa.Server(hbm_gb=2.0, max_atoms=400, max_slots=64)
    ids = []
    waves = [mols[:5], mols[5:9], mols[9:]]
    steps = 0
    for wave in waves:
        for z, x in wave:
            ids.append(srv.submit(z, x))
        for _ in range(3):
            srv.step()
            steps += 1
    while srv.step() > 0:
        steps += 1
        assert steps < 2000
    done = set()
    for _ in range(3):
        done.update(srv.finished())
    assert done == set(ids)
    for rid, (z, x) in zip(ids, mols):
        rs = srv.result(rid)
        assert rs["converged"]
        _same(xa.singlepoint(z, x), rs, f"request {rid} n={len(z)}")
    srv.close()


def test_server_in_solvent_same_alone_and_in_mix():
    """Continuous batching in implicit solvent (ALPB water): staggered requests
    give the bitwise results of the same molecules alone in a batch."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = w.mixed_batch(8, seed=33, lo=10, hi=90)
    srv = xa.Server(hbm_gb=1.0, max_atoms=200, max_slots=32).set_solvent("water", "p16", alpb=True)
    ids = [srv.submit(z, x) for z, x in mols[:4]]
    for _ in range(2):
        srv.step()
    ids += [srv.submit(z, x) for z, x in mols[4:]]
    steps = 0
    while srv.step() > 0:
        steps += 1
        assert steps < 2000
    for rid, (z, x) in zip(ids, mols):
        rs = srv.result(rid)
        assert rs["converged"] and rs["e_solv"] < 0.0
        _same(xa.singlepoint(z, x, solvent="water", solv_kernel="p16", alpb=True), rs, f"request {rid}")
    srv.close()


def test_server_admits_by_budget():
    """A budget that holds only a few molecules at a time still serves a
    long queue, and the results agree with the oracle."""
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(k, 30 + k) for k in (4, 9, 6, 12, 5, 8)]
    srv = xa.Server(hbm_gb=0.02, max_atoms=40, max_slots=8)
    ids = [srv.submit(z, x) for z, x in mols]
    peak = 0
    while srv.step() > 0:
        peak = max(peak, srv.pending()[1])
    assert 1 <= peak < len(mols)
    got = set(srv.finished())
    assert got == set(ids)
    for rid, (z, x) in zip(ids, mols):
        _cmp(srv.result(rid), po.singlepoint(z, x), f"budget request {rid}")
    srv.close()
