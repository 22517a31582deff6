"""GPU parity at the benchmark workloads' own sizes, and of the initial guess.

* The full 512-molecule batched workload (seeded water clusters / all-trans
  alkanes, 10-200 atoms, 20-392 AOs) as one batch: every molecule's energy,
  charges, gradient and SCC iteration count against the oracle's values in
  tests/golden/batch512_oracle.npz (tests/golden/make_batch_fixtures.py).
* Continuous batching over the serving mix (20-400 atoms, up to 800 AOs):
  requests submitted in waves while others iterate; each result bitwise the
  same as the molecule alone, and against the oracle's values in
  tests/golden/serve_oracle.npz.
* The device EEQ guess (one workgroup per molecule below 256 unknowns, column
  by column across the device above) against the oracle's om_eeq_guess and
  the reference's EEQ golden value for water (test/unit/test_eeq.f90:43-136).

Tolerances: energy 1e-6 Eh, charges 1e-5 e, gradient 1e-5 Eh/bohr; iteration
counts exact.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL_E, TOL_Q, TOL_G = 1e-6, 1e-5, 1e-5


def _engine():
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return xtb_amd


def _fixture(name):
    fx = np.load(os.path.join(HERE, "golden", name))
    off = np.concatenate([[0], np.cumsum(fx["natoms"])])
    return fx, off


def _check_vs_fixture(r, fx, off, m, tag):
    a, b = off[m], off[m + 1]
    q = fx["charges"][a:b]
    g = fx["gradient"][3 * a:3 * b].reshape(-1, 3)
    dE = abs(r["energy"] - fx["energy"][m])
    dq = float(np.max(np.abs(r["charges"] - q)))
    dg = float(np.max(np.abs(r["gradient"] - g)))
    assert r["converged"] and fx["converged"][m], tag
    assert r["natoms"] == fx["natoms"][m], tag
    assert r["iterations"] == fx["iterations"][m], f"{tag} iterations {r['iterations']} vs {fx['iterations'][m]}"
    assert dE <= TOL_E, f"{tag} dE={dE:.2e}"
    assert dq <= TOL_Q, f"{tag} dq={dq:.2e}"
    assert dg <= TOL_G, f"{tag} dg={dg:.2e}"
    assert abs(r["hl_gap_ev"] - fx["hl_gap_ev"][m]) <= 1e-4, tag
    return dE, dq, dg


def test_batch512_workload_vs_oracle_fixture():
    xa = _engine()
    from xtb_amd import workloads as w
    fx, off = _fixture("batch512_oracle.npz")
    mols = w.mixed_batch(len(fx["natoms"]), seed=int(fx["seed"]), lo=int(fx["lo"]), hi=int(fx["hi"]))
    assert [len(z) for z, _ in mols] == fx["natoms"].tolist()
    b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
    worst = np.zeros(3)
    for m in range(len(mols)):
        worst = np.maximum(worst, _check_vs_fixture(b.result(m), fx, off, m, f"mol{m} n={len(mols[m][0])}"))
    b.close()
    print(f"512 molecules: max dE={worst[0]:.2e} dq={worst[1]:.2e} dg={worst[2]:.2e}", flush=True)


def test_serving_mix_20_400_atoms_alone_and_vs_oracle():
    xa = _engine()
    from xtb_amd import workloads as w
    fx, off = _fixture("serve_oracle.npz")
    mols = w.mixed_batch(len(fx["natoms"]), seed=int(fx["seed"]), lo=int(fx["lo"]), hi=int(fx["hi"]))
    assert max(fx["natoms"]) > 300
    srv = xa.Server(hbm_gb=4.0, max_atoms=400, max_slots=64)
    ids = []
    steps = 0
    for wave in (mols[:8], mols[8:16], mols[16:]):
        for z, x in wave:
            ids.append(srv.submit(z, x))
        for _ in range(4):
            srv.step()
            steps += 1
    while srv.step() > 0:
        steps += 1
        assert steps < 5000
    got = set()
    for _ in range(3):
        got.update(srv.finished())
    assert got == set(ids)
    worst = np.zeros(3)
    for m, (rid, (z, x)) in enumerate(zip(ids, mols)):
        rs = srv.result(rid)
        tag = f"request {rid} n={len(z)}"
        worst = np.maximum(worst, _check_vs_fixture(rs, fx, off, m, tag))
        ra = xa.singlepoint(z, x)
        assert ra["iterations"] == rs["iterations"] and ra["energy"] == rs["energy"], tag
        for k in ("gradient", "charges", "qsh", "dipm", "qp", "wbo", "emo"):
            assert np.array_equal(ra[k], rs[k]), f"{tag} {k}"
    srv.close()
    print(f"serving mix: max dE={worst[0]:.2e} dq={worst[1]:.2e} dg={worst[2]:.2e}", flush=True)


def test_device_eeq_guess(golden):
    xa = _engine()
    from xtb_amd import workloads as w
    g = golden["eeq_water"]
    cases = [("eeq_water", (np.array(g["numbers"]), np.array(g["positions"])), 0.0),
             ("water32", w.water_cluster(32, 0), 0.0), ("alkane20", w.alkane(20, 6), 0.0),
             ("water96 (device-wide LU)", w.water_cluster(96, 5), 0.0),
             ("co_cnx6", (np.array(golden["co_cnx6"]["numbers"]), np.array(golden["co_cnx6"]["positions"])), -3.0),
             ("mindless04", (np.array(golden["mindless04"]["numbers"]), np.array(golden["mindless04"]["positions"])),
              0.0)]
    b = xa.Batch([(z, x, c, 0) for _, (z, x), c in cases])
    b.setup()
    for m, (tag, (z, x), c) in enumerate(cases):
        qd = b.charges(m)
        qo = po.eeq(z, x, charge=c)
        d = float(np.max(np.abs(qd - qo)))
        print(f"{tag}: max |q_device - q_oracle| = {d:.2e}", flush=True)
        assert d <= 1e-10, tag
        assert abs(qd.sum() - c) <= 1e-10, tag
    ref, tol = g["expect"]["q[0]"]
    assert abs(b.charges(0)[0] - ref) <= tol
    b.close()
