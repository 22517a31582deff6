"""GPU tests of the eigensolver's layout options and of failure reporting.

* The block-Jacobi problems on padded working copies (row stride rounded up to
  16 doubles, the default) and on the packed matrices (XTBAMD_JAC_PAD=0) give
  the same results: the padding changes only addresses, never the arithmetic.
* A diagonalization that gives up (the sweep bound forced down to one sweep
  from the host with XTBAMD_JAC_MAX_SWEEPS=1; every kernel still ends
  normally) fails its molecule: result["failed"] is the eigensolver's code,
  singlepoint raises, and the other molecules of the batch are untouched.
* The server refuses a request that can never fit its budget at submit time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _engine():
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return xtb_amd


def _same(a, b, tag):
    assert a["iterations"] == b["iterations"], tag
    assert a["energy"] == b["energy"], f"{tag}: {a['energy']!r} vs {b['energy']!r}"
    for k in ("gradient", "charges", "qsh", "dipm", "qp", "wbo", "emo"):
        assert np.array_equal(a[k], b[k]), f"{tag} {k}"


def test_padded_and_packed_jacobi_agree(monkeypatch):
    xa = _engine()
    from xtb_amd import workloads as w
    mols = [w.water_cluster(12, 5), w.alkane(20, 6), w.water_cluster(3, 7), w.alkane(45, 8), w.water_cluster(30, 9)]
    runs = []
    for pad in ("1", "0"):
        monkeypatch.setenv("XTBAMD_JAC_PAD", pad)
        b = xa.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
        runs.append(b.results())
        b.close()
    for k, (rp, ru) in enumerate(zip(*runs)):
        tag = f"mol{k} nao={rp['nao']}"
        print(tag, rp["energy"], ru["energy"], flush=True)
        assert rp["converged"] and not rp["failed"], tag
        _same(rp, ru, tag)


def test_diagonalization_that_gives_up_fails_the_molecule(monkeypatch):
    xa = _engine()
    from xtb_amd import workloads as w
    from xtb_amd.engine import FAILURES
    small = w.water_cluster(3, 7)     # 18 AOs: one-workgroup solver
    big = w.water_cluster(12, 5)      # 72 AOs: block Jacobi
    monkeypatch.setenv("XTBAMD_JAC_MAX_SWEEPS", "1")
    b = xa.Batch([(z, x, 0.0, 0) for z, x in (small, big)]).run()
    res = b.results()
    b.close()
    for r, tag in zip(res, ("18 AOs", "72 AOs")):
        print(tag, r["failed"], r["iterations"], r["converged"], flush=True)
        assert r["failed"] == 3, tag
        assert not r["converged"], tag
    assert FAILURES[3].startswith("diagonalization")
    with pytest.raises(xa.SolverFailure):
        xa.singlepoint(*big)
    monkeypatch.delenv("XTBAMD_JAC_MAX_SWEEPS")
    # the default bound converges both
    b = xa.Batch([(z, x, 0.0, 0) for z, x in (small, big)]).run()
    for r in b.results():
        assert r["converged"] and r["failed"] == 0
    b.close()


def test_server_refuses_request_beyond_budget():
    xa = _engine()
    from xtb_amd import workloads as w
    srv = xa.Server(hbm_gb=0.02, max_atoms=40, max_slots=8)
    z, x = w.water_cluster(40, 3)     # 120 atoms: above max_atoms
    with pytest.raises(RuntimeError, match="HBM budget"):
        srv.submit(z, x)
    z, x = w.water_cluster(4, 3)
    rid = srv.submit(z, x)
    while srv.step() > 0:
        pass
    assert srv.finished() == [rid]
    assert srv.result(rid)["converged"]
    srv.close()
