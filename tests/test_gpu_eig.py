"""The blocked eigensolver (Householder tridiagonalisation, divide and
conquer, back-transformation; replaces dsytrd/dstedc/dormtr of the reference's
dsyevd, src/mctc/lapack/eigensolve.F90:218-246) against the block-Jacobi path
and the oracle: every problem of at least XTBAMD_EIG_BLK_MIN AOs takes the
blocked path.  The eigenvectors of the two solvers differ by rounding (and by
rotations inside degenerate subspaces), so the single points agree to the
SCC's own noise, far inside the parity bar."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

MOLS = None


def _mols():
    from xtb_amd import workloads as w
    return [w.water_cluster(12, 5), w.alkane(20, 6), w.water_cluster(3, 7), w.alkane(45, 8), w.water_cluster(30, 9),
            w.water_cluster(40, 10)]


def _run(monkeypatch, blk_min, iters=None):
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    monkeypatch.setenv("XTBAMD_TUNING", "off")
    if blk_min is None:
        monkeypatch.delenv("XTBAMD_EIG_BLK_MIN", raising=False)
    else:
        monkeypatch.setenv("XTBAMD_EIG_BLK_MIN", str(blk_min))
    if iters is None:
        monkeypatch.delenv("XTBAMD_EIG_BLK_ITERS", raising=False)
    else:
        monkeypatch.setenv("XTBAMD_EIG_BLK_ITERS", str(iters))
    mols = _mols()
    b = xtb_amd.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
    res = b.results()
    b.close()
    return mols, res


def _clustered(n, rng):
    """Symmetric matrix with bands of nearly degenerate and exactly
    degenerate eigenvalues (as in large water clusters)."""
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([np.full(n // 5, c) + 1e-7 * rng.standard_normal(n // 5) for c in (-30.0, -20.0, -10.0, 0.0, 5.0)])
    ev = np.concatenate([ev, rng.standard_normal(n - len(ev))])
    ev[:4] = ev[4]
    a = (q * ev) @ q.T
    return 0.5 * (a + a.T)


def test_eigh_blocked_against_lapack():
    """xtbamd_eigh against numpy's LAPACK eigensolver: eigenvalues, the
    eigen-residual and orthogonality at 1e-12 of the matrix norm; the
    tridiagonal form has A's eigenvalues."""
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    rng = np.random.default_rng(3)
    mats = []
    for n in (2, 3, 40, 64, 65, 100, 129, 257):
        x = rng.standard_normal((n, n))
        mats.append(x + x.T)
    mats += [_clustered(300, rng), _clustered(530, rng), np.diag(np.arange(70.0))]
    out, tri = xtb_amd.eigh(mats, tridiagonal=True)
    for k, (a, (lam, w), (d, e)) in enumerate(zip(mats, out, tri)):
        n = a.shape[0]
        scale = np.max(np.abs(np.linalg.eigvalsh(a)))
        ref = np.linalg.eigvalsh(a)
        t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        dt = np.max(np.abs(np.linalg.eigvalsh(t) - ref)) / scale
        dl = np.max(np.abs(lam - ref)) / scale
        res = np.max(np.abs(a @ w.T - w.T * lam)) / scale
        orth = np.max(np.abs(w @ w.T - np.eye(n)))
        print(f"eigh {k} n={n}: tri {dt:.1e} lam {dl:.1e} res {res:.1e} orth {orth:.1e}", flush=True)
        assert dt < 1e-12 and dl < 1e-12 and res < 1e-12 and orth < 1e-12, k


@pytest.mark.parametrize("blk_min,iters", [(65, None), (200, None), (65, 2)])
def test_blocked_eigensolver_matches_jacobi_and_oracle(monkeypatch, blk_min, iters):
    """Every iteration on the blocked path (iters None), or only the first two
    SCC iterations (the block-Jacobi lanes from the third on)."""
    mols, ref = _run(monkeypatch, None)
    _, got = _run(monkeypatch, blk_min, iters)
    for k, (r, g) in enumerate(zip(ref, got)):
        tag = f"blk_min={blk_min} iters={iters} mol{k} nao={g['nao']}"
        dE = abs(r["energy"] - g["energy"])
        dg = float(np.max(np.abs(r["gradient"] - g["gradient"])))
        dq = float(np.max(np.abs(r["charges"] - g["charges"])))
        de = float(np.max(np.abs(r["emo"] - g["emo"])))
        print(f"{tag}: dE={dE:.1e} dg={dg:.1e} dq={dq:.1e} demo={de:.1e} it={g['iterations']}", flush=True)
        assert g["converged"] and not g["failed"], tag
        assert g["iterations"] == r["iterations"], tag
        assert dE < 1e-9 and dg < 1e-8 and dq < 1e-8 and de < 1e-7, tag
    # the largest molecule against the oracle (parity bar)
    z, x = mols[-1]
    o = po.singlepoint(z, x)
    g = got[-1]
    assert abs(g["energy"] - o["e_total"]) < 1e-6
    assert np.max(np.abs(g["gradient"] - o["gradient"])) < 1e-5
    assert np.max(np.abs(g["charges"] - o["q"])) < 1e-5
    assert g["iterations"] == o["iterations"]



def test_blocked_solve_graph_replays_equal_eager(monkeypatch):
    """The blocked solve captured once as a HIP graph and replayed in every
    SCC iteration (eager Jacobi, mixing and host synchronisations between the
    replays), against the same launches issued eagerly (XTBAMD_GRAPH=0):
    bitwise the same results, on two runs of the same batch in a row."""
    import xtb_amd
    runs = []
    for graph in ("0", "1"):
        monkeypatch.setenv("XTBAMD_GRAPH", graph)
        monkeypatch.setenv("XTBAMD_TUNING", "off")
        monkeypatch.setenv("XTBAMD_EIG_BLK_MIN", "65")
        mols = _mols()
        b = xtb_amd.Batch([(z, x, 0.0, 0) for z, x in mols])
        b.run()
        first = b.results()
        b.run()
        runs.append((first, b.results()))
        b.close()
    (e1, e2), (g1, g2) = runs
    for k in range(len(e1)):
        for ra, rb in ((e1[k], g1[k]), (e2[k], g2[k]), (e1[k], g2[k])):
            assert ra["iterations"] == rb["iterations"], k
            assert ra["energy"] == rb["energy"], k
            assert np.array_equal(ra["gradient"], rb["gradient"]), k
            assert np.array_equal(ra["charges"], rb["charges"]), k
