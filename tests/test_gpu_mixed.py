"""Mixed precision (XTBAMD_MIXED=1): the Fock transforms and the density
products on the FP32 matrix cores with an FP64 residual correction, the
eigensolver, mixing and energy sums in FP64.  The single points of C64H130
(the mid-size molecule of the precision comparison) and of a water cluster
against the FP64 path and the FP64 oracle, at the path's parity bar."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _run(monkeypatch, mixed, mols):
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    if mixed:
        monkeypatch.setenv("XTBAMD_MIXED", "1")
    else:
        monkeypatch.delenv("XTBAMD_MIXED", raising=False)
    b = xtb_amd.Batch([(z, x, 0.0, 0) for z, x in mols]).run()
    r = b.results()
    b.close()
    return r


def test_mixed_precision_against_fp64_and_oracle(monkeypatch):
    from xtb_amd import workloads as w
    mols = [w.alkane(64, 1), w.water_cluster(32, 0)]
    ref = _run(monkeypatch, False, mols)
    got = _run(monkeypatch, True, mols)
    for (z, x), r, g in zip(mols, ref, got):
        o = po.singlepoint(z, x)
        dE = abs(g["energy"] - r["energy"])
        dEo = abs(g["energy"] - o["e_total"])
        dg = float(np.max(np.abs(g["gradient"] - o["gradient"])))
        dq = float(np.max(np.abs(g["charges"] - o["q"])))
        print(f"mixed n={len(z)} nao={g['nao']}: dE(fp64 path)={dE:.2e} dE(oracle)={dEo:.2e} dg={dg:.2e} "
              f"dq={dq:.2e} it={g['iterations']}/{r['iterations']}", flush=True)
        assert g["converged"] and not g["failed"]
        assert dEo < 1e-6 and dg < 1e-5 and dq < 1e-5
