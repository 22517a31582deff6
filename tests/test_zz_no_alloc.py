"""xtbamd_singlepoint_noalloc (no device allocation, per-device scratch area)
against the allocating entry point (xtbamd_batch_new + xtbamd_batch_run +
xtbamd_batch_get) on the same molecules: two calls in a row, the second with
a smaller molecule, so the second call runs on the first call's leftovers."""
import numpy as np
import pytest

import xtb_amd
from xtb_amd import workloads

KEYS = ("energy", "gnorm", "hl_gap_ev", "e_es", "e_aes", "e_axc", "e_disp", "e_rep", "iterations", "converged",
        "nao", "nshells", "ihomo")
ARRAYS = ("gradient", "charges", "qsh", "dipm", "qp", "wbo", "emo", "focc", "dipole")


def _same(a, b):
    for k in KEYS:
        assert a[k] == b[k], (k, a[k], b[k])
    for k in ARRAYS:
        assert np.array_equal(a[k], b[k]), (k, float(np.max(np.abs(a[k] - b[k]))))


@pytest.mark.gpu
def test_noalloc_same_as_batch_entry_two_calls():
    cases = [("C24H50", workloads.alkane(24, 5)), ("(H2O)6", workloads.water_cluster(6, 11))]
    for name, (z, x) in cases:
        print(f"noalloc case {name}: {len(z)} atoms", flush=True)
        ref = xtb_amd.singlepoint(z, x)
        got = xtb_amd.singlepoint_noalloc(z, x)
        assert got["converged"] and got["failed"] == 0
        _same(got, ref)
    # a repeat of the larger molecule after the smaller one
    z, x = cases[0][1]
    _same(xtb_amd.singlepoint_noalloc(z, x), xtb_amd.singlepoint(z, x))


@pytest.mark.gpu
def test_noalloc_rejects_molecule_beyond_scratch():
    z, x = workloads.large_water_cluster(2024)   # 4200 AOs
    with pytest.raises(RuntimeError, match="scratch area"):
        xtb_amd.singlepoint_noalloc(z, x)
