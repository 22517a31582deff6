"""Pin the FP64 CPU oracle against the reference's own golden vectors.

Every expected value below is read from tests/golden/gfn2_golden.json, which
tests/golden/make_golden.py extracts from xtb's test suite (the file:line of
each case is stored in its "source" field).  Tolerances are the reference's.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

SP_CASES = ["h2o_scc", "propyne", "water_tetramer"] + [f"mindless{k:02d}" for k in range(1, 11)]


def _lookup(res, key):
    if key == "emo_homo":
        return res["emo"][res["ihomo"] - 1]
    if key.startswith("wbo_flat["):
        return res["wbo"].ravel()[int(key[9:-1])]
    name, _, idx = key.partition("[")
    val = res[name]
    for part in idx.rstrip("]").split("]["):
        if part:
            val = val[int(part)]
    return val


@pytest.mark.parametrize("case", SP_CASES)
def test_singlepoint_golden(golden, case):
    g = golden[case]
    r = po.singlepoint(g["numbers"], g["positions"], charge=g["charge"], uhf=g["uhf"],
                       guess=g["guess"], maxiter=g["maxiter"], accuracy=g["accuracy"],
                       etemp=g["etemp"])
    assert r["converged"]
    for key, (ref, tol) in g["expect"].items():
        val = _lookup(r, key)
        assert abs(val - ref) <= tol, f"{case} {key}: {val} vs {ref} (tol {tol})"


def test_eeq_water_golden(golden):
    g = golden["eeq_water"]
    q = po.eeq(g["numbers"], g["positions"])
    ref, tol = g["expect"]["q[0]"]
    assert abs(q[0] - ref) <= tol
    assert abs(q.sum()) < 1e-12


def test_repulsion_caffeine_golden(golden):
    g = golden["caffeine_repulsion"]
    e, grad = po.repulsion(g["numbers"], g["positions"])
    ref, tol = g["expect"]["e_rep"]
    assert abs(e - ref) <= tol
    # finite-difference gradient, step as in test/unit/test_repulsion.f90
    xyz = np.array(g["positions"])
    h = 1e-5
    for a, k in [(0, 0), (5, 1), (20, 2)]:
        xp = xyz.copy(); xp[a, k] += h
        xm = xyz.copy(); xm[a, k] -= h
        fd = (po.repulsion(g["numbers"], xp)[0] - po.repulsion(g["numbers"], xm)[0]) / (2 * h)
        assert abs(fd - grad[a, k]) < 1e-8


def test_gradient_finite_difference():
    """Analytic gradient against central differences at tight convergence."""
    at = [6, 7, 8, 1, 1, 1]
    xyz = np.array([[0.0, 0.0, 0.0], [2.5, 0.3, -0.2], [-1.2, 2.1, 0.4],
                    [-1.1, -1.6, 0.9], [3.5, -1.2, 0.3], [-2.9, 2.0, -0.5]])
    r = po.singlepoint(at, xyz, accuracy=1e-4, maxiter=200)
    h = 1e-4
    for a, k in [(0, 0), (1, 2), (2, 1), (4, 0)]:
        xp = xyz.copy(); xp[a, k] += h
        xm = xyz.copy(); xm[a, k] -= h
        ep = po.singlepoint(at, xp, accuracy=1e-4, maxiter=200)["e_total"]
        em = po.singlepoint(at, xm, accuracy=1e-4, maxiter=200)["e_total"]
        assert abs((ep - em) / (2 * h) - r["gradient"][a, k]) < 2e-8


def test_translation_invariance_and_sum_rules(golden):
    g = golden["propyne"]
    xyz = np.array(g["positions"])
    r1 = po.singlepoint(g["numbers"], xyz, maxiter=30, want_matrices=True)
    r2 = po.singlepoint(g["numbers"], xyz + np.array([1.7, -2.3, 0.9]), maxiter=30)
    assert abs(r1["e_total"] - r2["e_total"]) < 1e-9
    assert np.abs(r1["gradient"].sum(axis=0)).max() < 1e-9
    assert abs(r1["q"].sum()) < 1e-9
    assert abs(r1["focc"].sum() - 16.0) < 1e-9


def test_overlap_closed_form():
    """s-s overlap of two H 1s STO-3G shells against the Gaussian product formula."""
    from oracle.pyoracle import integrals
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]])
    S = integrals([1, 1], xyz)["S"]
    # rebuild the contracted 1s from the published STO-3G expansion scaled by zeta^2
    import re, os
    hdr = open(os.path.join(os.path.dirname(po.__file__), "..", "xtb_amd", "csrc", "gfn2_params.h")).read()
    def tab(name):
        m = re.search(r"k_" + name + r"\[\d+\] = \{(.*?)\};", hdr, re.S)
        return [float.fromhex(v.strip()) for v in m.group(1).split(",") if v.strip()]
    zeta = tab("slater_exponent")[0]
    a = np.array(tab("sto_alpha3")[:3]) * zeta ** 2
    c = np.array(tab("sto_coeff3")[:3]) * (2 * a / np.pi) ** 0.75
    r2 = 1.4 ** 2
    s = 0.0
    for i in range(3):
        for j in range(3):
            g = a[i] + a[j]
            s += c[i] * c[j] * (np.pi / g) ** 1.5 * np.exp(-a[i] * a[j] / g * r2)
    assert abs(S[0, 1] - s) < 1e-13
    assert abs(S[0, 0] - 1.0) < 1e-15
