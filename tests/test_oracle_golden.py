"""Pin the FP64 CPU oracle against the reference's own golden vectors.

Every expected value below is read from tests/golden/gfn2_golden.json, which
tests/golden/make_golden.py extracts from xtb's test suite (the file:line of
each case is stored in its "source" field).  Tolerances are the reference's.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

SP_CASES = ["h2o_scc", "propyne", "water_tetramer"] + [f"mindless{k:02d}" for k in range(1, 11)] + \
    ["feco5", "co_cnx6", "fe_cnx6"]


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


SOLV_CASES = ["gbsa_water_api"] + [f"mindless{k:02d}_solv" for k in range(1, 11)] + ["feco5_gbsa", "feco5_alpb"]


@pytest.mark.parametrize("case", SOLV_CASES)
def test_solvation_singlepoint_golden(golden, case):
    """GBSA and GBSA + ALPB shape term (Still kernel) in eleven solvents against
    the reference's unit-test values (test/unit/test_gfn2.f90:237-308, 745-828)."""
    g = golden[case]
    r = po.singlepoint(g["numbers"], g["positions"], charge=g["charge"], uhf=g["uhf"],
                       guess=g["guess"], maxiter=g["maxiter"], accuracy=g["accuracy"],
                       etemp=g["etemp"], solvent=g["solvent"], solv_kernel=g["solv_kernel"],
                       alpb=g["alpb"], threads=4)
    assert r["converged"]
    assert r["e_solv"] < 0.0
    for key, (ref, tol) in g["expect"].items():
        val = _lookup(r, key)
        assert abs(val - ref) <= tol, f"{case} {key}: {val} vs {ref} (tol {tol})"


def test_solvation_c_api_restart_chain(golden):
    """The C API sequence of the reference's example program: propyne in the gas
    phase, then GBSA (Still kernel, water) restarted from it, then ALPB (P16
    kernel, ALPB water parameters) restarted from that
    (test/api/c_api_example.c:114-158)."""
    g = golden["propyne"]
    z, x = g["numbers"], g["positions"]
    r0 = po.singlepoint(z, x, maxiter=30)
    r1 = po.singlepoint(z, x, maxiter=30, solvent="h2o", solv_kernel="still", restart=r0)
    r2 = po.singlepoint(z, x, maxiter=30, solvent="h2o", solv_kernel="p16", alpb=True, restart=r1)
    for r, e, q5, d2, w9 in ((r1, -8.38393864716134, 0.06090868805034, -0.35455233974705, 2.89453979224265),
                             (r2, -8.384076843892, 0.0644849340, -0.3641866008, 2.8932146955)):
        assert abs(r["e_total"] - e) <= 1e-9
        assert abs(r["q"][5] - q5) <= 1e-8
        assert abs(r["dipole"][2] - d2) <= 1e-6
        assert abs(r["wbo"].ravel()[9] - w9) <= 1e-8


@pytest.mark.parametrize("kernel,alpb", [("still", False), ("p16", True), ("still", True)])
def test_solvation_gradient_finite_difference(kernel, alpb):
    """Analytic gradient with implicit solvent (Born radii, surface and H-bond
    derivatives, ALPB shape term) against central differences, all components.
    The surface quadrature drops grid points whose accessibility is below 1e-6
    (src/solv/sasa.f90:125), so the energy has tiny steps; the geometry is one
    where no grid point crosses that threshold within the difference step."""
    at = [6, 7, 8, 1, 1, 1]
    xyz = np.array([[0.0, 0.0, 0.0], [2.5, 0.3, -0.2], [-1.2, 2.11, 0.4],
                    [-1.1, -1.6, 0.9], [3.5, -1.2, 0.3], [-2.9, 2.0, -0.5]])
    kw = dict(accuracy=1e-4, maxiter=200, solvent="water", solv_kernel=kernel, alpb=alpb)
    r = po.singlepoint(at, xyz, **kw)
    h = 1e-4
    for a, k in [(a, k) for a in range(6) for k in range(3)]:
        xp = xyz.copy(); xp[a, k] += h
        xm = xyz.copy(); xm[a, k] -= h
        ep = po.singlepoint(at, xp, **kw)["e_total"]
        em = po.singlepoint(at, xm, **kw)["e_total"]
        assert abs((ep - em) / (2 * h) - r["gradient"][a, k]) < 2e-8, (kernel, alpb, a, k)


def test_solvent_names_and_unknown_solvent():
    """Solvent aliases resolve like the reference's normalizeSolventName; an
    unknown solvent is an error, not a silent gas-phase run."""
    at = [8, 1, 1]
    xyz = np.array([[0.0, 0.0, -0.73], [1.43, 0.0, 0.37], [-1.43, 0.0, 0.37]])
    e1 = po.singlepoint(at, xyz, solvent="water")["e_total"]
    e2 = po.singlepoint(at, xyz, solvent="H2O")["e_total"]
    assert e1 == e2
    e3 = po.singlepoint(at, xyz, solvent="dichloromethane")["e_total"]
    e4 = po.singlepoint(at, xyz, solvent="ch2cl2")["e_total"]
    assert e3 == e4 and e3 != e1
    with pytest.raises(RuntimeError):
        po.singlepoint(at, xyz, solvent="no-such-solvent")
