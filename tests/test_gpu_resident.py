"""A results object keeps its engine on the device (src/api/results.f90:37-46,
reuse at src/api/interface.f90:83-101): after xtb_updateMolecule, a second
xtb_singlepoint with the same results object reuses the device arrays and
starts from the wavefunction the first one left on the device.  It must equal,
bitwise, the single point of a copied results object (wavefunction on the
host, new engine seeded from it), and agree with the oracle at the new
geometry within the path's tolerances."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)


def _api():
    import xtb_amd
    if xtb_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    lib = xtb_amd.library()
    for f in ("xtb_newEnvironment", "xtb_newCalculator", "xtb_newResults"):
        getattr(lib, f).restype = vp
        getattr(lib, f).argtypes = []
    lib.xtb_newMolecule.restype = vp
    lib.xtb_newMolecule.argtypes = [vp, ip, ip, dp, dp, ip, dp, vp]
    lib.xtb_copyResults.restype = vp
    lib.xtb_copyResults.argtypes = [vp]
    lib.xtb_loadGFN2xTB.argtypes = [vp, vp, vp, ctypes.c_char_p]
    lib.xtb_setVerbosity.argtypes = [vp, ctypes.c_int]
    lib.xtb_singlepoint.argtypes = [vp, vp, vp, vp]
    lib.xtb_updateMolecule.argtypes = [vp, vp, dp, dp]
    lib.xtb_checkEnvironment.argtypes = [vp]
    lib.xtb_checkEnvironment.restype = ctypes.c_int
    for f in ("xtb_getEnergy", "xtb_getGradient", "xtb_getCharges"):
        getattr(lib, f).argtypes = [vp, vp, dp]
    for f in ("xtb_delResults", "xtb_delMolecule", "xtb_delCalculator", "xtb_delEnvironment"):
        getattr(lib, f).argtypes = [ctypes.POINTER(vp)]
    return lib


def _p(a):
    return a.ctypes.data_as(dp)


def _sp(lib, env, mol, calc, res, n):
    lib.xtb_singlepoint(env, mol, calc, res)
    assert lib.xtb_checkEnvironment(env) == 0
    e = ctypes.c_double()
    g = np.zeros((n, 3))
    q = np.zeros(n)
    lib.xtb_getEnergy(env, res, ctypes.byref(e))
    lib.xtb_getGradient(env, res, _p(g))
    lib.xtb_getCharges(env, res, _p(q))
    return e.value, g, q


def _molecule(lib, env, z, x):
    n = ctypes.c_int(len(z))
    zz = np.ascontiguousarray(z, dtype=np.int32)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    mol = lib.xtb_newMolecule(env, ctypes.byref(n), zz.ctypes.data_as(ip), _p(xx), ctypes.byref(ctypes.c_double(0.0)),
                              ctypes.byref(ctypes.c_int(0)), None, None)
    assert mol
    return vp(mol)


def test_results_engine_resident_across_updates():
    from xtb_amd import workloads as w
    lib = _api()
    env = vp(lib.xtb_newEnvironment())
    lib.xtb_setVerbosity(env, 0)
    calc = vp(lib.xtb_newCalculator())
    z, x0 = w.water_cluster(8, 21)
    n = len(z)
    mol = _molecule(lib, env, z, x0)
    lib.xtb_loadGFN2xTB(env, mol, calc, None)
    res = vp(lib.xtb_newResults())
    _sp(lib, env, mol, calc, res, n)
    rng = np.random.default_rng(5)
    x = x0.copy()
    for step in range(3):
        # the copy carries the wavefunction on the host and no engine
        copy = vp(lib.xtb_copyResults(res))
        x = x + rng.uniform(-0.02, 0.02, x.shape)
        xc = np.ascontiguousarray(x)
        lib.xtb_updateMolecule(env, mol, _p(xc), None)
        assert lib.xtb_checkEnvironment(env) == 0
        e1, g1, q1 = _sp(lib, env, mol, calc, res, n)      # resident engine and wavefunction
        e2, g2, q2 = _sp(lib, env, mol, calc, copy, n)     # new engine, host guess
        print(f"step {step}: E={e1:.12f} resident-vs-host dE={abs(e1 - e2):.1e}", flush=True)
        assert e1 == e2 and np.array_equal(g1, g2) and np.array_equal(q1, q2)
        o = po.singlepoint(z, x)
        assert abs(e1 - o["e_total"]) < 1e-6
        assert np.max(np.abs(g1 - o["gradient"])) < 1e-5
        assert np.max(np.abs(q1 - o["q"])) < 1e-5
        lib.xtb_delResults(ctypes.byref(copy))
    # another molecule on the same results object: the engine is rebuilt
    z2, x2 = w.alkane(6, 3)
    mol2 = _molecule(lib, env, z2, x2)
    e3, g3, _ = _sp(lib, env, mol2, calc, res, len(z2))
    o = po.singlepoint(z2, x2)
    assert abs(e3 - o["e_total"]) < 1e-6 and np.max(np.abs(g3 - o["gradient"])) < 1e-5
    for h in (res,):
        lib.xtb_delResults(ctypes.byref(h))
    lib.xtb_delMolecule(ctypes.byref(mol))
    lib.xtb_delMolecule(ctypes.byref(mol2))
    lib.xtb_delCalculator(ctypes.byref(calc))
    lib.xtb_delEnvironment(ctypes.byref(env))
