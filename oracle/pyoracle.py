"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OmOptions(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_double), ("etemp", ctypes.c_double),
                ("maxiter", ctypes.c_int), ("guess", ctypes.c_int), ("lapack", ctypes.c_int),
                ("solvent", ctypes.c_char_p), ("solv_kernel", ctypes.c_int), ("solv_alpb", ctypes.c_int),
                ("init_qsh", ctypes.POINTER(ctypes.c_double)), ("init_dipm", ctypes.POINTER(ctypes.c_double)),
                ("init_qp", ctypes.POINTER(ctypes.c_double))]


_D = ctypes.POINTER(ctypes.c_double)


class OmResult(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nsh", ctypes.c_int), ("nao", ctypes.c_int),
                ("ndim_d4", ctypes.c_int), ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("ihomo", ctypes.c_int),
                ("e_total", ctypes.c_double), ("e_elec", ctypes.c_double),
                ("e_es", ctypes.c_double), ("e_aes", ctypes.c_double),
                ("e_axc", ctypes.c_double), ("e_disp", ctypes.c_double),
                ("e_rep", ctypes.c_double), ("e_atm", ctypes.c_double),
                ("hl_gap_ev", ctypes.c_double), ("gnorm", ctypes.c_double),
                ("dipole", ctypes.c_double * 3), ("e_solv", ctypes.c_double),
                ("gradient", _D), ("q", _D), ("qsh", _D), ("dipm", _D), ("qp", _D),
                ("wbo", _D), ("emo", _D), ("focc", _D), ("density", _D), ("coeff", _D)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(_D)


def dims(numbers):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    nsh, nao, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib().om_dims(len(at), at.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nsh),
                       ctypes.byref(nao), ctypes.byref(nd))
    if rc:
        raise ValueError("unsupported element")
    return nsh.value, nao.value, nd.value


def singlepoint(numbers, positions, charge=0.0, uhf=0, accuracy=1.0, etemp=300.0,
                maxiter=250, guess="eeq", lapack=False, want_matrices=False, threads=1,
                solvent=None, solv_kernel="still", alpb=False, restart=None):
    """Run the oracle; returns a dict of results (Hartree, bohr, eV for emo).

    solvent: implicit solvent name (None: gas phase); solv_kernel "still"
    (GFN2/GBSA parameter sets) or "p16" (GFN2/ALPB parameter sets); alpb adds
    the ALPB shape correction.  restart: a previous result dict whose qsh,
    dipm and qp seed the SCC (the reference's reused results object)."""
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    n = len(at)
    nsh, nao, _ = dims(at)
    lib().om_set_threads(int(threads))
    out = {
        "gradient": np.zeros((n, 3)), "q": np.zeros(n), "qsh": np.zeros(nsh),
        "dipm": np.zeros((n, 3)), "qp": np.zeros((n, 6)), "wbo": np.zeros((n, n)),
        "emo": np.zeros(nao), "focc": np.zeros(nao),
    }
    if want_matrices:
        out["density"] = np.zeros((nao, nao))
        out["coeff"] = np.zeros((nao, nao))
    res = OmResult()
    for k, v in out.items():
        setattr(res, k, _p(v))
    opt = OmOptions(accuracy, etemp, maxiter, 0 if guess == "eeq" else 1, 1 if lapack else 0,
                    solvent.encode() if solvent else None, 2 if solv_kernel == "p16" else 1, 1 if alpb else 0)
    if restart is not None:
        keep = [np.ascontiguousarray(restart[k], dtype=np.float64).ravel() for k in ("qsh", "dipm", "qp")]
        if keep[0].size != nsh or keep[1].size != 3 * n or keep[2].size != 6 * n:
            raise ValueError("restart data of another molecule")
        opt.init_qsh, opt.init_dipm, opt.init_qp = (_p(a) for a in keep)
    rc = lib().om_singlepoint(ctypes.byref(opt), n, at.ctypes.data_as(ctypes.c_void_p),
                              _p(xyz), ctypes.c_double(charge), ctypes.c_int(uhf),
                              ctypes.byref(res))
    if rc:
        raise RuntimeError(f"oracle single point failed with code {rc}")
    for k in ("n", "nsh", "nao", "ndim_d4", "iterations", "converged", "ihomo", "e_total",
              "e_elec", "e_es", "e_aes", "e_axc", "e_disp", "e_rep", "e_atm", "hl_gap_ev",
              "gnorm", "e_solv"):
        out[k] = getattr(res, k)
    out["dipole"] = np.array(res.dipole[:])
    return out


def integrals(numbers, positions, accuracy=1.0):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    _, nao, _ = dims(at)
    S = np.zeros((nao, nao)); D = np.zeros((3, nao, nao)); Q = np.zeros((6, nao, nao))
    H0 = np.zeros((nao, nao)); H0n = np.zeros((nao, nao)); cn = np.zeros(len(at))
    lib().om_integrals(len(at), at.ctypes.data_as(ctypes.c_void_p), _p(xyz), ctypes.c_double(accuracy),
                       _p(S), _p(D), _p(Q), _p(H0), _p(H0n), _p(cn))
    return dict(S=S, D=D, Q=Q, H0=H0, H0_noovlp=H0n, cn=cn)


def basis_maps(numbers):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    n = len(at)
    nsh, nao, _ = dims(at)
    sh_at = np.zeros(nsh, np.int32); sh_l = np.zeros(nsh, np.int32); sh_ao0 = np.zeros(nsh, np.int32)
    ao_at = np.zeros(nao, np.int32); ao_sh = np.zeros(nao, np.int32)
    at_sh0 = np.zeros(n, np.int32); at_ao0 = np.zeros(n, np.int32)
    args = [a.ctypes.data_as(ctypes.c_void_p) for a in (sh_at, sh_l, sh_ao0, ao_at, ao_sh, at_sh0, at_ao0)]
    lib().om_basis_maps(n, at.ctypes.data_as(ctypes.c_void_p), *args)
    return dict(sh_at=sh_at, sh_l=sh_l, sh_ao0=sh_ao0, ao_at=ao_at, ao_sh=ao_sh, at_sh0=at_sh0, at_ao0=at_ao0)


def eeq(numbers, positions, charge=0.0):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    q = np.zeros(len(at))
    rc = lib().om_eeq_guess(len(at), at.ctypes.data_as(ctypes.c_void_p), _p(xyz), ctypes.c_double(charge), _p(q))
    if rc:
        raise RuntimeError("EEQ failed")
    return q


def repulsion(numbers, positions):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    e = ctypes.c_double()
    g = np.zeros((len(at), 3))
    lib().om_repulsion(len(at), at.ctypes.data_as(ctypes.c_void_p), _p(xyz), ctypes.byref(e), _p(g))
    return e.value, g


def coulomb_matrix(numbers, positions):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    nsh, _, _ = dims(at)
    j = np.zeros((nsh, nsh))
    lib().om_coulomb_matrix(len(at), at.ctypes.data_as(ctypes.c_void_p), _p(xyz), _p(j))
    return j


def cn_gfn(numbers, positions):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    n = len(at)
    cn = np.zeros(n); d = np.zeros((n, n, 3))
    lib().om_cn_gfn(n, at.ctypes.data_as(ctypes.c_void_p), _p(xyz), _p(cn), _p(d))
    return cn, d


def d4_setup(numbers, positions):
    at = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    _, _, nd = dims(at)
    gw = np.zeros(nd); dm = np.zeros((nd, nd))
    lib().om_d4_setup(len(at), at.ctypes.data_as(ctypes.c_void_p), _p(xyz), _p(gw), _p(dm))
    return gw, dm
