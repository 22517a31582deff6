"""ctypes binding of the batch C ABI (include/xtb_amd_batch.h).

``singlepoint`` mirrors the behaviour of xtb_singlepoint for one molecule
(src/api/interface.f90:40-156); ``Batch`` runs many molecules as one batch.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libxtb_amd.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class Result(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("natoms", "nshells", "nao", "iterations", "converged", "failed", "ihomo")] + \
        [(k, ctypes.c_double) for k in ("energy", "gnorm", "hl_gap_ev", "e_elec", "e_es", "e_aes", "e_axc",
                                        "e_disp", "e_rep", "e_atm", "e_solv")] + \
        [("dipole", ctypes.c_double * 3)] + \
        [(k, _dp) for k in ("gradient", "charges", "qsh", "dipm", "qp", "wbo", "emo", "focc", "coeff", "density")]


def library():
    """Load libxtb_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with `python -m xtb_amd.build`")
    lib = ctypes.CDLL(LIBPATH)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.xtbamd_batch_new.restype = ctypes.c_void_p
    lib.xtbamd_batch_new.argtypes = [ctypes.c_int, ip, ip, _dp, _dp, ip, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int]
    lib.xtbamd_last_error.restype = ctypes.c_char_p
    for f in ("xtbamd_batch_setup", "xtbamd_batch_iterate", "xtbamd_batch_gradient", "xtbamd_batch_run",
              "xtbamd_batch_size"):
        getattr(lib, f).argtypes = [ctypes.c_void_p]
        getattr(lib, f).restype = ctypes.c_int
    lib.xtbamd_batch_set_guess.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
    lib.xtbamd_batch_set_solvent.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double]
    lib.xtbamd_solvent_known.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.xtbamd_batch_set_guess.restype = ctypes.c_int
    lib.xtbamd_batch_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Result)]
    lib.xtbamd_batch_get.restype = ctypes.c_int
    lib.xtbamd_batch_charges.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.xtbamd_batch_charges.restype = ctypes.c_int
    lib.xtbamd_batch_free.argtypes = [ctypes.c_void_p]
    lib.xtbamd_batch_free.restype = None
    lib.xtbamd_device_count.restype = ctypes.c_int
    lib.xtbamd_copy_probe.argtypes = [ctypes.c_longlong]
    lib.xtbamd_copy_probe.restype = ctypes.c_int
    lib.xtbamd_server_new.restype = ctypes.c_void_p
    lib.xtbamd_server_new.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int]
    lib.xtbamd_server_submit.restype = ctypes.c_longlong
    lib.xtbamd_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, _dp, ctypes.c_double, ctypes.c_int]
    lib.xtbamd_server_step.restype = ctypes.c_int
    lib.xtbamd_server_step.argtypes = [ctypes.c_void_p]
    lib.xtbamd_server_finished.restype = ctypes.c_int
    lib.xtbamd_server_finished.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    lib.xtbamd_server_result.restype = ctypes.c_int
    lib.xtbamd_server_result.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(Result)]
    lib.xtbamd_server_pending.restype = ctypes.c_int
    lib.xtbamd_server_pending.argtypes = [ctypes.c_void_p, ip, ip]
    lib.xtbamd_server_free.restype = None
    lib.xtbamd_server_free.argtypes = [ctypes.c_void_p]
    lib.xtbamd_server_set_solvent.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_double]
    lib.xtbamd_batch_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_longlong)]
    lib.xtbamd_batch_counters.restype = ctypes.c_int
    lib.xtbamd_batch_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.xtbamd_batch_profile.restype = ctypes.c_int
    lib.xtbamd_batch_profile_result.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.xtbamd_batch_profile_result.restype = ctypes.c_int
    lib.xtbamd_batch_profile_span.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.xtbamd_batch_profile_span.restype = ctypes.c_int
    lib.xtbamd_singlepoint_noalloc.argtypes = [ctypes.c_int, ip, _dp, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.POINTER(Result)]
    lib.xtbamd_singlepoint_noalloc.restype = ctypes.c_int
    lib.xtbamd_eigh.argtypes = [ctypes.c_int, ip, _dp, _dp, _dp, _dp]
    lib.xtbamd_eigh.restype = ctypes.c_int
    _lib = lib
    return lib


def device_count():
    return library().xtbamd_device_count()


def _err(lib):
    return lib.xtbamd_last_error().decode()


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Batch:
    """Many GFN2-xTB single points computed together on the GPU.

    molecules: list of (numbers, positions[bohr], charge, uhf) tuples.
    """

    def __init__(self, molecules, accuracy=1.0, etemp=300.0, maxiter=250, guess="eeq"):
        lib = library()
        self._lib = lib
        nat = np.array([len(m[0]) for m in molecules], dtype=np.int32)
        num = np.ascontiguousarray(np.concatenate([np.asarray(m[0], dtype=np.int32) for m in molecules]))
        pos = np.ascontiguousarray(np.concatenate([np.asarray(m[1], dtype=np.float64).reshape(-1, 3)
                                                   for m in molecules]))
        chg = np.array([float(m[2]) if len(m) > 2 else 0.0 for m in molecules], dtype=np.float64)
        uhf = np.array([int(m[3]) if len(m) > 3 else 0 for m in molecules], dtype=np.int32)
        g = {"eeq": 0, "sad": 1}[guess] if isinstance(guess, str) else int(guess)
        ip = ctypes.POINTER(ctypes.c_int)
        h = lib.xtbamd_batch_new(len(molecules), nat.ctypes.data_as(ip), num.ctypes.data_as(ip), _ptr(pos),
                                 _ptr(chg), uhf.ctypes.data_as(ip), accuracy, etemp, maxiter, g)
        if not h:
            raise RuntimeError(_err(lib))
        self._h = h
        self.natoms = nat

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(_err(self._lib))
        return rc

    def set_guess(self, m, qsh, dipm, qp):
        """Restart guess of molecule m from a previous result's qsh, dipm, qp."""
        qsh = np.ascontiguousarray(qsh, dtype=np.float64)
        dipm = np.ascontiguousarray(dipm, dtype=np.float64)
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        self._check(self._lib.xtbamd_batch_set_guess(self._h, m, _ptr(qsh), _ptr(dipm), _ptr(qp)))

    def set_solvent(self, solvent, kernel="still", alpb=False, state=1, temperature=298.15):
        """Implicit solvent for every molecule (None: gas phase).  kernel "still"
        uses the GFN2/GBSA parameter sets, "p16" the GFN2/ALPB sets; alpb adds
        the ALPB shape term (the C API's "gbsa" is still/False, "alpb" p16/True)."""
        k = {"still": 1, "p16": 2}[kernel]
        self._check(self._lib.xtbamd_batch_set_solvent(self._h, solvent.encode() if solvent else None, k,
                                                      1 if alpb else 0, int(state), float(temperature)))
        return self

    def setup(self):
        self._check(self._lib.xtbamd_batch_setup(self._h))

    def iterate(self):
        return self._check(self._lib.xtbamd_batch_iterate(self._h))

    def gradient(self):
        self._check(self._lib.xtbamd_batch_gradient(self._h))

    def run(self):
        self._check(self._lib.xtbamd_batch_run(self._h))
        return self

    def counters(self):
        """(SCC steps launched, block-Jacobi rounds) of the last run."""
        a, b = ctypes.c_int(), ctypes.c_longlong()
        self._check(self._lib.xtbamd_batch_counters(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    KERNELS = {"gemm": 1, "jacobi": 2, "apply": 3}

    def profile(self, *kernels):
        """Start HIP-event timing of kernel families ("gemm", "jacobi" pivots, "apply" tile updates)."""
        bits = 0
        for k in kernels:
            bits |= 1 << (self.KERNELS[k] - 1)
        self._check(self._lib.xtbamd_batch_profile(self._h, bits))

    def profile_result(self, kernel="gemm"):
        """(launches, summed ms, summed algorithmic work: flops, or bytes for "apply") since profile()."""
        n, t, w = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.xtbamd_batch_profile_result(self._h, self.KERNELS[kernel], ctypes.byref(n),
                                                          ctypes.byref(t), ctypes.byref(w)))
        return n.value, t.value, w.value

    def profile_span(self, kernel="apply"):
        """Summed execution spans (device wall clock) of the sampled launches,
        scaled to all launches, in ms (tile updates only)."""
        t = ctypes.c_double()
        self._check(self._lib.xtbamd_batch_profile_span(self._h, self.KERNELS[kernel], ctypes.byref(t)))
        return t.value

    def charges(self, m):
        """Current atomic charges of molecule m (after setup(): the initial guess)."""
        q = np.zeros(int(self.natoms[m]))
        self._check(self._lib.xtbamd_batch_charges(self._h, m, _ptr(q)))
        return q

    def __len__(self):
        return len(self.natoms)

    def result(self, m, matrices=False):
        lib = self._lib
        r = Result()
        self._check(lib.xtbamd_batch_get(self._h, m, ctypes.byref(r)))   # sizes first
        n, nsh, nao = r.natoms, r.nshells, r.nao
        out = {"gradient": np.zeros((n, 3)), "charges": np.zeros(n), "qsh": np.zeros(nsh),
               "dipm": np.zeros((n, 3)), "qp": np.zeros((n, 6)), "wbo": np.zeros((n, n)),
               "emo": np.zeros(nao), "focc": np.zeros(nao)}
        if matrices:
            out["coeff"] = np.zeros((nao, nao))
            out["density"] = np.zeros((nao, nao))
        for k, v in out.items():
            setattr(r, k, _ptr(v))
        self._check(lib.xtbamd_batch_get(self._h, m, ctypes.byref(r)))
        for k in ("natoms", "nshells", "nao", "iterations", "converged", "failed", "ihomo", "energy", "gnorm",
                  "hl_gap_ev", "e_elec", "e_es", "e_aes", "e_axc", "e_disp", "e_rep", "e_atm", "e_solv"):
            out[k] = getattr(r, k)
        out["dipole"] = np.array(r.dipole[:])
        return out

    def results(self, matrices=False):
        return [self.result(m, matrices) for m in range(len(self))]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xtbamd_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fetch(lib, get, check):
    """Sizes first (all array pointers NULL), then the arrays."""
    r = Result()
    check(get(ctypes.byref(r)))
    n, nsh, nao = r.natoms, r.nshells, r.nao
    out = {"gradient": np.zeros((n, 3)), "charges": np.zeros(n), "qsh": np.zeros(nsh),
           "dipm": np.zeros((n, 3)), "qp": np.zeros((n, 6)), "wbo": np.zeros((n, n)),
           "emo": np.zeros(nao), "focc": np.zeros(nao)}
    for k, v in out.items():
        setattr(r, k, _ptr(v))
    check(get(ctypes.byref(r)))
    for k in ("natoms", "nshells", "nao", "iterations", "converged", "failed", "ihomo", "energy", "gnorm",
              "hl_gap_ev", "e_elec", "e_es", "e_aes", "e_axc", "e_disp", "e_rep", "e_atm", "e_solv"):
        out[k] = getattr(r, k)
    out["dipole"] = np.array(r.dipole[:])
    return out


class Server:
    """Continuous batching of single points (xtbamd_server_* of include/xtb_amd_batch.h)."""

    def __init__(self, hbm_gb=8.0, max_atoms=400, max_slots=1024, accuracy=1.0, etemp=300.0, maxiter=250):
        lib = library()
        self._lib = lib
        h = lib.xtbamd_server_new(hbm_gb, max_atoms, max_slots, accuracy, etemp, maxiter)
        if not h:
            raise RuntimeError(_err(lib))
        self._h = h

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(_err(self._lib))
        return rc

    def submit(self, numbers, positions, charge=0.0, uhf=0):
        num = np.ascontiguousarray(numbers, dtype=np.int32)
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        rid = self._lib.xtbamd_server_submit(self._h, len(num), num.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             _ptr(pos), float(charge), int(uhf))
        if rid < 0:
            raise RuntimeError(_err(self._lib))
        return rid

    def set_solvent(self, solvent, kernel="still", alpb=False, state=1, temperature=298.15):
        """Implicit solvent of the requests admitted from now on (see Batch.set_solvent)."""
        k = {"still": 1, "p16": 2}[kernel]
        self._check(self._lib.xtbamd_server_set_solvent(self._h, solvent.encode() if solvent else None, k,
                                                       1 if alpb else 0, int(state), float(temperature)))
        return self

    def step(self):
        """Admit, one SCC iteration of all resident molecules, finish; returns requests not finished."""
        return self._check(self._lib.xtbamd_server_step(self._h))

    def finished(self, max_ids=4096):
        ids = (ctypes.c_longlong * max_ids)()
        n = self._check(self._lib.xtbamd_server_finished(self._h, ids, max_ids))
        return list(ids[:n])

    def result(self, rid):
        return _fetch(self._lib, lambda p: self._lib.xtbamd_server_result(self._h, rid, p), self._check)

    def pending(self):
        q, r = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.xtbamd_server_pending(self._h, ctypes.byref(q), ctypes.byref(r)))
        return q.value, r.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xtbamd_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# MolState failure codes (xtbamd_result.failed)
FAILURES = {2: "orthonormalization of the overlap matrix failed", 3: "diagonalization of Hamiltonian failed",
            4: "EEQ guess charges failed", 7: "solvation surface setup failed"}


class SolverFailure(RuntimeError):
    """A solver of the single point gave up (result["failed"] != 0); the
    result is attached as .result."""

    def __init__(self, result):
        super().__init__(FAILURES.get(result["failed"], f"single point failed (code {result['failed']})"))
        self.result = result


def singlepoint(numbers, positions, charge=0.0, uhf=0, accuracy=1.0, etemp=300.0, maxiter=250, guess="eeq",
                matrices=False, solvent=None, solv_kernel="still", alpb=False):
    """One single point; raises SolverFailure when a solver gave up."""
    b = Batch([(numbers, positions, charge, uhf)], accuracy, etemp, maxiter, guess)
    try:
        if solvent:
            b.set_solvent(solvent, solv_kernel, alpb)
        b.run()
        r = b.result(0, matrices)
    finally:
        b.close()
    if r["failed"]:
        raise SolverFailure(r)
    return r


def singlepoint_noalloc(numbers, positions, charge=0.0, uhf=0, accuracy=1.0, etemp=300.0, maxiter=250):
    """One single point through xtbamd_singlepoint_noalloc (no device
    allocation: the library's per-device scratch area); the same result as
    singlepoint() with the EEQ guess.  Raises SolverFailure when a solver gave up."""
    lib = library()
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    n = len(num)
    # shells and AOs per atom are at most 3 and 9 (s, p, d): arrays sized for
    # that bound, trimmed to the reported sizes
    out = {"gradient": np.zeros((n, 3)), "charges": np.zeros(n), "qsh": np.zeros(3 * n),
           "dipm": np.zeros((n, 3)), "qp": np.zeros((n, 6)), "wbo": np.zeros((n, n)),
           "emo": np.zeros(9 * n), "focc": np.zeros(9 * n)}
    r = Result()
    for k, v in out.items():
        setattr(r, k, _ptr(v))
    rc = lib.xtbamd_singlepoint_noalloc(n, num.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(pos),
                                        float(charge), int(uhf), accuracy, etemp, maxiter, ctypes.byref(r))
    if rc < 0:
        raise RuntimeError(_err(lib))
    for k in ("natoms", "nshells", "nao", "iterations", "converged", "failed", "ihomo", "energy", "gnorm",
              "hl_gap_ev", "e_elec", "e_es", "e_aes", "e_axc", "e_disp", "e_rep", "e_atm", "e_solv"):
        out[k] = getattr(r, k)
    out["dipole"] = np.array(r.dipole[:])
    out["qsh"] = out["qsh"][:r.nshells].copy()
    out["emo"] = out["emo"][:r.nao].copy()
    out["focc"] = out["focc"][:r.nao].copy()
    if out["failed"]:
        raise SolverFailure(out)
    return out


def eigh(matrices, tridiagonal=False):
    """Eigen-decomposition of symmetric matrices by the engine's blocked
    eigensolver (xtbamd_eigh): [(eigenvalues ascending, eigenvectors as rows)]
    (and the tridiagonal forms [(d, e)] when tridiagonal=True)."""
    lib = library()
    mats = [np.ascontiguousarray(m, dtype=np.float64) for m in matrices]
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    a = np.ascontiguousarray(np.concatenate([m.ravel() for m in mats]))
    ev = np.zeros(int(n.sum()))
    vec = np.zeros(a.size)
    tri = np.zeros(2 * int(n.sum()))
    rc = lib.xtbamd_eigh(len(mats), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(a), _ptr(ev), _ptr(vec),
                         _ptr(tri))
    if rc < 0:
        raise RuntimeError(_err(lib))
    out, tr, o, o2 = [], [], 0, 0
    for k in n:
        k = int(k)
        out.append((ev[o:o + k].copy(), vec[o2:o2 + k * k].reshape(k, k).copy()))
        tr.append((tri[2 * o:2 * o + k].copy(), tri[2 * o + k:2 * o + 2 * k - 1].copy()))
        o += k
        o2 += k * k
    return (out, tr) if tridiagonal else out
