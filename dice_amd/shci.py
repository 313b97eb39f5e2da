"""Python mirror of the reference's SHCI variational-step interface, bound to
the HIP library (dice_amd/lib/libshci_gpu.so) through its C ABI.

Names follow the reference's C++ entry points:

=====================================  =================================================
this module                            reference (sanshar/Dice)
=====================================  =================================================
``System.from_fcidump``                readIntegrals + twoIntHeatBath(SHM)::constructClass
                                       (SHCI/integral.cpp:43,228; SHCI/integral.h:125)
``System.energy``                      Determinant::Energy (SHCI/Determinants.cpp:252)
``System.hij``                         Hij (SHCI/Determinants.cpp:564)
``System.get_determinants_variational``  getDeterminantsVariationalApprox + sort/unique
                                       (SHCI/SHCIgetdeterminants.cpp:586; SHCI/SHCIbasics.cpp:984-1055)
``System.make_hamiltonian``            PopulateHelpers + MakeHfromSMHelpers2
                                       (SHCI/SHCImakeHamiltonian.cpp:290,486)
``System.hmult``                       Hmult2::operator() (SHCI/Hmult.h:52)
``System.davidson``                    davidson (SHCI/Davidson.cpp:111)
``do_variational``                     SHCIbasics::DoVariational via ``Dice input.dat``
                                       (SHCI/SHCIbasics.cpp:783; SHCI/SHCI.cpp:157-446)
=====================================  =================================================

Determinants are ``numpy.int64`` arrays of shape (n, words), the reference's
``Determinant::repr`` layout.  There is no CPU fallback: every call runs the
HIP kernels, and loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libshci_gpu.so")
EXE_PATH = os.path.join(_HERE, "bin", "shci_var")

PREC_FP64 = 0
PREC_FP32 = 1

_STATUS = {0: "OK", 1: "BAD_INPUT", 2: "OOM", 3: "NO_CONVERGENCE", 4: "HIP_ERROR", 5: "UNSUPPORTED"}

_lib = None


class RunOptions(ctypes.Structure):
    """shci_run_options of include/shci_gpu.h."""
    _fields_ = [("precision", ctypes.c_int), ("profile_sigma", ctypes.c_int), ("davidson_block", ctypes.c_int),
                ("autotune", ctypes.c_int), ("hip_graphs", ctypes.c_int), ("learned", ctypes.c_int),
                ("direct_sigma", ctypes.c_int), ("half_storage", ctypes.c_int)]


class VarParams(ctypes.Structure):
    """shci_var_params of include/shci_gpu.h."""
    _fields_ = [("nroots", ctypes.c_int), ("n_iter", ctypes.c_int), ("eps1", ctypes.POINTER(ctypes.c_double)),
                ("davidson_tol", ctypes.c_double), ("davidson_tol_loose", ctypes.c_double), ("dE", ctypes.c_double),
                ("seed", ctypes.c_uint), ("ncore", ctypes.c_int), ("nact", ctypes.c_int),
                ("reserved", ctypes.c_int * 4)]


class ShciError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (_STATUS.get(status, status), message))
        self.status = status


def library():
    """Load libshci_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("HIP library %s is not built; run `python -m dice_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    pI64, pI32, pD = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    pF = ctypes.POINTER(ctypes.c_float)
    sig = {
        "shci_create": ([ctypes.c_char_p, ctypes.POINTER(P)], I32),
        "shci_destroy": ([P], I32),
        "shci_create_from_integrals": ([I32, I32, pD, pD, D, pI32, ctypes.POINTER(P)], I32),
        "shci_do_variational": ([P, ctypes.POINTER(VarParams), pI64, I64, pD, pD], I32),
        "shci_last_error": ([P], ctypes.c_char_p),
        "shci_norb": ([P], I32),
        "shci_nelec": ([P], I32),
        "shci_det_words": ([P], I32),
        "shci_core_energy": ([P], D),
        "shci_set_precision": ([P, I32], I32),
        "shci_set_options": ([P, ctypes.POINTER(RunOptions)], I32),
        "shci_set_active_space": ([P, I32, I32], I32),
        "shci_add_restriction": ([P, I32, I32, pI32, I32], I32),
        "shci_det_energy": ([P, pI64, I64, pD], I32),
        "shci_hij": ([P, pI64, pI64, I64, pD], I32),
        "shci_set_dets": ([P, pI64, I64], I32),
        "shci_num_dets": ([P, pI64], I32),
        "shci_get_dets": ([P, pI64, I64], I32),
        "shci_screen": ([P, pD, D, pI64], I32),
        "shci_get_new_dets": ([P, pI64, I64], I32),
        "shci_append_new": ([P], I32),
        "shci_build_hamiltonian": ([P, pI64], I32),
        "shci_get_hamiltonian_rows": ([P, I64, pI64, pI32, pD, I64], I32),
        "shci_sigma": ([P, pD, pD, I32], I32),
        "shci_sigma_device": ([P, P, P, I32, I64], I32),
        "shci_sigma_device_noalloc": ([P, P, P, I32, I64], I32),
        "shci_davidson": ([P, pD, I32, I32, D, pD, ctypes.POINTER(ctypes.c_int)], I32),
        "shci_pt_deterministic": ([P, pD, D, D, pD, pD, pI64], I32),
        "shci_pt_semistochastic": ([P, pD, D, D, D, I32, I32, D, ctypes.c_uint, pD, pD, I32,
                                    ctypes.POINTER(ctypes.c_int), pD], I32),
        "shci_rdm": ([P, pD, pD, pD, pD, pD, pD], I32),
        "shci_run": ([ctypes.c_char_p, ctypes.c_char_p, I32, ctypes.POINTER(P)], I32),
        "shci_run_ex": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(RunOptions), ctypes.POINTER(P)], I32),
        "shci_run_iteration_phases": ([P, I32, pD, pD, pD, pI64], I32),
        "shci_run_final_davidson_seconds": ([P], D),
        "shci_profile_sigma_read": ([P, pI64, pD, pD], I32),
        "shci_run_iterations": ([P], I32),
        "shci_run_iteration": ([P, I32, pI64, pD, ctypes.POINTER(ctypes.c_int), pD], I32),
        "shci_run_energies": ([P, pD, ctypes.POINTER(ctypes.c_int)], I32),
        "shci_run_ci": ([P, pD, I64], I32),
        "shci_run_pt_energies": ([P, pD, ctypes.POINTER(ctypes.c_int)], I32),
        "shci_run_learned_count": ([P], I32),
        "shci_run_learned": ([P, I32, ctypes.POINTER(ctypes.c_int), pI64, pI64, pI64, pD, pD,
                              ctypes.POINTER(ctypes.c_int), pD], I32),
        "shci_run_learned_training": ([P, pI64, pF, pF, pD], I32),
        "shci_mlp_param_count": ([I32, I32], I64),
        "shci_mlp_init": ([pF, I32, I32, ctypes.c_uint64], I32),
        "shci_mlp_forward": ([pF, I32, I32, pF, I64, pF], I32),
        "shci_mlp_loss_grad": ([pF, I32, I32, pF, pF, I64, pF, pF], I32),
        "shci_mlp_train": ([pF, I32, I32, pF, pF, I64, I32, I32, ctypes.c_float, ctypes.c_uint64, pF, pF], I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class System:
    """One FCIDUMP on the device plus a variational space and its Hamiltonian."""

    def __init__(self, handle):
        self._h = handle
        self._lib = library()

    @classmethod
    def from_fcidump(cls, path):
        L = library()
        h = ctypes.c_void_p()
        st = L.shci_create(os.fsencode(path), ctypes.byref(h))
        if st != 0:
            raise ShciError(st, L.shci_last_error(None).decode())
        return cls(h)

    @classmethod
    def from_integrals(cls, h1, eri8, nelec, core_energy=0.0, orbsym=None):
        """From integrals in memory: h1 norb x norb, eri8 the 8-fold packed (pq|rs)
        (PySCF's ao2mo.restore(8, eri, norb) order)."""
        L = library()
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        eri8 = np.ascontiguousarray(eri8, dtype=np.float64)
        norb = h1.shape[0]
        npair = norb * (norb + 1) // 2
        if h1.shape != (norb, norb) or eri8.size != npair * (npair + 1) // 2:
            raise ValueError("h1 must be norb x norb and eri8 hold npair(npair+1)/2 values")
        sym = None if orbsym is None else np.ascontiguousarray(orbsym, dtype=np.int32)
        h = ctypes.c_void_p()
        st = L.shci_create_from_integrals(norb, int(nelec), _p(h1, ctypes.c_double), _p(eri8, ctypes.c_double),
                                          float(core_energy), None if sym is None else _p(sym, ctypes.c_int32),
                                          ctypes.byref(h))
        if st != 0:
            raise ShciError(st, L.shci_last_error(None).decode())
        return cls(h)

    def variational(self, dets, eps1, nroots=1, ci=None, davidson_tol=5e-5, davidson_tol_loose=5e-5, dE=1e-8,
                    seed=1, ncore=0, nact=1000000):
        """SHCIbasics::DoVariational on in-memory data (shci_do_variational): the
        iterations of the eps1 list from the space `dets`; returns the energies
        (core energy added).  The final space and records stay on the handle
        (dets(), run_ci(), run_iterations())."""
        d = self._dets(dets)
        eps = np.ascontiguousarray(eps1, dtype=np.float64)
        p = VarParams()
        p.nroots, p.n_iter, p.eps1 = nroots, len(eps), _p(eps, ctypes.c_double)
        p.davidson_tol, p.davidson_tol_loose, p.dE, p.seed = davidson_tol, davidson_tol_loose, dE, seed
        p.ncore, p.nact = ncore, nact
        c = None
        if ci is not None:
            c = np.asfortranarray(np.asarray(ci, dtype=np.float64).reshape(len(d), nroots))
        e = np.zeros(nroots)
        self._check(self._lib.shci_do_variational(self._h, ctypes.byref(p), _p(d, ctypes.c_int64), len(d),
                                                  None if c is None else _p(c, ctypes.c_double),
                                                  _p(e, ctypes.c_double)))
        return e

    def run_iterations(self):
        return VariationalResult.iterations(self)

    def run_energies(self):
        return VariationalResult.energies(self)

    def run_ci(self):
        """Final coefficients of the last run on this handle, n x nroots."""
        n = self.num_dets()
        nr = len(self.run_energies())
        out = np.zeros(n * nr)
        self._check(self._lib.shci_run_ci(self._h, _p(out, ctypes.c_double), len(out)))
        return out.reshape(nr, n).T

    def close(self):
        if self._h:
            self._lib.shci_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise ShciError(st, self._lib.shci_last_error(self._h).decode())

    norb = property(lambda self: self._lib.shci_norb(self._h))
    nelec = property(lambda self: self._lib.shci_nelec(self._h))
    words = property(lambda self: self._lib.shci_det_words(self._h))
    core_energy = property(lambda self: self._lib.shci_core_energy(self._h))

    def _dets(self, dets):
        d = np.ascontiguousarray(dets, dtype=np.int64)
        if d.ndim == 1:
            d = d.reshape(-1, self.words)
        if d.shape[1] != self.words:
            raise ValueError("determinants must have %d words" % self.words)
        return d

    def set_precision(self, prec):
        self._check(self._lib.shci_set_precision(self._h, prec))

    def set_options(self, precision=PREC_FP64, davidson_block=False, autotune=False, hip_graphs=False,
                    profile_sigma=False, direct_sigma=False):
        """Method options for the piecewise calls (see shci_run_options)."""
        o = RunOptions()
        o.direct_sigma = int(bool(direct_sigma))
        o.precision = precision
        o.profile_sigma = int(bool(profile_sigma))
        o.davidson_block = int(bool(davidson_block))
        o.autotune = int(bool(autotune))
        o.hip_graphs = int(bool(hip_graphs))
        self._check(self._lib.shci_set_options(self._h, ctypes.byref(o)))

    def set_active_space(self, ncore, nact):
        self._check(self._lib.shci_set_active_space(self._h, ncore, nact))

    def add_restriction(self, min_elec, max_elec, spatial_orbs):
        o = np.ascontiguousarray(spatial_orbs, dtype=np.int32)
        self._check(self._lib.shci_add_restriction(self._h, min_elec, max_elec, _p(o, ctypes.c_int32), len(o)))

    def energy(self, dets):
        d = self._dets(dets)
        out = np.zeros(len(d))
        self._check(self._lib.shci_det_energy(self._h, _p(d, ctypes.c_int64), len(d), _p(out, ctypes.c_double)))
        return out

    def hij(self, bra, ket):
        b, k = self._dets(bra), self._dets(ket)
        out = np.zeros(len(b))
        self._check(self._lib.shci_hij(self._h, _p(b, ctypes.c_int64), _p(k, ctypes.c_int64), len(b),
                                       _p(out, ctypes.c_double)))
        return out

    def set_dets(self, dets):
        d = self._dets(dets)
        self._check(self._lib.shci_set_dets(self._h, _p(d, ctypes.c_int64), len(d)))

    def num_dets(self):
        n = ctypes.c_int64()
        self._check(self._lib.shci_num_dets(self._h, ctypes.byref(n)))
        return n.value

    def dets(self):
        n = self.num_dets()
        out = np.zeros((n, self.words), dtype=np.int64)
        self._check(self._lib.shci_get_dets(self._h, _p(out, ctypes.c_int64), n))
        return out

    def get_determinants_variational(self, cmax, eps1):
        """New determinants (sorted, unique, not yet held) selected by heat-bath screening."""
        c = np.ascontiguousarray(cmax, dtype=np.float64)
        m = ctypes.c_int64()
        self._check(self._lib.shci_screen(self._h, _p(c, ctypes.c_double), eps1, ctypes.byref(m)))
        out = np.zeros((m.value, self.words), dtype=np.int64)
        self._check(self._lib.shci_get_new_dets(self._h, _p(out, ctypes.c_int64), m.value))
        return out

    def append_new(self):
        self._check(self._lib.shci_append_new(self._h))

    def make_hamiltonian(self):
        """Build H; returns the element count of the reference's lower-triangular rows."""
        nnz = ctypes.c_int64()
        self._check(self._lib.shci_build_hamiltonian(self._h, ctypes.byref(nnz)))
        return nnz.value

    def hamiltonian_rows(self, row_begin=0):
        """(rowptr, cols, vals) in the reference layout: diagonal first, then J < I ascending."""
        n = self.num_dets()
        nnz = self.make_hamiltonian()
        rp = np.zeros(n - row_begin + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz)
        self._check(self._lib.shci_get_hamiltonian_rows(self._h, row_begin, _p(rp, ctypes.c_int64),
                                                        _p(cols, ctypes.c_int32), _p(vals, ctypes.c_double), nnz))
        return rp, cols[:rp[-1]], vals[:rp[-1]]

    def hmult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        k = 1 if x.ndim == 1 else x.shape[1]
        xf = np.asfortranarray(x.reshape(len(x), k))
        y = np.zeros_like(xf, order="F")
        self._check(self._lib.shci_sigma(self._h, _p(xf, ctypes.c_double), _p(y, ctypes.c_double), k))
        return y.reshape(x.shape)

    def davidson(self, x0, max_copies=None, tol=5e-5):
        x = np.asfortranarray(np.array(x0, dtype=np.float64, ndmin=2).T if np.ndim(x0) == 1 else x0, dtype=np.float64)
        nroots = x.shape[1]
        e = np.zeros(nroots)
        it = ctypes.c_int()
        mc = max_copies or nroots + 4
        self._check(self._lib.shci_davidson(self._h, _p(x, ctypes.c_double), nroots, mc, tol,
                                            _p(e, ctypes.c_double), ctypes.byref(it)))
        return e, x, it.value


    def pt_deterministic(self, ci, e0, eps2):
        """Deterministic EN-PT2 of one root over the held determinants
        (SHCIbasics::DoPerturbativeDeterministic, SHCI/SHCIbasics.cpp:453-728):
        ci = the root's coefficients, e0 = its variational energy without the
        core energy.  Returns (E_PT, psi1 norm, number of external determinants)."""
        c = np.ascontiguousarray(ci, dtype=np.float64)
        if len(c) != self.num_dets():
            raise ValueError("need one coefficient per held determinant")
        e, psi, nx = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._check(self._lib.shci_pt_deterministic(self._h, _p(c, ctypes.c_double), e0, eps2, ctypes.byref(e),
                                                    ctypes.byref(psi), ctypes.byref(nx)))
        return e.value, psi.value, nx.value


    def rdm(self, bra, ket=None, spin=False):
        """RDMs of <bra| ... |ket> over the held determinants (SHCIrdm::EvaluateRDM
        and EvaluateOneRDM, SHCI/SHCIrdm.cpp:863-1100): (g1[P,R], g2[P,Q,R,S]) with
        g2[P,Q,R,S] = sum_st <a+_Ps a+_Qt a_St a_Rs> (the reference's
        s2RDM(P*n+Q, R*n+S)); with spin, also the spin-orbital 1-RDM and the
        pair-indexed spin 2-RDM (npair x npair, pair(p,q) = p(p+1)/2 + q, p > q)."""
        b = np.ascontiguousarray(bra, dtype=np.float64)
        k = b if ket is None else np.ascontiguousarray(ket, dtype=np.float64)
        if len(b) != self.num_dets() or len(k) != len(b):
            raise ValueError("need one coefficient per held determinant")
        n = self.norb
        nso, npair = 2 * n, n * (2 * n + 1)
        g1, g2 = np.zeros((n, n)), np.zeros((n,) * 4)
        s1 = np.zeros((nso, nso)) if spin else None
        s2 = np.zeros((npair, npair)) if spin else None
        self._check(self._lib.shci_rdm(self._h, _p(b, ctypes.c_double), _p(k, ctypes.c_double),
                                       _p(g1, ctypes.c_double), _p(g2, ctypes.c_double),
                                       _p(s1, ctypes.c_double) if spin else None,
                                       _p(s2, ctypes.c_double) if spin else None))
        return (g1, g2, s1, s2) if spin else (g1, g2)

    def pt_semistochastic(self, ci, e0, eps2, eps2_large, sample_n, n_iter, target_error, seed):
        """Semistochastic EN-PT2 of one root (SHCI/SHCIbasics.cpp:55-452): returns
        (the reference function's value, per-iteration estimates, iterations, standard error)."""
        c = np.ascontiguousarray(ci, dtype=np.float64)
        if len(c) != self.num_dets():
            raise ValueError("need one coefficient per held determinant")
        e, err, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        per = np.zeros(max(n_iter, 1))
        self._check(self._lib.shci_pt_semistochastic(self._h, _p(c, ctypes.c_double), e0, eps2, eps2_large, sample_n,
                                                     n_iter, target_error, seed, ctypes.byref(e),
                                                     _p(per, ctypes.c_double), len(per), ctypes.byref(it),
                                                     ctypes.byref(err)))
        return e.value, per[:it.value].copy(), it.value, err.value


class VariationalResult:
    def __init__(self, handle):
        self._h = handle
        self._lib = library()

    def __del__(self):
        try:
            self._lib.shci_destroy(self._h)
        except Exception:
            pass

    def iterations(self):
        out = []
        for i in range(self._lib.shci_run_iterations(self._h)):
            nd, e, it, t = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
            self._lib.shci_run_iteration(self._h, i, ctypes.byref(nd), ctypes.byref(e), ctypes.byref(it),
                                         ctypes.byref(t))
            out.append({"ndets": nd.value, "energy": e.value, "davidson": it.value, "seconds": t.value})
        return out

    def energies(self):
        e = np.zeros(64)
        n = ctypes.c_int()
        self._lib.shci_run_energies(self._h, _p(e, ctypes.c_double), ctypes.byref(n))
        return e[:n.value].copy()

    def pt_energies(self):
        """E_var + E_PT per root when the run did the deterministic PT step (empty otherwise)."""
        e = np.zeros(64)
        n = ctypes.c_int()
        self._lib.shci_run_pt_energies(self._h, _p(e, ctypes.c_double), ctypes.byref(n))
        return e[:n.value].copy()

    def ci(self):
        """Final coefficients, n x nroots."""
        n = ctypes.c_int64()
        self._lib.shci_num_dets(self._h, ctypes.byref(n))
        nr = len(self.energies())
        out = np.zeros(n.value * nr)
        self._lib.shci_run_ci(self._h, _p(out, ctypes.c_double), len(out))
        return out.reshape(nr, n.value).T

    def phases(self):
        """Per iteration: (screen s, Hamiltonian s, Davidson s, lower-triangle element count)."""
        out = []
        for i in range(self._lib.shci_run_iterations(self._h)):
            a, b, c, z = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
            self._lib.shci_run_iteration_phases(self._h, i, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                ctypes.byref(z))
            out.append((a.value, b.value, c.value, z.value))
        return out

    def learned(self):
        """Learned screening form: per ranked iteration {iteration, candidates, evaluated, accepted,
        sampled_recall, fallback, seconds}, and the training record."""
        L = self._lib
        out = []
        for k in range(L.shci_run_learned_count(self._h)):
            it, fb = ctypes.c_int(), ctypes.c_int()
            c, e, a = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            rc, sf, sec = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            L.shci_run_learned(self._h, k, ctypes.byref(it), ctypes.byref(c), ctypes.byref(e), ctypes.byref(a),
                               ctypes.byref(rc), ctypes.byref(sf), ctypes.byref(fb), ctypes.byref(sec))
            out.append({"iteration": it.value, "candidates": c.value, "evaluated": e.value, "accepted": a.value,
                        "sampled_recall": rc.value, "sampled_fraction": sf.value, "fallback": bool(fb.value),
                        "seconds": sec.value})
        n, f, l, sec = ctypes.c_int64(), ctypes.c_float(), ctypes.c_float(), ctypes.c_double()
        L.shci_run_learned_training(self._h, ctypes.byref(n), ctypes.byref(f), ctypes.byref(l), ctypes.byref(sec))
        return out, {"samples": n.value, "first_loss": f.value, "last_loss": l.value, "seconds": sec.value}

    def final_davidson_seconds(self):
        return self._lib.shci_run_final_davidson_seconds(self._h)

    def sigma_profile(self):
        """(calls, kernel milliseconds, algorithmic bytes) of the timed sigma launches."""
        c, ms, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self._lib.shci_profile_sigma_read(self._h, ctypes.byref(c), ctypes.byref(ms), ctypes.byref(b))
        return c.value, ms.value, b.value

    def dets(self):
        n = ctypes.c_int64()
        self._lib.shci_num_dets(self._h, ctypes.byref(n))
        w = self._lib.shci_det_words(self._h)
        out = np.zeros((n.value, w), dtype=np.int64)
        self._lib.shci_get_dets(self._h, _p(out, ctypes.c_int64), n.value)
        return out


def do_variational(input_path, log_path=None, precision=PREC_FP64, profile_sigma=False, davidson_block=False,
                   autotune=False, hip_graphs=False, learned=False, half_storage=None):
    """Run the variational step of ``Dice input.dat``; returns a VariationalResult.

    precision: PREC_FP64 (the drop-in) or PREC_FP32 (FP32 values, FP64
    accumulation and refinement); davidson_block / autotune / hip_graphs
    select the batching form, learned the learned screening form;
    half_storage True / False: the Hamiltonian's lower triangle alone / both
    triangles (None: the library's choice per form; see shci_run_options)."""
    L = library()
    h = ctypes.c_void_p()
    opt = RunOptions()
    opt.precision = precision
    opt.profile_sigma = 1 if profile_sigma else 0
    opt.davidson_block = 1 if davidson_block else 0
    opt.autotune = 1 if autotune else 0
    opt.hip_graphs = 1 if hip_graphs else 0
    opt.half_storage = 0 if half_storage is None else (1 if half_storage else -1)
    opt.learned = {"control": 2, "forced": 3}.get(learned, 1 if learned else 0) if isinstance(learned, str) \
        else (1 if learned else 0)
    st = L.shci_run_ex(os.fsencode(input_path), os.fsencode(log_path) if log_path else None, ctypes.byref(opt),
                       ctypes.byref(h))
    if st != 0:
        msg = L.shci_last_error(h).decode() if h else L.shci_last_error(None).decode()
        if h:
            L.shci_destroy(h)
        raise ShciError(st, msg)
    return VariationalResult(h)


# ---- candidate-ranking MLP of the learned screening form (no reference counterpart) ----

def _mlp_check(L, st):
    if st != 0:
        raise ShciError(st, L.shci_last_error(None).decode())


def mlp_param_count(din, hidden):
    return int(library().shci_mlp_param_count(din, hidden))


def mlp_init(din, hidden, seed=0):
    """Xavier-uniform parameters (float32, flat layout W1, b1, W2, b2, W3, b3)."""
    L = library()
    p = np.zeros(mlp_param_count(din, hidden), dtype=np.float32)
    _mlp_check(L, L.shci_mlp_init(_p(p, ctypes.c_float), din, hidden, seed))
    return p


def mlp_forward(params, din, hidden, X):
    L = library()
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, din)
    y = np.zeros(len(Xc), dtype=np.float32)
    _mlp_check(L, L.shci_mlp_forward(_p(P, ctypes.c_float), din, hidden, _p(Xc, ctypes.c_float), len(Xc),
                                     _p(y, ctypes.c_float)))
    return y


def mlp_loss_grad(params, din, hidden, X, t):
    L = library()
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, din)
    tc = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros_like(P)
    loss = ctypes.c_float()
    _mlp_check(L, L.shci_mlp_loss_grad(_p(P, ctypes.c_float), din, hidden, _p(Xc, ctypes.c_float),
                                       _p(tc, ctypes.c_float), len(Xc), ctypes.byref(loss), _p(g, ctypes.c_float)))
    return loss.value, g


def mlp_train(params, din, hidden, X, t, epochs, batch, lr=1e-3, seed=0):
    """Adam training on the GPU; returns (new params, first minibatch loss, last minibatch loss)."""
    L = library()
    P = np.array(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, din)
    tc = np.ascontiguousarray(t, dtype=np.float32)
    f, l = ctypes.c_float(), ctypes.c_float()
    _mlp_check(L, L.shci_mlp_train(_p(P, ctypes.c_float), din, hidden, _p(Xc, ctypes.c_float), _p(tc, ctypes.c_float),
                                   len(Xc), epochs, batch, lr, seed, ctypes.byref(f), ctypes.byref(l)))
    return P, f.value, l.value
