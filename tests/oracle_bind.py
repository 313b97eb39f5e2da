"""ctypes binding of the CPU oracle (oracle/build/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it.  It is
built from oracle/ with make when missing.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")
EXE = os.path.join(ORACLE_DIR, "build", "shci_ref")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not (os.path.exists(LIB) and os.path.exists(EXE)):
        build()
    L = ctypes.CDLL(LIB)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    pI64, pD = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)
    sig = {
        "oref_last_error": ([], ctypes.c_char_p),
        "oref_create": ([ctypes.c_char_p], P),
        "oref_destroy": ([P], None),
        "oref_norb": ([P], I32),
        "oref_nelec": ([P], I32),
        "oref_det_words": ([P], I32),
        "oref_core_energy": ([P], D),
        "oref_max_entry": ([P], D),
        "oref_i2_size": ([P], I64),
        "oref_get_integrals": ([P, pD, pD, pD, pD], None),
        "oref_hb_size": ([P, I32], I64),
        "oref_hb_get": ([P, I32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int16), pI64], None),
        "oref_singles": ([P, pD], None),
        "oref_set_active": ([P, I32, I32], None),
        "oref_energy": ([P, pI64], D),
        "oref_hij": ([P, pI64, pI64], D),
        "oref_screen": ([P, pI64, I64, pD, D, pI64, I64, pI64, I64], I64),
        "oref_csr_append": ([P, pI64, I64], I64),
        "oref_csr_get": ([P, pI64, ctypes.POINTER(ctypes.c_int32), pD], None),
        "oref_sigma": ([P, pD, pD], None),
        "oref_sigma_direct": ([P, pD, pD], None),
        "oref_rdm": ([P, pD, pD, pD, pD, pD, pD], None),
        "oref_davidson": ([P, pD, I32, pD, I32, D, pD], I32),
        "oref_sym_eig": ([I32, pD, pD, pD], None),
        "oref_alias_draw_index": ([I32, I32], I32),
        "oref_rand_max": ([], I32),
        "oref_run": ([ctypes.c_char_p, ctypes.c_char_p, I32], P),
        "oref_run_status": ([P], I32),
        "oref_run_error": ([P], ctypes.c_char_p),
        "oref_run_niter": ([P], I32),
        "oref_run_nroots": ([P], I32),
        "oref_run_iter": ([P, I32, pI64, pD, ctypes.POINTER(ctypes.c_int)], None),
        "oref_run_energies": ([P, pD], None),
        "oref_run_ndets": ([P], I64),
        "oref_run_iter_ndets": ([P, I32], I64),
        "oref_run_dets": ([P, I32, pI64, I32], None),
        "oref_run_ci": ([P, pD], None),
        "oref_run_free": ([P], None),
        "oref_run_pt": ([P, pD], I32),
        "oref_add_pt_restriction": ([P, I32, I32, ctypes.POINTER(ctypes.c_int), I32], None),
        "oref_pt_deterministic": ([P, pI64, I64, pD, D, D, pD, pI64], D),
        "oref_pt_semistochastic": ([P, pI64, I64, pD, D, D, D, I32, I32, D, ctypes.c_uint, pD, I32,
                                    ctypes.POINTER(ctypes.c_int), pD], D),
        "oref_mlp_forward": ([ctypes.POINTER(ctypes.c_float), I32, I32, ctypes.POINTER(ctypes.c_float), I64, I32,
                              pD], None),
        "oref_mlp_loss_grad": ([ctypes.POINTER(ctypes.c_float), I32, I32, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), I64, I32, pD], D),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class OracleSystem:
    def __init__(self, path):
        self.L = load()
        self.h = self.L.oref_create(os.fsencode(path))
        if not self.h:
            raise RuntimeError(self.L.oref_last_error().decode())
        self.words = self.L.oref_det_words(self.h)
        self.norb = self.L.oref_norb(self.h)
        self.nelec = self.L.oref_nelec(self.h)
        self._n = 0

    def __del__(self):
        try:
            self.L.oref_destroy(self.h)
        except Exception:
            pass

    def set_active(self, ncore, nact):
        self.L.oref_set_active(self.h, ncore, nact)

    def energy(self, dets):
        d = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.words)
        return np.array([self.L.oref_energy(self.h, _p(d[i], ctypes.c_int64)) for i in range(len(d))])

    def hij(self, bra, ket):
        b = np.ascontiguousarray(bra, dtype=np.int64).reshape(-1, self.words)
        k = np.ascontiguousarray(ket, dtype=np.int64).reshape(-1, self.words)
        return np.array([self.L.oref_hij(self.h, _p(b[i], ctypes.c_int64), _p(k[i], ctypes.c_int64))
                         for i in range(len(b))])

    def screen(self, dets, cmax, eps1, sorted_dets=None):
        d = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.words)
        if sorted_dets is None:
            sorted_dets = sort_dets(d)
        s = np.ascontiguousarray(sorted_dets, dtype=np.int64).reshape(-1, self.words)
        c = np.ascontiguousarray(cmax, dtype=np.float64)
        cap = 1 << 16
        while True:
            out = np.zeros((cap, self.words), dtype=np.int64)
            m = self.L.oref_screen(self.h, _p(d, ctypes.c_int64), len(d), _p(c, ctypes.c_double), eps1,
                                   _p(s, ctypes.c_int64), len(s), _p(out, ctypes.c_int64), cap)
            if m <= cap:
                return out[:m]
            cap = m

    def csr_append(self, dets):
        d = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.words)
        nnz = self.L.oref_csr_append(self.h, _p(d, ctypes.c_int64), len(d))
        self._n = len(d)
        self._nnz = nnz
        return nnz

    def csr(self):
        rp = np.zeros(self._n + 1, dtype=np.int64)
        cols = np.zeros(self._nnz, dtype=np.int32)
        vals = np.zeros(self._nnz)
        self.L.oref_csr_get(self.h, _p(rp, ctypes.c_int64), _p(cols, ctypes.c_int32), _p(vals, ctypes.c_double))
        return rp, cols, vals

    def sigma(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.L.oref_sigma(self.h, _p(x, ctypes.c_double), _p(y, ctypes.c_double))
        return y

    def sigma_direct(self, x):
        """HmultDirect: every element evaluated on the fly, Hij(D_I, D_J) for row I."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.L.oref_sigma_direct(self.h, _p(x, ctypes.c_double), _p(y, ctypes.c_double))
        return y

    def rdm(self, bra, ket=None, spin=False):
        """SHCIrdm::EvaluateOneRDM / EvaluateRDM over the appended determinants:
        (g1, g2[P,Q,R,S]) and with spin the spin-orbital 1-RDM and the
        pair-indexed spin 2-RDM."""
        b = np.ascontiguousarray(bra, dtype=np.float64)
        k = b if ket is None else np.ascontiguousarray(ket, dtype=np.float64)
        n, nso = self.norb, 2 * self.norb
        npair = nso * (nso + 1) // 2
        g1, g2 = np.zeros((n, n)), np.zeros((n,) * 4)
        s1 = np.zeros((nso, nso)) if spin else None
        s2 = np.zeros((npair, npair)) if spin else None
        self.L.oref_rdm(self.h, _p(b, ctypes.c_double), _p(k, ctypes.c_double), _p(g1, ctypes.c_double),
                        _p(g2, ctypes.c_double), _p(s1, ctypes.c_double) if spin else None,
                        _p(s2, ctypes.c_double) if spin else None)
        return (g1, g2, s1, s2) if spin else (g1, g2)

    def add_pt_restriction(self, min_elec, max_elec, spatial_orbs):
        o = np.ascontiguousarray(spatial_orbs, dtype=np.int32)
        self.L.oref_add_pt_restriction(self.h, min_elec, max_elec, o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       len(o))

    def pt_deterministic(self, dets, ci, e0, eps2):
        """Deterministic EN-PT2 of one root: (E_PT, psi1 norm, number of external determinants)."""
        d = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.words)
        c = np.ascontiguousarray(ci, dtype=np.float64)
        psi, nx = ctypes.c_double(), ctypes.c_int64()
        e = self.L.oref_pt_deterministic(self.h, _p(d, ctypes.c_int64), len(d), _p(c, ctypes.c_double), e0, eps2,
                                         ctypes.byref(psi), ctypes.byref(nx))
        return e, psi.value, nx.value

    def pt_semistochastic(self, dets, ci, e0, eps2, eps2_large, sample_n, n_iter, target_error, seed):
        """Semistochastic EN-PT2 of one root: (returned value, per-iteration
        estimates, iterations, standard error)."""
        d = np.ascontiguousarray(dets, dtype=np.int64).reshape(-1, self.words)
        c = np.ascontiguousarray(ci, dtype=np.float64)
        per = np.zeros(max(n_iter, 1))
        it, err = ctypes.c_int(), ctypes.c_double()
        r = self.L.oref_pt_semistochastic(self.h, _p(d, ctypes.c_int64), len(d), _p(c, ctypes.c_double), e0, eps2,
                                          eps2_large, sample_n, n_iter, target_error, seed, _p(per, ctypes.c_double),
                                          len(per), ctypes.byref(it), ctypes.byref(err))
        return r, per[:it.value].copy(), it.value, err.value

    def davidson(self, x0, diag, max_copies, tol):
        x = np.array(np.atleast_2d(np.asarray(x0, dtype=np.float64).T).T, order="F")
        if x.ndim == 1:
            x = x[:, None]
        nroots = x.shape[1]
        xx = np.ascontiguousarray(x.T)  # root-major = column-major n x nroots
        d = np.ascontiguousarray(diag, dtype=np.float64)
        e = np.zeros(nroots)
        it = self.L.oref_davidson(self.h, _p(xx, ctypes.c_double), nroots, _p(d, ctypes.c_double), max_copies, tol,
                                  _p(e, ctypes.c_double))
        return e, xx.T, it


class OracleRun:
    def __init__(self, input_path, log_path=None, record=False):
        self.L = load()
        self.h = self.L.oref_run(os.fsencode(input_path), os.fsencode(log_path) if log_path else None,
                                 1 if record else 0)
        st = self.L.oref_run_status(self.h)
        if st != 0:
            msg = self.L.oref_run_error(self.h).decode()
            raise RuntimeError("oracle run failed (%d): %s" % (st, msg))

    def __del__(self):
        try:
            self.L.oref_run_free(self.h)
        except Exception:
            pass

    def iterations(self):
        out = []
        nr = max(self.L.oref_run_nroots(self.h), 1)
        for i in range(self.L.oref_run_niter(self.h)):
            nd, e, it = ctypes.c_int64(), np.zeros(nr), ctypes.c_int()
            self.L.oref_run_iter(self.h, i, ctypes.byref(nd), _p(e, ctypes.c_double), ctypes.byref(it))
            out.append({"ndets": nd.value, "energy": float(e[0]), "energies": e.copy(), "davidson": it.value})
        return out

    def pt_energies(self):
        """E_var + E_PT per root when the run did the deterministic PT step (empty otherwise)."""
        e = np.zeros(64)
        n = self.L.oref_run_pt(self.h, _p(e, ctypes.c_double))
        return e[:n].copy()

    def energies(self):
        e = np.zeros(self.L.oref_run_nroots(self.h))
        self.L.oref_run_energies(self.h, _p(e, ctypes.c_double))
        return e

    def dets(self, words, it=-1):
        n = self.L.oref_run_ndets(self.h) if it < 0 else self.L.oref_run_iter_ndets(self.h, it)
        out = np.zeros((n, words), dtype=np.int64)
        self.L.oref_run_dets(self.h, it, _p(out, ctypes.c_int64), words)
        return out

    def ci(self):
        n = self.L.oref_run_ndets(self.h)
        nr = self.L.oref_run_nroots(self.h)
        out = np.zeros(n * nr)
        self.L.oref_run_ci(self.h, _p(out, ctypes.c_double))
        return out.reshape(nr, n).T


def sort_dets(d):
    """Sort determinants in the reference's signed word order (highest word first)."""
    d = np.asarray(d, dtype=np.int64)
    keys = [d[:, w] for w in range(d.shape[1])]  # lexsort: last key primary
    return d[np.lexsort(keys)]


def mlp_forward(params, din, hidden, X, bf16=True):
    """The oracle's MLP forward pass (double; bf16-rounded matmul operands when bf16)."""
    L = load()
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, din)
    y = np.zeros(len(Xc))
    L.oref_mlp_forward(_p(P, ctypes.c_float), din, hidden, _p(Xc, ctypes.c_float), len(Xc), 1 if bf16 else 0,
                       _p(y, ctypes.c_double))
    return y


def mlp_loss_grad(params, din, hidden, X, t, bf16=True):
    L = load()
    P = np.ascontiguousarray(params, dtype=np.float32)
    Xc = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, din)
    tc = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(len(P))
    loss = L.oref_mlp_loss_grad(_p(P, ctypes.c_float), din, hidden, _p(Xc, ctypes.c_float), _p(tc, ctypes.c_float),
                                len(Xc), 1 if bf16 else 0, _p(g, ctypes.c_double))
    return loss, g


def spatial_integrals(o):
    """(h1 n x n, eri n^4 chemist (ij|kl), core) of an OracleSystem."""
    L = o.L
    n = L.oref_norb(o.h)
    nso = 2 * n
    I1 = np.zeros(nso * nso)
    I2 = np.zeros(L.oref_i2_size(o.h))
    L.oref_get_integrals(o.h, _p(I1, ctypes.c_double), _p(I2, ctypes.c_double), None, None)
    h1 = I1.reshape(nso, nso)[0::2, 0::2].copy()

    def pair(i, j):
        return max(i, j) * (max(i, j) + 1) // 2 + min(i, j)

    eri = np.zeros((n,) * 4)
    for i in range(n):
        for j in range(n):
            ij = pair(i, j)
            for k in range(n):
                for l in range(n):
                    kl = pair(k, l)
                    A, B = max(ij, kl), min(ij, kl)
                    eri[i, j, k, l] = I2[A * (A + 1) // 2 + B]
    return h1, eri, L.oref_core_energy(o.h)
