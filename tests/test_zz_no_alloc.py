"""shci_sigma_device_noalloc: sigma with the fixed per-GPU scratch region in
place of the handle's buffers.  Its results equal shci_sigma_device's bit for
bit (same kernels, same summation order) and the oracle's sigma to 1e-12
relative; two calls in a row on spaces of different sizes (the second
smaller) must both be right, since the region keeps the previous call's
contents.  The first call on each handle is the allocation-free one, so it
cannot lean on buffers an earlier shci_sigma_device left behind."""
import ctypes
import os

import numpy as np
import pytest

from oracle_bind import OracleSystem

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _space(path, occ, n, W):
    import helpers
    o = OracleSystem(path)
    dets = helpers.occ_to_det(occ, W)[None, :]
    eps = 1e-2
    while len(dets) < n:
        new = o.screen(dets, np.ones(len(dets)), eps)
        if len(new):
            dets = np.concatenate([dets, new])
        eps *= 0.5
    return dets[:n]


def _hip():
    """The HIP runtime, for device buffers of the test (the product calls
    allocate nothing)."""
    L = ctypes.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipFree.argtypes = [ctypes.c_void_p]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


def _dev(H, nbytes):
    p = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def _sigma_pair(S, x):
    """(noalloc result, shci_sigma_device result) for the n x k matrix x."""
    H = _hip()
    n, k = x.shape
    xf = np.asfortranarray(x, dtype=np.float64)
    nb = n * k * 8
    dx, dy1, dy2 = _dev(H, nb), _dev(H, nb), _dev(H, nb)
    try:
        assert H.hipMemcpy(dx, xf.ctypes.data, nb, 1) == 0  # host to device
        nan = np.full((n, k), np.nan, order="F")
        for d in (dy1, dy2):
            assert H.hipMemcpy(d, nan.ctypes.data, nb, 1) == 0
        L = S._lib
        st = L.shci_sigma_device_noalloc(S._h, dx, dy1, k, n)
        assert st == 0, L.shci_last_error(S._h)
        assert H.hipDeviceSynchronize() == 0
        st = L.shci_sigma_device(S._h, dx, dy2, k, n)
        assert st == 0, L.shci_last_error(S._h)
        assert H.hipDeviceSynchronize() == 0
        out = []
        for d in (dy1, dy2):
            y = np.zeros((n, k), order="F")
            assert H.hipMemcpy(y.ctypes.data, d, nb, 2) == 0  # device to host
            out.append(y)
        return out[0], out[1]
    finally:
        for d in (dx, dy1, dy2):
            H.hipFree(d)


@pytest.mark.parametrize("storage", ["full", "half"])
@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_sigma_matches_allocating_sigma(prec, storage, tmp_path, monkeypatch):
    from dice_amd import shci, synth
    # both triangles, or the tiled copy of the stored half (fixed-point
    # transposed sums)
    monkeypatch.setenv("SHCI_STORAGE", storage)
    big = str(tmp_path / "c0")
    synth.write_molecular_fcidump(big, 0, 16, 8, 8, 64)
    small = str(tmp_path / "w2")  # 34 orbitals: two-word determinants
    synth.write_molecular_fcidump(small, 5, 34, 6, 6, 12)
    cases = [(big, synth.hf_occupation(8, 8), 3000, 1, 3), (small, synth.hf_occupation(6, 6), 1200, 2, 1)]
    for path, occ, n, W, k in cases:  # the second call is on the smaller space
        print("noalloc sigma: %s n=%d k=%d precision=%d storage=%s" % (os.path.basename(path), n, k, prec, storage),
              flush=True)
        dets = _space(path, occ, n, W)
        S = shci.System.from_fcidump(path)
        S.set_precision(prec)
        S.set_dets(dets)
        S.make_hamiltonian()
        x = np.random.default_rng(n).normal(size=(len(dets), k))
        y_fixed, y_alloc = _sigma_pair(S, x)
        np.testing.assert_array_equal(y_fixed, y_alloc)
        if prec == 0:
            o = OracleSystem(path)
            o.csr_append(dets)
            for r in range(k):
                yo = o.sigma(np.ascontiguousarray(x[:, r]))
                assert np.max(np.abs(y_fixed[:, r] - yo)) <= 1e-12 * np.max(np.abs(yo))
        S.close()
