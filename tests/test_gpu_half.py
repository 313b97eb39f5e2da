"""Half storage (each connected pair once, in row max(I, J): the reference's
lower-triangular rows without the diagonal, SHCI/SHCImakeHamiltonian.cpp:549-711),
the layout for spaces whose two-triangle Hamiltonian does not fit.  Sigma
applies every stored element to both rows as Hmult2 does
(SHCI/Hmult.h:110-121), the transposed half by atomic adds (by default on a
tiled copy of the stored half, in 64-bit fixed point, so in a fixed result;
or in FP64 on the stored rows).  Bar: the rows
equal the two-triangle build's lower part bit for bit; sigma equals the
oracle's to 1e-12 relative (its sums are in no fixed order); runs give the
oracle's determinant sets and energies to 1e-10 Ha."""
import os

import numpy as np
import pytest

import helpers
from oracle_bind import OracleRun, OracleSystem

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _space(path, occ, n, W):
    o = OracleSystem(path)
    dets = helpers.occ_to_det(occ, W)[None, :]
    eps = 1e-2
    while len(dets) < n:
        new = o.screen(dets, np.ones(len(dets)), eps)
        if len(new):
            dets = np.concatenate([dets, new])
        eps *= 0.5
    return dets[:n]


@pytest.mark.parametrize("mode", ["tiled", "rows"])
@pytest.mark.parametrize("case", ["c0", "w2"])
def test_half_storage_rows_and_sigma(case, mode, tmp_path, monkeypatch):
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    if case == "c0":
        synth.write_molecular_fcidump(path, 0, 16, 8, 8, 64)
        occ, W = synth.hf_occupation(8, 8), 1
    else:
        synth.write_molecular_fcidump(path, 5, 34, 6, 6, 12)
        occ, W = synth.hf_occupation(6, 6), 2
    dets = _space(path, occ, 3000, W)
    monkeypatch.setenv("SHCI_STORAGE", "full")
    full = shci.System.from_fcidump(path)
    full.set_dets(dets)
    rp, cols, vals = full.hamiltonian_rows()
    monkeypatch.setenv("SHCI_STORAGE", "half")
    if mode == "rows":
        monkeypatch.setenv("SHCI_HALF_SIGMA", "rows")
    half = shci.System.from_fcidump(path)
    half.set_dets(dets)
    hrp, hcols, hvals = half.hamiltonian_rows()
    np.testing.assert_array_equal(rp, hrp)
    np.testing.assert_array_equal(cols, hcols)
    np.testing.assert_array_equal(vals, hvals)
    o = OracleSystem(path)
    o.csr_append(dets)
    x = np.random.default_rng(3).normal(size=(len(dets), 3))
    y = half.hmult(x)
    for r in range(3):
        yo = o.sigma(np.ascontiguousarray(x[:, r]))
        assert np.max(np.abs(y[:, r] - yo)) <= 1e-12 * np.max(np.abs(yo))


@pytest.mark.parametrize("case", ["c0", "w2"])
def test_half_storage_tiled_sigma_is_deterministic(case, tmp_path, monkeypatch):
    """The tiled half-storage sigma sums the transposed half in 64-bit fixed
    point: repeated calls, and a copy whose every segment sums in LDS first
    (a different split between LDS and direct integer atomics), give the
    same bits, within 1e-12 of the oracle."""
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    if case == "c0":
        synth.write_molecular_fcidump(path, 0, 16, 8, 8, 64)
        occ, W = synth.hf_occupation(8, 8), 1
    else:
        synth.write_molecular_fcidump(path, 5, 34, 6, 6, 12)
        occ, W = synth.hf_occupation(6, 6), 2
    dets = _space(path, occ, 3000, W)
    x = np.random.default_rng(5).normal(size=(len(dets), 2))
    x[7, 0] = 50.0  # one large value sets the scale of the fixed point
    monkeypatch.setenv("SHCI_STORAGE", "half")
    ys = []
    for stage_min in ("4", "1"):
        monkeypatch.setenv("SHCI_HALF_STAGE_MIN", stage_min)
        h = shci.System.from_fcidump(path)
        h.set_dets(dets)
        h.make_hamiltonian()
        ys += [h.hmult(x), h.hmult(x)]
        del h
    for y in ys[1:]:
        np.testing.assert_array_equal(y, ys[0])
    o = OracleSystem(path)
    o.csr_append(dets)
    for r in range(2):
        yo = o.sigma(np.ascontiguousarray(x[:, r]))
        assert np.max(np.abs(ys[0][:, r] - yo)) <= 1e-12 * np.max(np.abs(yo))


def test_half_storage_tiled_sigma_scales(tmp_path, monkeypatch):
    """The fixed point follows the scale of x: sigma of x scaled by 2^-600,
    1e-150, 1e150 and 2^600 equals the oracle's to 1e-12 relative (a power
    of two scales the result exactly), and x = 0 gives y = 0."""
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 0, 16, 8, 8, 64)
    dets = _space(path, synth.hf_occupation(8, 8), 3000, 1)
    monkeypatch.setenv("SHCI_STORAGE", "half")
    h = shci.System.from_fcidump(path)
    h.set_dets(dets)
    h.make_hamiltonian()
    o = OracleSystem(path)
    o.csr_append(dets)
    x = np.random.default_rng(11).normal(size=len(dets))
    y1 = h.hmult(x)
    for scale in (2.0 ** -600, 1e-150, 1e150, 2.0 ** 600):
        xs = x * scale
        y = h.hmult(xs)
        yo = o.sigma(np.ascontiguousarray(xs))
        assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo)), scale
        if scale in (2.0 ** -600, 2.0 ** 600):
            np.testing.assert_array_equal(y, y1 * scale)
    np.testing.assert_array_equal(h.hmult(np.zeros(len(dets))), np.zeros(len(dets)))


@pytest.mark.parametrize("mode", ["tiled", "rows"])
@pytest.mark.parametrize("prec", [0, 1])
def test_half_storage_run_matches_oracle(prec, mode, tmp_path, monkeypatch):
    """A whole schedule (incremental builds append new rows in place) on a
    12-orbital system: the oracle's determinant set after every iteration and
    its energies (FP64); the FP32 form within 1e-4 Ha with the same space."""
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 9, 12, 4, 4, 32)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(4, 4), [(0, 3e-3), (2, 1e-3)], 5,
                      extra=["prefix %s" % tmp_path])
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    monkeypatch.setenv("SHCI_STORAGE", "half")
    if mode == "rows":
        monkeypatch.setenv("SHCI_HALF_SIGMA", "rows")
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"), precision=prec)
    assert [x["ndets"] for x in r.iterations()] == [x["ndets"] for x in o.iterations()]
    np.testing.assert_array_equal(r.dets(), o.dets(1))
    tol = 1e-10 if prec == 0 else 1e-4
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=tol)
