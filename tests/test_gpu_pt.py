"""Deterministic EN-PT2 on the GPU (SHCI/SHCIbasics.cpp:453-728) against the
CPU oracle, through the C ABI.

Bar: E_PT and the psi1 norm to 1e-12 relative (the numerators are
bit-identical; only the order of the final sum differs), the same number of
external determinants; whole runs to 1e-10 Ha, and the reference's trusted
o2_det_direct energy (E_var + E_PT) to 2e-9 Ha.
"""
import os
import struct

import numpy as np
import pytest

import helpers
from oracle_bind import OracleRun, OracleSystem
from test_gpu_parity import GOLD, _sample, systems  # noqa: F401  (module fixture)

pytestmark = pytest.mark.gpu


def _case(name, systems, n, seed):
    base, near = _sample(name, systems, n, seed)
    dets = np.concatenate([base, near])
    dets = dets[np.unique(dets, axis=0, return_index=True)[1]]
    rng = np.random.default_rng(seed)
    ci = rng.normal(0, 1, len(dets))
    ci /= np.linalg.norm(ci)
    return dets, ci


def _both(name, systems, dets, ci, eps2, restrict=None):
    from dice_amd import shci
    path = systems[name][0]
    g, o = shci.System.from_fcidump(path), OracleSystem(path)
    if restrict:
        g.add_restriction(*restrict)
        o.add_pt_restriction(*restrict)
    e0 = float(np.min(o.energy(dets))) - 0.5
    g.set_dets(dets)
    return g.pt_deterministic(ci, e0, eps2), o.pt_deterministic(dets, ci, e0, eps2)


@pytest.mark.parametrize("name,eps2", [("h8", 1e-6), ("c0", 1e-4), ("w2", 1e-4), ("hub", 1e-4)])
def test_pt_matches_oracle(name, eps2, systems):
    dets, ci = _case(name, systems, 200, 17)
    (eg, pg, ng), (eo, po, no) = _both(name, systems, dets, ci, eps2)
    print(name, "E_PT gpu %.15e oracle %.15e external %d" % (eg, eo, no))
    assert no > 100 and ng == no
    assert abs(eg - eo) <= 1e-12 * abs(eo)
    assert abs(pg - po) <= 1e-12 * abs(po)


def test_pt_partitions_agree(systems, monkeypatch):
    """Hash partitions (forced to 3 passes) give the single-pass result."""
    from dice_amd import shci
    dets, ci = _case("c0", systems, 200, 23)
    g = shci.System.from_fcidump(systems["c0"][0])
    g.set_dets(dets)
    e0 = float(np.min(g.energy(dets))) - 0.5
    one = g.pt_deterministic(ci, e0, 1e-4)
    monkeypatch.setenv("SHCI_PT_PARTITIONS", "3")
    three = g.pt_deterministic(ci, e0, 1e-4)
    assert three[2] == one[2]
    assert abs(three[0] - one[0]) <= 1e-12 * abs(one[0])


def test_pt_restrictions_match_oracle(systems):
    """`restrict` limits the PT space too (SHCI/input.cpp:267-276): at most 6
    electrons in the first four spatial orbitals of c0."""
    dets, ci = _case("c0", systems, 150, 29)
    free = _both("c0", systems, dets, ci, 1e-4)[1]
    (eg, pg, ng), (eo, po, no) = _both("c0", systems, dets, ci, 1e-4, restrict=(0, 6, [0, 1, 2, 3]))
    assert ng == no and no < free[2]
    assert abs(eg - eo) <= 1e-12 * abs(eo)


def _o2_input(tmp_path, extra=""):
    from test_oracle_pins import O2_INPUT
    fcidump = os.path.join(GOLD, "o2", "FCIDUMP")
    p = tmp_path / "in.dat"
    p.write_text(O2_INPUT + "orbitals %s\nprefix %s\n" % (fcidump, tmp_path) + extra)
    return str(p)


def test_run_o2_direct_pt_trusted(tmp_path, golden_dir):
    """tests/SHCI/o2_det_direct: direct Davidson leaves a smaller space, the PT
    step adds -2.28e-5 Ha, and shci.e = E_var + E_PT reproduces the stored
    trusted_hci.e to 2e-9 Ha."""
    import json
    from dice_amd import shci
    g = json.load(open(os.path.join(golden_dir, "golden.json")))
    inp = _o2_input(tmp_path, "directdavidson\n")
    r = shci.do_variational(inp, str(tmp_path / "log"))
    ept = r.pt_energies()
    assert len(ept) == 1
    assert ept[0] - r.energies()[0] < -2e-5
    assert abs(ept[0] - g["trusted_energy"]["o2_det_direct"]) < 2e-9
    with open(tmp_path / "shci.e", "rb") as f:
        assert struct.unpack("<d", f.read(8))[0] == ept[0]
    log = open(tmp_path / "log").read()
    assert "PERTURBATION THEORY STEP" in log and "PTEnergy: %18.10f" % ept[0] in log


def test_run_pt_two_roots_matches_oracle(tmp_path):
    """A two-root run with `deterministic` and epsilon2 1e-5 on an 8-orbital
    synthetic molecule: per-root E_var + E_PT equal the oracle's run."""
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 5, 8, 4, 4, 32)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(4, 4), [(0, 2e-2)], 3, nroots=2,
                      extra=["deterministic", "epsilon2 1e-5", "prefix %s" % tmp_path])
    r = shci.do_variational(inp, str(tmp_path / "g.log"))
    o = OracleRun(inp, str(tmp_path / "o.log"))
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-10)
    assert len(r.pt_energies()) == 2
    np.testing.assert_allclose(r.pt_energies(), o.pt_energies(), rtol=0, atol=1e-10)
    assert np.all(r.pt_energies() < r.energies())


# ---- semistochastic EN-PT2 (SHCI/SHCIbasics.cpp:55-452) ----

@pytest.mark.parametrize("name,eps2,large", [("c0", 1e-4, 1e-3), ("w2", 1e-4, 1000.0), ("hub", 1e-4, 1e-3)])
def test_pt_semistochastic_matches_oracle(name, eps2, large, systems):
    """Same seed, same draws (the C library's rand() on the host in both):
    every iteration's estimate agrees with the oracle's to 1e-12 relative
    (the per-determinant sums are bit-identical; only the final sum's order
    differs), and so do the iteration count and the returned value."""
    from dice_amd import shci
    dets, ci = _case(name, systems, 200, 31)
    path = systems[name][0]
    g, o = shci.System.from_fcidump(path), OracleSystem(path)
    e0 = float(np.min(o.energy(dets))) - 0.5
    g.set_dets(dets)
    # (a negative target error runs all 12 iterations: when every term passes
    # both thresholds the estimates are all equal and the error is 0 only up
    # to rounding)
    rg, pg, ig, sg = g.pt_semistochastic(ci, e0, eps2, large, 50, 12, -1.0, 11)
    ro, po, io_, so = o.pt_semistochastic(dets, ci, e0, eps2, large, 50, 12, -1.0, 11)
    assert ig == io_ == 12
    np.testing.assert_allclose(pg, po, rtol=1e-12, atol=0)
    assert abs(rg - ro) <= 1e-12 * abs(ro)


def test_run_semistochastic_pt_matches_oracle(tmp_path):
    """A whole run with the semistochastic PT step on the 8-orbital synthetic
    molecule: E_var and E_var + E_PT equal the oracle's run."""
    from dice_amd import shci, synth
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 5, 8, 4, 4, 32)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(4, 4), [(0, 2e-2)], 3,
                      extra=["nPTiter 200", "epsilon2 1e-4", "epsilon2Large 1e-3", "targetError 6e-4",
                             "prefix %s" % tmp_path])
    r = shci.do_variational(inp, str(tmp_path / "g.log"))
    o = OracleRun(inp, str(tmp_path / "o.log"))
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-10)
    np.testing.assert_allclose(r.pt_energies(), o.pt_energies(), rtol=0, atol=1e-9)
    assert "Semistochastic PT calculation converged" in open(tmp_path / "g.log").read()
