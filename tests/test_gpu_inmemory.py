"""The in-memory boundary: shci_create_from_integrals and shci_do_variational
take what SHCIbasics::DoVariational takes (SHCI/SHCIbasics.h:89-93:
integrals, determinants, coefficients, schedule) instead of the FCIDUMP and
input.dat files.  On the same integrals and schedule they give the
file-driven run bit for bit."""
import os

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _file_run(tmp_path, fcidump, occ, extra, nroots=1):
    from dice_amd import shci
    inp = tmp_path / "input.dat"
    inp.write_text("nocc %d\n%s\nend\norbitals %s\nnroots %d\nschedule\n0 1e-6\nend\nmaxiter 12\n"
                   "davidsonTol 1e-8\ndE 1e-10\nseed 1\nnoio\nsampleN 200\nnPTiter 0\nprefix %s\n%s"
                   % (len(occ), " ".join(map(str, occ)), fcidump, nroots, tmp_path, extra))
    return shci.do_variational(str(inp), str(tmp_path / "gpu.log"))


@pytest.mark.parametrize("nroots", [1, 2])
def test_in_memory_run_equals_file_run(tmp_path, nroots):
    from dice_amd import shci
    fcidump = os.path.join(GOLD, "h8_631g", "FCIDUMP")
    norb, nelec, h1, eri8, ecore, sym = helpers.read_fcidump_arrays(fcidump)
    occ = list(range(nelec))
    r = _file_run(tmp_path, fcidump, occ, "ncore 0\nnact 8\n", nroots)
    S = shci.System.from_integrals(h1, eri8, nelec, ecore, sym)
    assert (S.norb, S.nelec, S.core_energy) == (norb, nelec, ecore)
    hf = helpers.occ_to_det(occ, S.words)[None, :]
    e = S.variational(hf, [1e-6] * 12, nroots=nroots, davidson_tol=1e-8, davidson_tol_loose=5e-5, dE=1e-10, seed=1,
                      ncore=0, nact=8)
    np.testing.assert_array_equal(e, r.energies())
    np.testing.assert_array_equal(S.dets(), r.dets())
    np.testing.assert_array_equal(S.run_ci(), r.ci())
    assert [(x["ndets"], x["energy"], x["davidson"]) for x in S.run_iterations()] == \
        [(x["ndets"], x["energy"], x["davidson"]) for x in r.iterations()]
    if nroots == 1:
        import json
        with open(os.path.join(GOLD, "golden.json")) as f:
            g = json.load(f)
        assert abs(e[0] - g["h8_631g"]["final_energy"]) < 6e-11
    S.close()


def test_in_memory_run_continues_from_given_space(tmp_path):
    """Started from the space and coefficients an earlier call returned
    (DoVariational's in/out Dets and ci), a second call with a smaller eps1
    keeps the given determinants first, in their order, adds the selected
    ones after them, and lands within the Davidson tolerance's reach of one
    run whose schedule has both steps."""
    from dice_amd import shci
    fcidump = os.path.join(GOLD, "h8_631g", "FCIDUMP")
    norb, nelec, h1, eri8, ecore, sym = helpers.read_fcidump_arrays(fcidump)
    S = shci.System.from_integrals(h1, eri8, nelec, ecore, sym)
    hf = helpers.occ_to_det(list(range(nelec)), S.words)[None, :]
    S.variational(hf, [1e-3], dE=1e-10, seed=1)
    d1, c1 = S.dets(), S.run_ci()
    assert len(d1) > 1
    e2 = S.variational(d1, [1e-4], ci=c1, dE=1e-10, seed=1)
    d2 = S.dets()
    assert len(d2) > len(d1)
    np.testing.assert_array_equal(d2[:len(d1)], d1)
    T = shci.System.from_integrals(h1, eri8, nelec, ecore, sym)
    e = T.variational(hf, [1e-3, 1e-4], dE=1e-10, seed=1)
    assert abs(e[0] - e2[0]) < 1e-5
