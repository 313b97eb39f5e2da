"""Time reversal (Treversal; SHCI/Determinants.cpp:32-121,
SHCI/SHCImakeHamiltonian.cpp:344-483, SHCI/SHCIbasics.cpp:730-776,984-996) on
the GPU: standard forms in the screening, sigma in the adapted basis as
U^T H_E U over the plain Hamiltonian of the determinants and their flipped
copies, and the unpacked state.

Bar: the reference's trusted energies of tests/SHCI/o2_det_trev,
o2_det_trev_direct and restart_trev (1e-9 / 2e-9 Ha); per iteration the
oracle's determinant counts and energies (1e-10 Ha; the oracle stores the
reference's elements, the GPU forms them from the plain ones); the unpacked
vectors in the +1 space of the spin flip.
"""
import json
import os
import struct

import numpy as np
import pytest

from oracle_bind import OracleRun
from test_gpu_parity import GOLD

pytestmark = pytest.mark.gpu


def _gold():
    return json.load(open(os.path.join(GOLD, "golden.json")))


def _input(text, fcidump, prefix, drop=(), extra=""):
    keep = [l for l in text.splitlines()
            if not any(l.strip().lower().startswith(k.lower()) for k in ("orbitals", "prefix") + tuple(drop))]
    return "\n".join(keep) + "\norbitals %s\nprefix %s\n%s" % (fcidump, prefix, extra)


def _both(tmp_path, text, name="in.dat"):
    from dice_amd import shci
    res = {}
    for who in ("gpu", "oracle"):
        d = tmp_path / who
        d.mkdir(exist_ok=True)
        p = d / name
        p.write_text(text.replace("@PREFIX@", str(d)))
        if who == "gpu":
            res[who] = shci.do_variational(str(p), str(d / "log"))
        else:
            res[who] = OracleRun(str(p), str(d / "log"))
    return res


@pytest.mark.parametrize("case", ["o2_det_trev", "o2_det_trev_direct"])
def test_run_trev_o2_trusted(case, tmp_path):
    g = _gold()
    fcidump = os.path.join(GOLD, "o2", "FCIDUMP")
    r = _both(tmp_path, _input(g[case + "_input"], fcidump, "@PREFIX@"))
    gi, oi = r["gpu"].iterations(), r["oracle"].iterations()
    print(case, [(x["ndets"], x["energy"]) for x in gi])
    assert [x["ndets"] for x in gi] == [x["ndets"] for x in oi]
    assert max(abs(a["energy"] - b["energy"]) for a, b in zip(gi, oi)) < 1e-10
    tol = 1e-9 if case == "o2_det_trev" else 2e-9
    assert abs(r["gpu"].pt_energies()[0] - g["trusted_energy"][case]) < tol
    with open(tmp_path / "gpu" / "shci.e", "rb") as f:
        assert abs(struct.unpack("<d", f.read(8))[0] - g["trusted_energy"][case]) < tol
    assert len(r["gpu"].dets()) == len(r["oracle"].dets(1))


def test_run_trev_restart_chain_trusted(tmp_path):
    """tests/SHCI/restart_trev on the GPU: input2 writes the packed
    checkpoint, input3 restarts; shci.e = trusted_hci.e to 1e-9 Ha."""
    from dice_amd import shci
    g = _gold()
    fcidump = os.path.join(GOLD, "restart", "FCIDUMP")
    for k, name in enumerate(("restart_trev_input2", "restart_trev_input3")):
        p = tmp_path / ("in%d.dat" % k)
        p.write_text(_input(g[name], fcidump, tmp_path))
        shci.do_variational(str(p), str(tmp_path / ("log%d" % k)))
    with open(tmp_path / "shci.e", "rb") as f:
        e = struct.unpack("<d", f.read(8))[0]
    print("restart_trev", e, e - g["trusted_energy"]["restart_trev"])
    assert abs(e - g["trusted_energy"]["restart_trev"]) < 1e-9


def test_run_trev_roots_and_unpacked_state(tmp_path):
    """H2H2 with four roots and Treversal 1 (the spin-flip +1 roots): the
    GPU's roots equal the oracle's (1e-10 Ha), the unpacked determinant list
    is the oracle's, and the unpacked vectors satisfy F c = c (1e-12) and are
    eigenvectors of H (the DoRDM energy line equals the root's energy)."""
    g = _gold()
    fcidump = os.path.join(GOLD, "h2h2", "FCIDUMP")
    text = _input(g["h2he_rdm_input"], fcidump, "@PREFIX@",
                  drop=("nroots", "DoThreeRDM", "DoOneRDM", "DoSpinOneRDM"), extra="nroots 4\nTreversal 1\n")
    r = _both(tmp_path, text)
    assert np.max(np.abs(r["gpu"].energies() - r["oracle"].energies())) < 1e-10
    gd, od = r["gpu"].dets(), r["oracle"].dets(1)
    np.testing.assert_array_equal(gd, od)
    c = r["gpu"].ci()
    ints = [int(x[0]) & ((1 << 64) - 1) for x in gd]
    index = {v: k for k, v in enumerate(ints)}
    for k, x in enumerate(ints):
        f = ((x & 0x5555555555555555) << 1) | ((x & 0xAAAAAAAAAAAAAAAA) >> 1)
        p = (-1) ** bin(((x & 0x5555555555555555) << 1) & x).count("1") * (-1) ** bin(x & 0xAAAAAAAAAAAAAAAA).count("1")
        assert np.max(np.abs(c[index[f]] - p * c[k])) < 1e-12
    log = open(tmp_path / "gpu" / "log").read()
    es = [float(b.split()[0]) for b in log.split("E from 2RDM:")[1:]]
    assert len(es) == 4 and np.max(np.abs(np.array(es) - r["gpu"].energies())) < 1e-8


def test_run_trev_fp32_form_and_refusals(tmp_path):
    """The mixed-precision form with time reversal ends at the FP64 run's
    energy (1e-9 Ha: the final Davidson refines with hi + lo values) on the
    same determinants; the block Davidson and the learned screening are
    refused with Treversal."""
    from dice_amd import shci
    g = _gold()
    fcidump = os.path.join(GOLD, "o2", "FCIDUMP")
    text = _input(g["o2_det_trev_input"], fcidump, tmp_path, drop=("deterministic",), extra="nPTiter 0\n")
    p = tmp_path / "in.dat"
    p.write_text(text)
    r64 = shci.do_variational(str(p), str(tmp_path / "log64"))
    r32 = shci.do_variational(str(p), str(tmp_path / "log32"), precision=shci.PREC_FP32)
    assert abs(r32.energies()[0] - r64.energies()[0]) < 1e-9
    np.testing.assert_array_equal(r32.dets(), r64.dets())
    for kw in (dict(davidson_block=True), dict(learned=True)):
        with pytest.raises(shci.ShciError):
            shci.do_variational(str(p), str(tmp_path / "logx"), **kw)
