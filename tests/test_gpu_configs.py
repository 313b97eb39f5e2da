"""The benchmark configurations' systems through the whole variational
iteration, HIP path against the CPU oracle, and the drop-in executable run
as a process the way the reference's PySCF driver runs Dice.

* The 24-orbital system (12 alpha + 12 beta electrons, 96 Cholesky vectors)
  at eps1 5e-3: the first iteration from HF (27,685 determinants).
* The 40-orbital system (20 + 20 electrons, 160 Cholesky vectors, two-word
  determinants, two roots) at eps1 0.1: the first iteration from HF (57,963
  determinants; at 5e-3 it selects 217,459 and the oracle needs three
  minutes on eight threads).
* The 24-orbital system's full benchmark space (six iterations at eps1
  5e-3, 1,177,638 determinants): one sigma against the oracle's.

Bar (FP64): the determinant set bit for bit, every root's energy to 1e-10 Ha
of the oracle.  The FP32-value form: the same determinant set and every
energy within 1e-4 Ha of the FP64 run.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle_bind import OracleRun

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(autouse=True)
def _oracle_threads(monkeypatch):
    # the oracle's rank-parallel loops on the host's cores (same results as
    # one thread: test_threaded_oracle_matches_serial)
    monkeypatch.setenv("ORACLE_THREADS", str(max(1, min(16, os.cpu_count() or 1))))


def _first_iteration(tmp_path, name, eps1, nroots):
    from dice_amd import synth
    c, n, na, nb, nchol = synth.MOLECULAR[name]
    fcidump = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, fcidump, synth.hf_occupation(na, nb), [(0, eps1)], 1, nroots=nroots,
                      extra=["prefix %s" % tmp_path])
    return inp, 2 * n // 64 + 1


def _check_against_oracle(tmp_path, inp, words, nroots, ndets):
    from dice_amd import shci
    print("oracle run: %s" % inp, flush=True)
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    print("GPU run, FP64", flush=True)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    oi, gi = o.iterations(), r.iterations()
    assert [x["ndets"] for x in gi] == [x["ndets"] for x in oi]
    assert gi[-1]["ndets"] == ndets
    np.testing.assert_array_equal(r.dets(), o.dets(words))
    e64 = np.asarray(r.energies())
    assert len(e64) == nroots
    np.testing.assert_allclose(e64, o.energies(), rtol=0, atol=1e-10)
    print("GPU run, FP32 values", flush=True)
    r32 = shci.do_variational(inp, str(tmp_path / "gpu32.log"), precision=shci.PREC_FP32)
    np.testing.assert_array_equal(r32.dets(), r.dets())
    d = np.abs(np.asarray(r32.energies()) - e64)
    print("FP32 form: |E - E_fp64| per root = %s" % d, flush=True)
    assert np.all(d < 1e-4)


def test_c1_first_iteration(tmp_path):
    inp, words = _first_iteration(tmp_path, "c1", 5e-3, 1)
    _check_against_oracle(tmp_path, inp, words, 1, 27685)


def test_c3_first_iteration_two_roots(tmp_path):
    inp, words = _first_iteration(tmp_path, "c3", 0.1, 2)
    assert words == 2
    _check_against_oracle(tmp_path, inp, words, 2, 57963)


def _table_rows(text):
    """(iteration, root, #dets, energy, #Davidson) rows of the variational table
    (SHCI/SHCIbasics.cpp:1181-1190)."""
    rows = []
    for line in text.splitlines():
        p = line.split()
        if len(p) == 7 and p[0].isdigit() and p[1].isdigit():
            try:
                rows.append((int(p[0]), int(p[1]), int(p[3]), float(p[4]), int(p[5])))
            except ValueError:
                pass
    return rows


def test_drop_in_executable_as_a_process(tmp_path):
    """`shci_var input.dat > output.dat` in the working directory, as
    scripts/shci.py:1257-1278 runs Dice, then `shci.e` read as
    scripts/shci.py:1291-1302 reads it (nroots raw doubles): equal to the
    oracle executable's run of the same input to 1e-10 Ha, with the same
    table rows (determinant and Davidson counts) on stdout.  The h8 input
    asks for the semistochastic PT step (sampleN without nPTiter 0), so
    shci.e holds E_var + E_PT; a fixed seed makes both runs draw the same
    samples (the reference's default seed is the wall time,
    SHCI/input.cpp:80)."""
    import json
    exe = os.path.join(ROOT, "dice_amd", "bin", "shci_var")
    ref = os.path.join(ROOT, "oracle", "build", "shci_ref")
    assert os.path.exists(exe), "build the library first (python -m dice_amd.build)"
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)
    outs = {}
    for tag, prog in (("gpu", exe), ("ref", ref)):
        d = tmp_path / tag
        d.mkdir()
        (d / "input.dat").write_text(g["h8_631g"]["input"] + "\norbitals %s\nprefix %s\nseed 7\n"
                                     % (os.path.join(GOLD, "h8_631g", "FCIDUMP"), d))
        with open(d / "output.dat", "w") as out:
            p = subprocess.run([prog, "input.dat"], cwd=str(d), stdout=out, stderr=subprocess.STDOUT, timeout=300)
        text = (d / "output.dat").read_text()
        assert p.returncode == 0, text[-2000:]
        raw = (d / "shci.e").read_bytes()
        assert len(raw) == 8, len(raw)
        outs[tag] = (np.array(struct.unpack("d", raw)), _table_rows(text))
    (eg, rows_g), (er, rows_r) = outs["gpu"], outs["ref"]
    np.testing.assert_allclose(eg, er, rtol=0, atol=1e-10)
    assert len(rows_g) == len(rows_r) > 0
    assert abs(rows_g[-1][3] - g["h8_631g"]["iterations"][-1]["energy"]) < 6e-11
    for a, b in zip(rows_g, rows_r):
        assert a[:3] == b[:3] and a[4] == b[4], (a, b)
        assert abs(a[3] - b[3]) < 1e-9


def test_c1_final_space_sigma_matches_oracle(tmp_path):
    """At the benchmark's full size: the 24-orbital system through its six
    iterations at eps1 5e-3 (1,177,638 determinants, 0.91e9 lower-triangle
    elements), then the oracle's Hamiltonian build and sigma (the
    reference's MakeHfromSMHelpers2 and Hmult2, rank-parallel loops on the
    host threads) on the same final space: the GPU sigma (the default: the
    tiled copy of the stored lower triangle, sums in fixed point) equals it
    to 1e-12 relative, and the number of stored elements agrees."""
    from dice_amd import synth, shci
    from oracle_bind import OracleSystem
    c, n, na, nb, nchol = synth.MOLECULAR["c1"]
    fcidump = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, fcidump, synth.hf_occupation(na, nb), [(0, 5e-3)], 6, extra=["prefix %s" % tmp_path])
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    dets = r.dets()
    assert len(dets) == 1177638
    G = shci.System.from_fcidump(fcidump)
    G.set_dets(dets)
    nnz_lower = G.make_hamiltonian()
    x = np.random.default_rng(4).normal(size=len(dets))
    y = G.hmult(x)
    G.close()
    print("oracle build on %d determinants" % len(dets), flush=True)
    o = OracleSystem(fcidump)
    onnz = o.csr_append(dets)
    yo = o.sigma(x)
    rel = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    print("final space: max |y_gpu - y_oracle| / max |y_oracle| = %.3e" % rel, flush=True)
    assert rel <= 1e-12
    assert onnz == nnz_lower

