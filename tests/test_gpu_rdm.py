"""RDMs on the GPU (SHCIrdm::EvaluateOneRDM / EvaluateRDM,
SHCI/SHCIrdm.cpp:863-1100) against the oracle (oracle/rdm_ref.py) through the
C ABI, and whole runs with DoOneRDM / DoRDM / DoSpinOneRDM / dospinrdm /
TransitionRDM against the reference's own fixtures.

Bar: the GPU RDMs equal the oracle's to 1e-12 absolute (the same sums of
products of two coefficients, added by FP64 atomics in another order); the
files of whole runs match PySCF's FCI RDMs and the reference's trusted O2
2-RDM to the bars of the reference's test scripts (test_onepdm.py,
test_twopdm.py: L2 norm of the difference / number of elements), and the
oracle's RDMs of the run's own state to 1e-10 (the files keep 12 digits).
"""
import importlib.util
import json
import os

import numpy as np
import pytest

from test_gpu_parity import GOLD, systems  # noqa: F401  (module fixture)
from test_gpu_pt import _case

pytestmark = pytest.mark.gpu


def _rdm_ref():
    spec = importlib.util.spec_from_file_location(
        "rdm_ref", os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "rdm_ref.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _gold():
    return json.load(open(os.path.join(GOLD, "golden.json")))


def _input(text, fcidump, prefix, drop=(), extra=""):
    keep = [l for l in text.splitlines()
            if not any(l.strip().lower().startswith(k.lower()) for k in ("orbitals", "prefix") + tuple(drop))]
    return "\n".join(keep) + "\norbitals %s\nprefix %s\n%s" % (fcidump, prefix, extra)


def _run(tmp_path, text, name="in.dat"):
    from dice_amd import shci
    p = tmp_path / name
    p.write_text(text)
    return shci.do_variational(str(p), str(tmp_path / "log"))


@pytest.mark.parametrize("name", ["h8", "c0", "w2", "hub"])
def test_rdm_matches_oracle(name, systems):
    """Random determinants plus singles and doubles of them, unrelated bra and
    ket: spatial and spin-orbital 1- and 2-RDMs equal the oracle's."""
    from dice_amd import shci
    rdm = _rdm_ref()
    dets, bra = _case(name, systems, 150, 31)
    ket = np.random.default_rng(7).normal(0, 1, len(dets))
    ket /= np.linalg.norm(ket)
    norb = systems[name][1]
    g = shci.System.from_fcidump(systems[name][0])
    g.set_dets(dets)
    g1, g2, s1, s2 = g.rdm(bra, ket, spin=True)
    o1, oG = rdm.spin_rdms(dets, bra, norb, ket)
    r1, r2 = rdm.spatial_from_spin(o1, oG)
    print(name, len(dets), "max |g2| %.3e" % np.abs(r2).max())
    assert np.abs(r2).max() > 1e-3
    assert np.max(np.abs(g1 - r1)) < 1e-12 and np.max(np.abs(g2 - r2)) < 1e-12
    assert np.max(np.abs(s1 - o1)) < 1e-12
    assert np.max(np.abs(s2 - rdm.spin_pair_matrix(oG))) < 1e-12
    # the same held determinants, bra = ket: hermitian
    h1, h2 = g.rdm(bra)
    assert np.max(np.abs(h1 - h1.T)) < 1e-13 and np.max(np.abs(h2 - h2.transpose(2, 3, 0, 1))) < 1e-13


def test_run_h2he_rdm_files_match_pyscf(tmp_path):
    """tests/SHCI/H2He_rdm (without DoThreeRDM): spatial1RDM.0.0.txt and
    spatialRDM.0.0.txt match PySCF's FCI RDMs at the reference test's bar
    (1e-9), spin1RDM.0.0.txt sums over spin to the spatial 1-RDM
    (test_spin1RDM.py, 1e-9), and the log's "E from 2RDM" is E_var."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "h2h2", "FCIDUMP")
    r = _run(tmp_path, _input(g["h2he_rdm_input"], fcidump, tmp_path, drop=("DoThreeRDM",)))
    one = rdm.read_dice_1rdm(str(tmp_path / "spatial1RDM.0.0.txt"))
    two = rdm.read_dice_2rdm(str(tmp_path / "spatialRDM.0.0.txt"))
    ref1 = np.load(os.path.join(GOLD, "h2he_rdm", "pyscf_1RDM.npy"), allow_pickle=False)
    ref2 = np.load(os.path.join(GOLD, "h2he_rdm", "pyscf_2RDM.npy"), allow_pickle=False).transpose(0, 2, 1, 3)
    assert np.linalg.norm(one - ref1) / one.size < 1e-9
    assert np.linalg.norm(two - ref2) / two.size < 1e-9
    spin = rdm.read_dice_1rdm(str(tmp_path / "spin1RDM.0.0.txt"))
    assert spin.shape == (12, 12)
    assert np.linalg.norm(spin[0::2, 0::2] + spin[1::2, 1::2] - one) / one.size < 1e-9
    log = open(tmp_path / "log").read()
    assert "Calculating 1-RDM" in log and "Calculating 2-RDM" in log
    e = float(log.split("E from 2RDM:")[1].split()[0])
    assert abs(e - r.energies()[0]) < 1e-8


def test_run_h2co_rdm_files_match_pyscf(tmp_path):
    """tests/SHCI/h2co_rdm (FCI of 8 orbitals, 10 electrons; without
    DoThreeRDM): the 1-RDM to 1e-8 and the 2-RDM to 1e-9, the reference
    test's bars."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "h2co_rdm", "FCIDUMP")
    _run(tmp_path, _input(g["h2co_rdm_input"], fcidump, tmp_path, drop=("DoThreeRDM",)))
    one = rdm.read_dice_1rdm(str(tmp_path / "spatial1RDM.0.0.txt"))
    two = rdm.read_dice_2rdm(str(tmp_path / "spatialRDM.0.0.txt"))
    ref1 = np.load(os.path.join(GOLD, "h2co_rdm", "pyscf_1RDM.npy"), allow_pickle=False)
    ref2 = np.load(os.path.join(GOLD, "h2co_rdm", "pyscf_2RDM.npy"), allow_pickle=False).transpose(0, 2, 1, 3)
    print("h2co 1-RDM %.3e 2-RDM %.3e" % (np.linalg.norm(one - ref1) / one.size, np.linalg.norm(two - ref2) / two.size))
    assert np.linalg.norm(one - ref1) / one.size < 1e-8
    assert np.linalg.norm(two - ref2) / two.size < 1e-9


def test_run_o2_stoc_2rdm_trusted(tmp_path):
    """tests/SHCI/o2_stoc as the reference runs it (semistochastic PT after
    the variational step, which disables the PT RDM): spatialRDM.0.0.txt
    matches trusted2RDM.txt at the reference test's bar (1e-8)."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "o2", "FCIDUMP")
    r = _run(tmp_path, _input(g["o2_stoc_input"], fcidump, tmp_path))
    two = rdm.read_dice_2rdm(str(tmp_path / "spatialRDM.0.0.txt"))
    ref = rdm.read_dice_2rdm(os.path.join(GOLD, "o2_stoc", "trusted2RDM.txt"))
    assert np.linalg.norm(two - ref) / two.size < 1e-8
    log = open(tmp_path / "log").read()
    assert "(We cannot perform PT RDM with stochastic PT. Disabling RDM.)" in log
    assert len(r.pt_energies()) == 1


def test_run_ueg_spin1rdm_trace(tmp_path):
    """tests/SHCI/spin1rdm (uniform electron gas, 19 orbitals): the spin
    1-RDM summed over spin is the spatial 1-RDM (test_spin1RDM.py, 1e-8), and
    both match the oracle's 1-RDMs of the run's state (1e-10 spatial, the
    spin file's 8 digits)."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "ueg", "FCIDUMP")
    r = _run(tmp_path, _input(g["spin1rdm_input"], fcidump, tmp_path))
    one = rdm.read_dice_1rdm(str(tmp_path / "spatial1RDM.0.0.txt"))
    spin = rdm.read_dice_1rdm(str(tmp_path / "spin1RDM.0.0.txt"))
    assert np.linalg.norm(spin[0::2, 0::2] + spin[1::2, 1::2] - one) / one.size < 1e-8
    assert abs(np.trace(one) - 14) < 1e-9
    assert not os.path.exists(tmp_path / "spatialRDM.0.0.txt")
    dets, ci = r.dets(), r.ci()[:, 0]
    print("ueg", len(dets))
    o1 = rdm.spin_one_rdm(dets, ci, 19)
    assert np.max(np.abs(one - (o1[0::2, 0::2] + o1[1::2, 1::2]))) < 1e-10
    big, small = np.abs(o1) > 2e-6, np.abs(o1) < 5e-7  # the file keeps |v| > 1e-6, 8 digits
    assert np.max(np.abs(spin - o1)[big]) < 1e-7 and np.all(spin[small] == 0.0)


def test_run_transition_rdms_match_oracle(tmp_path):
    """TransitionRDM with three roots of H2H2: spatial1RDM.i.j.txt and
    spatialRDM.i.j.txt (i > j, bra = root i, ket = root j) hold the oracle's
    transition RDMs of the run's roots (1e-10); the transition 1-RDMs have
    zero trace (orthogonal roots) and the diagonal files' energies are the
    roots' energies."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "h2h2", "FCIDUMP")
    text = _input(g["h2he_rdm_input"], fcidump, tmp_path, drop=("DoThreeRDM", "nroots", "DoSpinOneRDM"),
                  extra="nroots 3\nTransitionRDM\n")
    r = _run(tmp_path, text)
    dets, ci = r.dets(), r.ci()
    for i in range(3):
        for j in range(i + 1):
            o1, o2 = rdm.rdms(dets, ci[:, i], 6, ci[:, j])
            one = rdm.read_dice_1rdm(str(tmp_path / ("spatial1RDM.%d.%d.txt" % (i, j))))
            two = rdm.read_dice_2rdm(str(tmp_path / ("spatialRDM.%d.%d.txt" % (i, j))))
            assert np.max(np.abs(one - o1)) < 1e-10 and np.max(np.abs(two - o2)) < 1e-10
            if i != j:
                assert abs(np.trace(one)) < 1e-8
    log = open(tmp_path / "log").read()
    es = [float(b.split()[0]) for b in log.split("E from 2RDM:")[1:]]
    assert len(es) == 6
    diag = [es[0], es[2], es[5]]  # root i: transitions to j < i first, then (i, i)
    assert np.max(np.abs(np.array(diag) - r.energies())) < 1e-8


def test_run_direct_rdm_and_spin_rdm(tmp_path):
    """directdavidson with dospinrdm: the 2-RDM files only (the reference's
    direct-mode branch), equal to the stored-Hamiltonian run's; spinRDM.0.0.txt
    holds the oracle's pair-indexed spin 2-RDM (1e-10, above the file's 1e-6
    cut)."""
    rdm = _rdm_ref()
    g = _gold()
    fcidump = os.path.join(GOLD, "h2h2", "FCIDUMP")
    drop = ("DoThreeRDM", "DoOneRDM", "DoSpinOneRDM", "DoRDM")
    a, b = tmp_path / "stored", tmp_path / "direct"
    a.mkdir()
    b.mkdir()
    _run(a, _input(g["h2he_rdm_input"], fcidump, a, drop=drop, extra="dospinrdm\n"))
    rb = _run(b, _input(g["h2he_rdm_input"], fcidump, b, drop=drop, extra="dospinrdm\ndirectdavidson\n"))
    ta = rdm.read_dice_2rdm(str(a / "spatialRDM.0.0.txt"))
    tb = rdm.read_dice_2rdm(str(b / "spatialRDM.0.0.txt"))
    assert np.max(np.abs(ta - tb)) < 1e-9
    assert os.path.exists(a / "spatial1RDM.0.0.txt") and not os.path.exists(b / "spatial1RDM.0.0.txt")
    _, G = rdm.spin_rdms(rb.dets(), rb.ci()[:, 0], 6)
    found = {}
    for line in open(b / "spinRDM.0.0.txt"):
        p, q, r_, s, v = line.split()
        found[(int(p), int(q), int(r_), int(s))] = float(v)
    for p in range(1, 12):
        for q in range(p):
            for r_ in range(1, 12):
                for s in range(r_):
                    v = G[p, q, r_, s]
                    if abs(v) > 1e-6:
                        assert abs(found.pop((p, q, r_, s)) - v) < 1e-10
    assert not found


def test_run_rdm_refusals(tmp_path):
    """DoThreeRDM and DoRDM with deterministic PT (the relaxed PT 2-RDM) are
    refused; relaxedRDM keeps the variational RDMs and runs the PT."""
    from dice_amd import shci
    g = _gold()
    fcidump = os.path.join(GOLD, "h2h2", "FCIDUMP")
    with pytest.raises(shci.ShciError):
        _run(tmp_path, _input(g["h2he_rdm_input"], fcidump, tmp_path), name="three.dat")
    base = _input(g["h2he_rdm_input"], fcidump, tmp_path, drop=("DoThreeRDM", "nPTiter"))
    with pytest.raises(shci.ShciError):
        _run(tmp_path, base + "deterministic\n", name="det.dat")
    r = _run(tmp_path, base + "deterministic\nrelaxedRDM\n", name="relaxed.dat")
    assert len(r.pt_energies()) == 1 and os.path.exists(tmp_path / "spatialRDM.0.0.txt")
