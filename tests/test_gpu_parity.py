"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: bit-exact for determinants, CSR structure and Hamiltonian values;
1e-12 relative for sigma (summation order differs); 1e-10 Ha for energies.
"""
import os

import numpy as np
import pytest

import helpers
from oracle_bind import OracleRun, OracleSystem, sort_dets

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def systems(tmp_path_factory):
    from dice_amd import synth
    d = tmp_path_factory.mktemp("fcidumps")
    out = {"h8": (os.path.join(GOLD, "h8_631g", "FCIDUMP"), 16, 4, 4)}
    p = str(d / "c0")
    synth.write_molecular_fcidump(p, 0, 16, 8, 8, 64)
    out["c0"] = (p, 16, 8, 8)
    p = str(d / "w2")  # 34 orbitals: two 64-bit words
    synth.write_molecular_fcidump(p, 5, 34, 6, 6, 12)
    out["w2"] = (p, 34, 6, 6)
    p = str(d / "hub")
    synth.write_hubbard_fcidump(p, 6, 6, 1.0, 4.0, 18, 18)
    out["hub"] = (p, 36, 18, 18)
    return out


def _gpu(path):
    from dice_amd import shci
    return shci.System.from_fcidump(path)


def _sample(name, systems, n, seed):
    path, norb, na, nb = systems[name]
    rng = np.random.default_rng(seed)
    W = helpers.words_for(norb)
    base = helpers.random_dets(norb, na, nb, n, rng, W)
    # add near neighbours so singles / doubles are well represented
    near = []
    for d in base[: n // 2]:
        occ = helpers.det_to_occ(d)
        near.append(helpers.occ_to_det(helpers.excite(occ, norb, rng, 1 + (len(near) % 2)), W))
    return base, np.array(near)


@pytest.mark.parametrize("name", ["h8", "c0", "w2", "hub"])
def test_energy_and_hij_bitwise(name, systems):
    path = systems[name][0]
    g, o = _gpu(path), OracleSystem(path)
    base, near = _sample(name, systems, 400, 11)
    np.testing.assert_array_equal(g.energy(base), o.energy(base))
    np.testing.assert_array_equal(g.hij(base[: len(near)], near), o.hij(base[: len(near)], near))
    np.testing.assert_array_equal(g.hij(near, base[: len(near)]), o.hij(near, base[: len(near)]))


@pytest.mark.parametrize("name,eps", [("h8", 1e-4), ("c0", 1e-3), ("w2", 1e-3), ("hub", 2e-4)])
def test_screening_det_set(name, eps, systems):
    path = systems[name][0]
    g, o = _gpu(path), OracleSystem(path)
    base, near = _sample(name, systems, 300, 5)
    dets = np.concatenate([base, near])
    dets = dets[np.unique(dets, axis=0, return_index=True)[1]]
    rng = np.random.default_rng(3)
    cmax = np.abs(rng.normal(0, 0.2, len(dets))) + 1e-3
    g.set_dets(dets)
    got = g.get_determinants_variational(cmax, eps)
    ref = o.screen(dets, cmax, eps)
    assert len(got) == len(ref) and len(ref) > 0
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, sort_dets(got))


def _space(name, systems, n, seed):
    """A connected determinant space: repeated screening from the reference."""
    path, norb, na, nb = systems[name]
    o = OracleSystem(path)
    W = o.words
    if name == "hub":
        from dice_amd import synth
        ref = synth.neel_occupation(6, 6)
    else:
        from dice_amd import synth
        ref = synth.hf_occupation(na, nb)
    dets = helpers.occ_to_det(ref, W)[None, :]
    eps = 1e-2
    while len(dets) < n:
        new = o.screen(dets, np.ones(len(dets)), eps)
        if len(new):
            dets = np.concatenate([dets, new])
        eps *= 0.5
    return dets[:n]


@pytest.mark.parametrize("name,n", [("h8", 1500), ("c0", 3000), ("w2", 2500), ("hub", 2500)])
def test_hamiltonian_rows_and_sigma(name, n, systems):
    path = systems[name][0]
    dets = _space(name, systems, n, 1)
    g, o = _gpu(path), OracleSystem(path)
    g.set_dets(dets)
    rp, cols, vals = g.hamiltonian_rows()
    o.csr_append(dets)
    orp, ocols, ovals = o.csr()
    np.testing.assert_array_equal(np.diff(rp), np.diff(orp))
    for i in range(len(dets)):
        a, b = rp[i], rp[i + 1]
        # diagonal first, then J < I; compare after sorting the oracle's row
        assert cols[a] == ocols[orp[i]] == i
        assert vals[a] == ovals[orp[i]]
        order = np.argsort(ocols[orp[i] + 1:orp[i + 1]], kind="stable")
        np.testing.assert_array_equal(cols[a + 1:b], ocols[orp[i] + 1:orp[i + 1]][order])
        np.testing.assert_array_equal(vals[a + 1:b], ovals[orp[i] + 1:orp[i + 1]][order])
    x = np.random.default_rng(2).normal(size=len(dets))
    y = g.hmult(x)
    yo = o.sigma(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))


def test_hamiltonian_rows_large_alpha_groups(tmp_path):
    """Alpha groups far larger than a row's beta singles (16 orbitals, one
    alpha and eight beta electrons: three alpha strings x all 12,870 beta
    strings), where the build looks the row's beta singles up in the
    beta-ordered members of each alpha group one excitation away instead of
    scanning them: rows and values equal the oracle's, and sigma agrees."""
    from dice_amd import synth
    import itertools
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 7, 16, 1, 8, 24)
    dets = []
    for a in (0, 1, 2):
        for b in itertools.combinations(range(16), 8):
            dets.append((1 << (2 * a)) | sum(1 << (2 * x + 1) for x in b))
    dets = np.array(dets, dtype=np.int64).reshape(-1, 1)
    g, o = _gpu(path), OracleSystem(path)
    g.set_dets(dets)
    rp, cols, vals = g.hamiltonian_rows()
    o.csr_append(dets)
    orp, ocols, ovals = o.csr()
    np.testing.assert_array_equal(np.diff(rp), np.diff(orp))
    for i in range(len(dets)):
        a, b = rp[i], rp[i + 1]
        order = np.argsort(ocols[orp[i] + 1:orp[i + 1]], kind="stable")
        np.testing.assert_array_equal(cols[a + 1:b], ocols[orp[i] + 1:orp[i + 1]][order])
        np.testing.assert_array_equal(vals[a + 1:b], ovals[orp[i] + 1:orp[i + 1]][order])
    x = np.random.default_rng(5).normal(size=len(dets))
    y, yo = g.hmult(x), o.sigma(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))


# (system, precision, screened sources, eps1 of the two rounds): a few
# thousand to a few tens of thousands of new determinants per round
@pytest.mark.parametrize("name,prec,nsrc,eps", [("h8", 0, 60, (3e-3, 1.5e-3)), ("w2", 0, 5, (1e-1, 5e-2)),
                                                ("hub", 0, 60, (3e-3, 1.5e-3)), ("c0", 1, 20, (1e-2, 5e-3))])
def test_incremental_hamiltonian(name, prec, nsrc, eps, systems):
    """Two rounds of screen -> append -> build extend the Hamiltonian in place
    (new rows, and the new columns of old rows); the rows equal those of a
    build from scratch over the same determinants, element for element."""
    path = systems[name][0]
    dets = _space(name, systems, 800, 1)
    g = _gpu(path)
    g.set_precision(prec)
    g.set_dets(dets)
    g.make_hamiltonian()
    rng = np.random.default_rng(8)
    for e1 in eps:
        cmax = np.zeros(g.num_dets())
        cmax[:nsrc] = np.abs(rng.normal(0, 0.3, nsrc)) + 1e-3
        new = g.get_determinants_variational(cmax, e1)
        assert len(new) > 0
        g.append_new()
        rp, cols, vals = g.hamiltonian_rows()
        f = _gpu(path)
        f.set_precision(prec)
        f.set_dets(g.dets())
        frp, fcols, fvals = f.hamiltonian_rows()
        np.testing.assert_array_equal(rp, frp)
        np.testing.assert_array_equal(cols, fcols)
        np.testing.assert_array_equal(vals, fvals)
        x = rng.normal(size=g.num_dets())
        y, yf = g.hmult(x), f.hmult(x)
        assert np.max(np.abs(y - yf)) <= 1e-12 * np.max(np.abs(yf))


@pytest.mark.parametrize("name,prec", [("c0", 0), ("hub", 0), ("w2", 0), ("c0", 1)])
def test_sigma_layouts_agree(name, prec, systems, monkeypatch):
    """Every sigma layout gives sigma over H in determinant order ("external")
    to 1e-12 relative with FP64 values (1e-6 with FP32 values, where the
    relabelled paths also round x to FP32), for one vector and for two: the
    tiled copy ("tiled": blocks of 256 relabelled rows with entries grouped
    by column tile, one launch per vector), the row-ordered relabelled copy
    ("rows": one vector, or the SpMM of two), both with same-beta columns
    gathered from a beta-major copy of x, every column alpha-major
    ("alpha"), and the two-triangle default (the tiled copy, large segments
    gathering from an LDS copy of their x tile; "tiled-global": every gather
    from global memory); and the tiled copy of the lower triangle alone
    ("half", the transposed half in fixed point; "half-split" with split
    columns).  Each layout repeats its
    own result bit for bit, staging every segment gives the default's bits,
    and the tiled copy built from H directly equals the one built from the
    row copy bit for bit."""
    path = systems[name][0]
    dets = _space(name, systems, 3000, 1)
    x = np.random.default_rng(5).normal(size=(len(dets), 2))
    out = {}
    for layout in ("default", "tiled", "tiled-from-H", "tiled-global", "tiled-persistent", "rows", "external", "alpha",
                   "half", "half-split"):
        for var in ("SHCI_SIGMA_LAYOUT", "SHCI_SIGMA_TILE_BUILD", "SHCI_SIGMA_TILE_STAGE", "SHCI_SIGMA_TILE_STAGE_MIN",
                    "SHCI_SIGMA_TILE_PERSISTENT", "SHCI_HALF_SPLIT"):
            monkeypatch.delenv(var, raising=False)
        monkeypatch.setenv("SHCI_STORAGE", "half" if layout.startswith("half") else "full")
        if layout == "half-split":
            monkeypatch.setenv("SHCI_HALF_SPLIT", "1")
        if layout == "tiled-from-H":
            monkeypatch.setenv("SHCI_SIGMA_LAYOUT", "tiled")
            monkeypatch.setenv("SHCI_SIGMA_TILE_BUILD", "direct")
        elif layout == "tiled-global":
            monkeypatch.setenv("SHCI_SIGMA_TILE_STAGE", "0")
        elif layout == "tiled-persistent":
            monkeypatch.setenv("SHCI_SIGMA_TILE_PERSISTENT", "1")
        elif layout == "tiled":
            # every segment staged, down to one instruction
            monkeypatch.setenv("SHCI_SIGMA_LAYOUT", "tiled")
            monkeypatch.setenv("SHCI_SIGMA_TILE_STAGE_MIN", "1")
        elif layout not in ("default", "half", "half-split"):
            monkeypatch.setenv("SHCI_SIGMA_LAYOUT", layout)
        g = _gpu(path)
        g.set_precision(prec)
        g.set_dets(dets)
        g.make_hamiltonian()
        out[layout] = (g.hmult(x), g.hmult(x[:, 0].copy()))
        # deterministic: the same call again gives the same bits
        np.testing.assert_array_equal(out[layout][1], g.hmult(x[:, 0].copy()))
    tol = 1e-12 if prec == 0 else 1e-6
    ref = out["external"]
    for layout in ("default", "tiled", "tiled-from-H", "tiled-global", "tiled-persistent", "rows", "alpha", "half",
                   "half-split"):
        for k in (0, 1):
            assert np.max(np.abs(out[layout][k] - ref[k])) <= tol * np.max(np.abs(ref[k])), (layout, k)
    # the staged kernel's sums do not depend on which segments it stages
    # (the default stages large ones, "tiled" all of them); the tiled copy
    # built from H through the permutation (when the row copy does not fit
    # beside it) holds the same entries in the same order
    for k in (0, 1):
        np.testing.assert_array_equal(out["default"][k], out["tiled"][k])
        np.testing.assert_array_equal(out["tiled-from-H"][k], out["default"][k])
        # persistent workgroups take the blocks in another order, each block
        # summed the same way
        np.testing.assert_array_equal(out["tiled-persistent"][k], out["default"][k])


@pytest.mark.parametrize("storage", ["full", "half"])
@pytest.mark.parametrize("name", ["c0", "w2"])
def test_tiled_sigma_bitwise_across_handles_and_growth(name, storage, systems, monkeypatch):
    """The tiled sigma's sums do not depend on scheduling: a handle whose
    tiled copy was first built for a smaller space and then rebuilt for the
    full one gives the bits of a fresh handle on the full space (both
    triangles stored, or the lower one with the transposed half in fixed
    point)."""
    monkeypatch.setenv("SHCI_SIGMA_LAYOUT", "tiled")
    monkeypatch.setenv("SHCI_STORAGE", storage)
    path = systems[name][0]
    dets = _space(name, systems, 3000, 1)
    x = np.random.default_rng(9).normal(size=len(dets))
    g = _gpu(path)
    g.set_dets(dets[:1000])
    g.make_hamiltonian()
    g.hmult(x[:1000])
    g.set_dets(dets)
    g.make_hamiltonian()
    y1 = g.hmult(x)
    f = _gpu(path)
    f.set_dets(dets)
    f.make_hamiltonian()
    np.testing.assert_array_equal(y1, f.hmult(x))


@pytest.mark.parametrize("name,nroots", [("c0", 1), ("hub", 2)])
def test_davidson_energies(name, nroots, systems):
    path = systems[name][0]
    dets = _space(name, systems, 2000, 1)
    g, o = _gpu(path), OracleSystem(path)
    g.set_dets(dets)
    g.make_hamiltonian()
    o.csr_append(dets)
    x0 = np.zeros((len(dets), nroots))
    x0[0, 0] = 1.0
    if nroots > 1:
        x0[1:, 1] = np.random.default_rng(4).normal(size=len(dets) - 1)
    e, x, it = g.davidson(x0.copy(), tol=1e-8)
    diag = o.energy(dets)
    eo, xo, ito = o.davidson(x0.copy(), diag, nroots + 4, 1e-8)
    np.testing.assert_allclose(e, eo, rtol=0, atol=1e-10)
    assert abs(it - ito) <= 2


def _write_input(tmp_path, text):
    p = tmp_path / "input.dat"
    p.write_text(text)
    return str(p)


def _compare_runs(inp, tmp_path, words, nroots=1):
    from dice_amd import shci
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    oi, gi = o.iterations(), r.iterations()
    assert [x["ndets"] for x in oi] == [x["ndets"] for x in gi]
    for a, b in zip(oi, gi):
        assert abs(a["energy"] - b["energy"]) < 1e-10
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-10)
    np.testing.assert_array_equal(r.dets(), o.dets(words))
    return o, r


def test_run_h8_golden(tmp_path):
    import json
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)
    inp = _write_input(tmp_path, g["h8_631g"]["input"] + "\norbitals %s\nprefix %s\n" % (
        os.path.join(GOLD, "h8_631g", "FCIDUMP"), tmp_path))
    o, r = _compare_runs(inp, tmp_path, 1)
    its = r.iterations()
    for mine, ref in zip(its, g["h8_631g"]["iterations"]):
        assert mine["ndets"] == ref["ndets"]
        assert abs(mine["energy"] - ref["energy"]) < 6e-11
        assert mine["davidson"] == ref["davidson"]
    assert abs(r.energies()[0] - g["h8_631g"]["final_energy"]) < 6e-11


def test_run_n2_ras(tmp_path):
    import json
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)
    text = g["n2_RAS_input"].replace("orbitals ./FCIDUMP", "orbitals %s" % os.path.join(GOLD, "n2_RAS", "FCIDUMP"))
    inp = _write_input(tmp_path, text + "\nprefix %s\n" % tmp_path)
    o, r = _compare_runs(inp, tmp_path, 1)
    assert abs(r.energies()[0] - g["trusted_energy"]["n2_RAS"]) < 1e-9


def test_run_restart_trusted(tmp_path):
    """524k determinants; the reference's trusted energy to 1e-9 (no oracle run)."""
    import json
    from dice_amd import shci
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)
    text = "\n".join(l for l in g["restart_input3"].splitlines() if l.strip() != "restart")
    text += "\norbitals %s\nprefix %s\n" % (os.path.join(GOLD, "restart", "FCIDUMP"), tmp_path)
    inp = _write_input(tmp_path, text)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    assert [x["ndets"] for x in r.iterations()][-1] == 524307
    assert abs(r.energies()[0] - g["trusted_energy"]["restart"]) < 1e-9


def test_run_c0_first_iteration(tmp_path, systems):
    """The 16-orbital workload's first iteration (5,695 determinants; the next
    one selects 650k and takes the CPU oracle about ten minutes)."""
    from dice_amd import synth
    path = systems["c0"][0]
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(8, 8), [(0, 1e-3)], 1)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    o, r = _compare_runs(inp, tmp_path, 1)
    assert r.iterations()[-1]["ndets"] == 5695


def test_run_small_molecular_per_iteration_det_sets(tmp_path):
    """A 12-orbital system of the same construction through a full schedule:
    the determinant set after every iteration is identical to the oracle's."""
    from dice_amd import synth, shci
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 9, 12, 4, 4, 32)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(4, 4), [(0, 3e-3), (2, 1e-3)], 5)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    oi, gi = o.iterations(), r.iterations()
    assert [x["ndets"] for x in oi] == [x["ndets"] for x in gi]
    for a, b in zip(oi, gi):
        assert abs(a["energy"] - b["energy"]) < 1e-10
    np.testing.assert_array_equal(r.dets(), o.dets(1))
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-10)


def test_run_hubbard_two_roots(tmp_path):
    from dice_amd import synth
    p = str(tmp_path / "FCIDUMP")
    synth.write_hubbard_fcidump(p, 4, 4, 1.0, 4.0, 8, 8)
    inp = str(tmp_path / "input.dat")
    # three iterations (29,937 determinants): the oracle needs ~1.5 s; a
    # fourth would take it minutes
    synth.write_input(inp, p, synth.neel_occupation(4, 4), [(0, 1e-3)], 3, nroots=2)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    from dice_amd import shci
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-9)


def test_run_mixed_precision_energy_gap(tmp_path):
    """FP32 Hamiltonian values with FP64 accumulation and the FP64-accurate
    refinement: the variational energy stays within 0.1 mHa of the FP64 path
    (the bar of the mixed-precision form); the FP64 path is the oracle-pinned
    one."""
    from dice_amd import synth, shci
    path = str(tmp_path / "FCIDUMP")
    synth.write_molecular_fcidump(path, 9, 12, 4, 4, 32)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(4, 4), [(0, 3e-3), (2, 1e-3)], 5)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    e64 = shci.do_variational(inp, str(tmp_path / "a.log")).energies()[0]
    r32 = shci.do_variational(inp, str(tmp_path / "b.log"), precision=shci.PREC_FP32)
    e32 = r32.energies()[0]
    print("mixed-precision energy gap: %.3e Ha" % (e32 - e64))
    assert abs(e32 - e64) < 1e-4
    # the refinement makes the final energy FP64-accurate for its space
    g = shci.System.from_fcidump(path)
    g.set_dets(r32.dets())
    g.make_hamiltonian()
    assert abs(e32 - g.core_energy - g.davidson(np.eye(len(r32.dets()), 1), tol=1e-8)[0][0]) < 1e-7


# ---- batching form: SpMM, block Davidson, hipGraph replay, autotuned launch shapes ----

@pytest.mark.parametrize("name,prec", [("c0", 0), ("hub", 0), ("w2", 0), ("c0", 1)])
def test_spmm_matches_column_sigma(name, prec, systems, monkeypatch):
    """sigma of k columns at once (interleaved SpMM over both stored
    triangles, k = 2..8 and 11 = 8 + 3) equals k single-vector sigmas to
    1e-12 relative."""
    monkeypatch.setenv("SHCI_STORAGE", "full")
    path = systems[name][0]
    dets = _space(name, systems, 3000, 1)
    g = _gpu(path)
    g.set_precision(prec)
    g.set_dets(dets)
    g.make_hamiltonian()
    x = np.random.default_rng(6).normal(size=(len(dets), 11))
    one = np.stack([g.hmult(x[:, j]) for j in range(11)], axis=1)
    for k in (2, 3, 4, 5, 8, 11):
        y = g.hmult(x[:, :k])
        assert np.max(np.abs(y - one[:, :k])) <= 1e-12 * np.max(np.abs(one)), k


@pytest.mark.parametrize("name,nroots", [("c0", 1), ("hub", 2), ("hub", 4)])
def test_block_davidson_energies(name, nroots, systems):
    """The block Davidson converges to the reference Davidson's eigenvalues
    (oracle, tolerance 1e-8 on the residual norm)."""
    path = systems[name][0]
    dets = _space(name, systems, 2000, 1)
    g, o = _gpu(path), OracleSystem(path)
    g.set_options(davidson_block=True)
    g.set_dets(dets)
    g.make_hamiltonian()
    o.csr_append(dets)
    x0 = np.zeros((len(dets), nroots))
    x0[0, 0] = 1.0
    rng = np.random.default_rng(4)
    for r in range(1, nroots):
        x0[1:, r] = rng.normal(size=len(dets) - 1)
    e, x, it = g.davidson(x0.copy(), tol=1e-8)
    eo, xo, ito = o.davidson(x0.copy(), o.energy(dets), nroots + 4, 1e-8)
    np.testing.assert_allclose(e, eo, rtol=0, atol=1e-9)
    # Ritz vectors: unit norm, residual small
    hx = g.hmult(x)
    for r in range(nroots):
        assert abs(np.linalg.norm(x[:, r]) - 1.0) < 1e-10
        assert np.linalg.norm(hx[:, r] - e[r] * x[:, r]) < 1e-7


def test_block_davidson_graph_replays_match_eager(systems):
    """Iterations replayed from captured hipGraphs give the same eigenpairs,
    bit for bit, as the same iterations launched eagerly, over three calls
    (graphs captured in the first call are replayed in the later ones, with
    eager work and synchronisations in between)."""
    path = systems["hub"][0]
    dets = _space("hub", systems, 2500, 1)
    res = {}
    for graphs in (False, True):
        g = _gpu(path)
        g.set_options(davidson_block=True, hip_graphs=graphs)
        g.set_dets(dets)
        g.make_hamiltonian()
        out = []
        for call in range(3):
            x0 = np.zeros((len(dets), 3))
            x0[call, 0] = 1.0
            x0[:, 1] = np.random.default_rng(call).normal(size=len(dets))
            x0[:, 2] = np.random.default_rng(10 + call).normal(size=len(dets))
            e, x, it = g.davidson(x0, tol=1e-7)
            _ = g.hmult(x[:, 0])  # eager work between graph replays
            out.append((e, x, it))
        res[graphs] = out
    for (e0, x0, i0), (e1, x1, i1) in zip(res[False], res[True]):
        assert i0 == i1
        np.testing.assert_array_equal(e0, e1)
        np.testing.assert_array_equal(x0, x1)


def test_autotuned_sigma_matches_default(systems, monkeypatch):
    """With autotune on (both triangles stored: the row copy's launch shapes),
    the Hamiltonian build times every launch shape and keeps the fastest;
    sigma results do not depend on the shape (1e-12)."""
    monkeypatch.setenv("SHCI_STORAGE", "full")
    path = systems["c0"][0]
    dets = _space("c0", systems, 24000, 1)
    x = np.random.default_rng(7).normal(size=(len(dets), 3))
    ys = []
    for tune in (False, True):
        g = _gpu(path)
        g.set_options(autotune=tune, davidson_block=True)
        g.set_dets(dets)
        g.make_hamiltonian()
        ys.append((g.hmult(x[:, 0]), g.hmult(x)))
    for a, b in zip(ys[0], ys[1]):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


def _dense_lowest(path, dets, k):
    """The k lowest eigenvalues of the oracle's Hamiltonian over dets (dense)."""
    import scipy.linalg as sl
    o = OracleSystem(path)
    o.csr_append(dets)
    rp, cols, vals = o.csr()
    n = len(dets)
    H = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    H[rows, cols] = vals
    H[cols, rows] = vals
    return sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])


@pytest.mark.parametrize("nroots", [2, 4])
def test_run_batching_form_vs_dense(nroots, tmp_path):
    """The whole variational step with the batching form (block Davidson, SpMM,
    hipGraphs, autotune) on a 2x4 Hubbard lattice (4+4 electrons) selected
    down to eps1 1e-7: its energies equal the lowest eigenvalues of the
    oracle's Hamiltonian over the selected determinants (dense eigensolver,
    1e-8 Ha), degenerate roots included.  (The reference's single-vector
    Davidson, which the plain path reproduces, can converge to a higher state
    or cycle to its iteration cap when roots are degenerate.)"""
    from dice_amd import synth, shci
    p = str(tmp_path / "FCIDUMP")
    synth.write_hubbard_fcidump(p, 2, 4, 1.0, 4.0, 4, 4)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, p, synth.neel_occupation(2, 4), [(0, 1e-4), (2, 1e-7)], 5, nroots=nroots,
                      davidson_tol=1e-8)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    r = shci.do_variational(inp, str(tmp_path / "b.log"), davidson_block=True, autotune=True, hip_graphs=True)
    ev = _dense_lowest(p, r.dets(), nroots) + shci.System.from_fcidump(p).core_energy
    print("block", r.energies(), "dense", ev, "n", len(r.dets()))
    np.testing.assert_allclose(np.sort(r.energies()), ev, rtol=0, atol=1e-8)


# ---- learned screening form ----

def test_learned_form_gain_over_untrained_network(tmp_path, monkeypatch):
    """The learned screening form on the 24-orbital synthetic system (eps1
    1e-2, five iterations): the network trained on the first two
    iterations' screening values keeps the candidates predicted within 0.6
    log10 units of eps1 and evaluates a small share of them exactly, while
    the same pipeline with the untrained network (control) would evaluate
    nearly all and falls back to full screening.  The energy gap of the
    learned run to the plain run is reported."""
    from dice_amd import synth, shci
    monkeypatch.setenv("SHCI_LEARNED_MARGIN", "0.6")
    path = str(tmp_path / "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c1"]
    synth.write_molecular_fcidump(path, c, n, na, nb, nchol)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(na, nb), [(0, 1e-2)], 5)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    plain = shci.do_variational(inp, str(tmp_path / "p.log")).energies()[0]
    r = shci.do_variational(inp, str(tmp_path / "l.log"), learned="forced")
    recs, train = r.learned()
    ctl = shci.do_variational(inp, str(tmp_path / "c.log"), learned="control")
    crecs, _ = ctl.learned()
    print("training", train)
    for a, b in zip(recs, crecs):
        print("trained", a, "control", b)
    print("energy gap learned - plain: %.3e Ha" % (r.energies()[0] - plain))
    assert train["samples"] > 1000 and train["last_loss"] < train["first_loss"] / 10
    # the control falls back at its first ranked iteration and screens in full
    # from then on; the trained network ranks every later iteration
    assert len(recs) >= 2 and len(crecs) == 1
    assert crecs[0]["fallback"] and crecs[0]["sampled_fraction"] > 0.9
    for a in recs:
        assert not a["fallback"]
        assert a["sampled_fraction"] < 0.5 * crecs[0]["sampled_fraction"]
        assert a["sampled_recall"] > 0.9
    # the control falls back to full screening every time: exact energy
    assert abs(ctl.energies()[0] - plain) < 1e-10
    # fewer determinants can only raise the variational energy
    assert r.energies()[0] >= plain - 1e-9


def test_learned_form_gate_turns_itself_off(tmp_path):
    """The gated learned form (the default) on a workload whose exact
    screen costs less than ranking its candidate records would: after the
    second iteration's exact screen it counts the records per source, turns
    itself off before that iteration logs anything (no training), leaves one
    fallback record, and gives the plain run's energy and space exactly."""
    from dice_amd import synth, shci
    path = str(tmp_path / "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c1"]
    synth.write_molecular_fcidump(path, c, n, na, nb, nchol)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(na, nb), [(0, 1e-2)], 4)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    p = shci.do_variational(inp, str(tmp_path / "p.log"))
    r = shci.do_variational(inp, str(tmp_path / "g.log"), learned=True)
    recs, train = r.learned()
    assert len(recs) == 1 and recs[0]["fallback"] and recs[0]["evaluated"] == 0 and recs[0]["candidates"] > 0
    assert recs[0]["iteration"] == 1
    assert train["seconds"] == 0.0
    np.testing.assert_array_equal(r.dets(), p.dets())
    assert r.energies()[0] == p.energies()[0]


def test_learned_form_default_margin_meets_the_energy_bar(tmp_path):
    """The learned form forced on at its default margin (0.8 log10 units
    below eps1) on the 24-orbital system (eps1 1e-2, five iterations): the
    energy stays within 0.1 mHa of the plain run, and every ranked
    iteration leaves a record (ranked, or fallen back to full screening when
    the network would keep more than half of the candidates).  At margin 0.6
    the same run ranks and misses 4.3 mHa (the trade-off the default
    avoids)."""
    from dice_amd import synth, shci
    path = str(tmp_path / "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c1"]
    synth.write_molecular_fcidump(path, c, n, na, nb, nchol)
    inp = str(tmp_path / "input.dat")
    synth.write_input(inp, path, synth.hf_occupation(na, nb), [(0, 1e-2)], 5)
    with open(inp, "a") as f:
        f.write("prefix %s\n" % tmp_path)
    plain = shci.do_variational(inp, str(tmp_path / "p.log")).energies()[0]
    r = shci.do_variational(inp, str(tmp_path / "l.log"), learned="forced")
    recs, train = r.learned()
    gap = r.energies()[0] - plain
    print("default margin: energy gap %.3e Ha; records %s" % (gap, recs))
    assert len(recs) >= 1 and train["samples"] > 1000
    assert abs(gap) < 1e-4


# ---- direct mode (no stored Hamiltonian) ----

@pytest.mark.parametrize("name", ["h8", "c0", "w2", "hub"])
def test_direct_sigma_matches_oracle(name, systems):
    """Sigma with every element evaluated from the string tables (HmultDirect
    orientation Hij(D_I, D_J)) equals the oracle's direct product and the
    stored-matrix product to 1e-12 relative, for one and for three vectors."""
    path = systems[name][0]
    dets = _space(name, systems, 2500, 1)
    g, o = _gpu(path), OracleSystem(path)
    g.set_options(direct_sigma=True)
    g.set_dets(dets)
    g.make_hamiltonian()
    o.csr_append(dets)
    x = np.random.default_rng(9).normal(size=(len(dets), 3))
    y = g.hmult(x)
    for r in range(3):
        yo = o.sigma_direct(np.ascontiguousarray(x[:, r]))
        assert np.max(np.abs(y[:, r] - yo)) <= 1e-12 * np.max(np.abs(yo))
        assert np.max(np.abs(y[:, r] - o.sigma(np.ascontiguousarray(x[:, r])))) <= 1e-12 * np.max(np.abs(yo))


O2_INPUT = ("nocc 12\n0 2 4 6 8 10 1 3 5 7 9 11\nend\nepsilon2 1e-07\nsampleN 200\ndavidsonTol 5e-05\n"
            "dE 1e-08\nschedule\n0 0.001\n3 0.0001\nend\ndeterministic\nnoio\nmaxiter 9\nnPTiter 0\n")


def test_run_o2_det_trusted(tmp_path):
    """tests/SHCI/o2_det: the trusted E_var to 1e-9 Ha, and the oracle's run
    (determinant sets, energies) on the GPU."""
    import json
    with open(os.path.join(GOLD, "golden.json")) as f:
        g = json.load(f)
    inp = _write_input(tmp_path, O2_INPUT + "orbitals %s\nprefix %s\n" % (os.path.join(GOLD, "o2", "FCIDUMP"),
                                                                       tmp_path))
    o, r = _compare_runs(inp, tmp_path, 1)
    assert abs(r.energies()[0] - g["trusted_energy"]["o2_det_var_plus_pt"]) < 1e-9


def test_run_direct_davidson_o2(tmp_path):
    """`directdavidson` (HmultDirect + davidsonDirect) on the o2_det_direct
    input: the oracle's determinant sets, per-iteration energies and
    Davidson iteration counts, and the direct mode's log lines."""
    from dice_amd import shci
    inp = _write_input(tmp_path, O2_INPUT + "orbitals %s\nprefix %s\ndirectdavidson\n" % (
        os.path.join(GOLD, "o2", "FCIDUMP"), tmp_path))
    o = OracleRun(inp, str(tmp_path / "oracle.log"), record=True)
    r = shci.do_variational(inp, str(tmp_path / "gpu.log"))
    oi, gi = o.iterations(), r.iterations()
    assert [x["ndets"] for x in oi] == [x["ndets"] for x in gi]
    assert [x["davidson"] for x in oi] == [x["davidson"] for x in gi]
    for a, b in zip(oi, gi):
        assert abs(a["energy"] - b["energy"]) < 1e-10
    np.testing.assert_allclose(r.energies(), o.energies(), rtol=0, atol=1e-10)
    np.testing.assert_array_equal(r.dets(), o.dets(1))
    log = open(tmp_path / "gpu.log").read()
    assert "New size of determinant space" in log and "nIter  Root               Energy" in log
    # progress lines of the final davidsonDirect match the oracle's: iteration
    # and root exactly, energy to 1e-10 relative, residual norm to 1e-6
    # relative (summation order), time column aside
    def prog(path):
        return [l.split()[:4] for l in open(path) if l[:5].strip().isdigit() and len(l.split()) == 5]
    a, b = prog(tmp_path / "gpu.log"), prog(tmp_path / "oracle.log")
    assert len(a) == len(b) > 0
    for u, v in zip(a, b):
        assert u[:2] == v[:2]
        assert abs(float(u[2]) - float(v[2])) <= 1e-10 * abs(float(v[2]))
        assert abs(float(u[3]) - float(v[3])) <= 1e-6 * abs(float(v[3]))


def test_run_output_files_match_oracle(tmp_path, monkeypatch):
    """writeBestDeterminants / printAllDeterminants (SHCI/SHCI.cpp:448-505) on
    the h8 golden input: the drop-in's dets.bin holds the oracle's 101
    determinants with the same coefficients (1e-9), in non-increasing |c|, and
    its log lists every determinant."""
    import json
    from dice_amd import shci
    g = json.load(open(os.path.join(GOLD, "golden.json")))
    fcidump = os.path.join(GOLD, "h8_631g", "FCIDUMP")
    res = {}
    for who in ("gpu", "oracle"):
        d = tmp_path / who
        d.mkdir()
        inp = d / "dice.dat"
        inp.write_text(g["h8_631g"]["input"] + "\norbitals %s\nprefix %s\nwriteBestDeterminants 101\n"
                       "printAllDeterminants\nnPTiter 0\n" % (fcidump, d))
        monkeypatch.chdir(d)
        if who == "gpu":
            r = shci.do_variational(str(inp), str(d / "log"))
            n = len(r.dets())
        else:
            OracleRun(str(inp), str(d / "log"))
        res[who] = helpers.read_dets_bin(str(d / "dets.bin"))
        log = open(d / "log").read()
        block = log.split("Printing all determinants")[1].split("Printing most important")[0]
        res[who + "_lines"] = len([l for l in block.splitlines()
                                   if l.strip() and l.strip()[0] in "-0123456789" and "e" in l.split()[0]])
    (ng, rg), (no, ro) = res["gpu"], res["oracle"]
    assert ng == no == 16 and len(rg) == len(ro) == 101
    mine = {s: c for c, s in rg}
    ref = {s: c for c, s in ro}
    assert set(mine) == set(ref)
    for s, c in ref.items():
        assert abs(mine[s] - c) < 1e-9
    for (c0, _), (c1, _) in zip(rg, rg[1:]):
        assert abs(c1) <= abs(c0)
    assert res["gpu_lines"] == res["oracle_lines"] == n


def test_run_restart_chain_trusted(tmp_path):
    """tests/SHCI/restart as the reference runs it, on the GPU: input2 (io)
    checkpoints every iteration, input3 (restart) continues the schedule to
    eps1 5e-5; shci.e matches trusted_hci.e (runner tolerance 5e-5)."""
    import json
    import struct
    from dice_amd import shci
    g = json.load(open(os.path.join(GOLD, "golden.json")))
    fcidump = os.path.join(GOLD, "restart", "FCIDUMP")
    for k, name in enumerate(("restart_input2", "restart_input3")):
        text = "\n".join(l for l in g[name].splitlines() if not l.strip().lower().startswith("prefix"))
        p = tmp_path / ("in%d.dat" % k)
        p.write_text(text + "\norbitals %s\nprefix %s\n" % (fcidump, tmp_path))
        shci.do_variational(str(p), str(tmp_path / ("log%d" % k)))
    with open(tmp_path / "shci.e", "rb") as f:
        e = struct.unpack("<d", f.read(8))[0]
    assert abs(e - g["trusted_energy"]["restart"]) < 1e-6, e


def test_run_restart_from_oracle_checkpoint_moves_to_pt(tmp_path):
    """The checkpoint format is shared: the GPU restarts from the oracle's
    checkpoint of an O2 run at eps1 1e-2, skips the iterations (converged at
    or below its own eps1 2e-2) and gives the oracle's E_var and E_var + E_PT
    (1e-10 Ha)."""
    from dice_amd import shci
    from test_oracle_pins import O2_INPUT
    fcidump = os.path.join(GOLD, "o2", "FCIDUMP")
    base = O2_INPUT.replace("deterministic\n", "").replace("noio\n", "").replace(
        "schedule\n0 0.001\n3 0.0001\nend\n", "schedule\n0 0.01\nend\n")
    (tmp_path / "a.dat").write_text(base + "io\norbitals %s\nprefix %s\n" % (fcidump, tmp_path))
    OracleRun(str(tmp_path / "a.dat"), str(tmp_path / "a.log"))
    b = base.replace("schedule\n0 0.01\nend\n", "schedule\n0 0.02\nend\n")
    (tmp_path / "b.dat").write_text(b + "restart\ndeterministic\nnoio\norbitals %s\nprefix %s\n" % (fcidump, tmp_path))
    rg = shci.do_variational(str(tmp_path / "b.dat"), str(tmp_path / "g.log"))
    ro = OracleRun(str(tmp_path / "b.dat"), str(tmp_path / "o.log"))
    assert "moving to perturbative part" in open(tmp_path / "g.log").read()
    np.testing.assert_allclose(rg.energies(), ro.energies(), rtol=0, atol=1e-10)
    np.testing.assert_allclose(rg.pt_energies(), ro.pt_energies(), rtol=0, atol=1e-10)
    assert rg.pt_energies()[0] < rg.energies()[0] - 1e-6
