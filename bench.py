"""Benchmark: SHCI variational iterations on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]

Workload: the 24-orbital synthetic system (12 alpha + 12 beta electrons, 96
Cholesky vectors, random stream 1000+1), heat-bath selection from the HF
determinant with eps1 = 5e-3 for six variational iterations (the space grows
to about 1.18 million determinants, 0.91e9 stored lower-triangle elements).
One step = one complete variational step: the six iterations (screening,
merge, Hamiltonian build, Davidson) and the final tight Davidson, with the
integrals resident on the GPU.  `value` is the mean wall time of one
variational iteration over the timed steps.

The roofline object is the sigma = H x kernel of the Davidson, timed with
HIP events on the library's stream over the timed steps.  The Hamiltonian's
lower triangle alone is stored (the library's default for the single-vector
Davidson): kSigmaTiledSym over a tiled copy of it (256-row blocks of the
alpha-major relabelled rows, entries grouped by 1024-column tiles), each
stored element applied to its row and, in 64-bit fixed point, to its
column, with its x permutation and the final addition into y.  Achieved
bandwidth uses algorithmic bytes per launch: stored values + column
indices, row offsets, diagonal, x read once and y written once.  `traffic`
is the HBM
bytes per sigma launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
this benchmark (scaled by a copy kernel of known size), read from
profiles/sigma_traffic_c1_<precision>.json (see profiles/sigma_traffic_c1.md) when that summary exists.

pt_deterministic: the deterministic EN-PT2 step (after the variational one)
on the final space at eps2 = 1e-4, the semistochastic one (deterministic
part at 1e-4, 50 sampled iterations down to 1e-6), and GPU vs CPU oracle on
the first iteration's space at eps2 = 1e-3.  rdm (inside it): the spatial
1- and 2-RDMs of the final state (one call with the connections stored), and
GPU vs the C restatement of EvaluateOneRDM / EvaluateRDM (one thread) on the
first iteration's space.

cpu_baseline: the CPU oracle (a restatement of the reference algorithm whose
rank-parallel loops run on OpenMP threads, up to 16) on a bounded sample of
the same workload: the first variational iteration from HF (27,685
determinants) and the final tight Davidson; the GPU time of the same sample
is reported beside it.
"""
import argparse
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
EPS1 = 5e-3
NITER = 6
PT_EPS2 = 1e-4
PT_SAMPLE_EPS2 = 1e-3  # keeps the CPU oracle's PT sample to tens of seconds
PT_STOCH_EPS2 = 1e-6


def traffic_bytes(precision):
    """HBM bytes per sigma launch from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", "sigma_traffic_c1_%s.json" % precision)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["traffic_bytes_per_dispatch"], 0)


def workload_inputs(directory):
    from dice_amd import synth
    os.makedirs(directory, exist_ok=True)
    fcidump = os.path.join(directory, "FCIDUMP")
    if not os.path.exists(fcidump):
        c, n, na, nb, nchol = synth.MOLECULAR["c1"]
        synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    full = os.path.join(directory, "input.dat")
    synth.write_input(full, fcidump, synth.hf_occupation(12, 12), [(0, EPS1)], NITER, extra=["prefix " + directory])
    sample = os.path.join(directory, "sample.dat")
    synth.write_input(sample, fcidump, synth.hf_occupation(12, 12), [(0, EPS1)], 1, extra=["prefix " + directory])
    return full, sample


def run_form(shci, full, log, prec, steps, warmup, tag=None, profile=True, **method):
    """Warm up, then time `steps` complete variational steps in one precision
    (method: the batching / learned options of do_variational)."""
    tag = tag or ("fp32" if prec == shci.PREC_FP32 else "fp64")
    for w in range(warmup):
        print("# %s warmup %d/%d" % (tag, w + 1, warmup), file=sys.stderr, flush=True)
        shci.do_variational(full, log, precision=prec, **method)
    R = {"iter_times": [], "step_times": [], "calls": 0, "ms": 0.0, "nbytes": 0.0, "energies": [], "per_iter": None,
         "sigma_loose": None}
    for k in range(steps):
        t0 = time.time()
        r = shci.do_variational(full, log, precision=prec, profile_sigma=profile, **method)
        wall = time.time() - t0
        ph = r.phases()
        its = r.iterations()
        tt = [a + b + c for a, b, c, _ in ph]
        R["iter_times"] += tt
        R["step_times"].append(sum(tt) + r.final_davidson_seconds())
        c, m, b = r.sigma_profile()
        R["calls"] += c
        R["ms"] += m
        R["nbytes"] += b
        R["energies"].append(float(r.energies()[0]))
        R["all_energies"] = [float(e) for e in r.energies()]
        if method.get("learned"):
            R["learned"] = r.learned()
        R["per_iter"] = [{"ndets": it["ndets"], "nnz_lower": p[3], "davidson_iters": it["davidson"],
                          "screen_s": round(p[0], 4), "hamiltonian_s": round(p[1], 4), "davidson_s": round(p[2], 4)}
                         for it, p in zip(its, ph)]
        print("# %s step %d: %.3f s (wall %.3f s), E = %.10f; per iteration (screen, H, Davidson s / iters): %s"
              % (tag, k + 1, R["step_times"][-1], wall, R["energies"][-1],
                 " ".join("%.3f,%.3f,%.3f/%d" % (p[0], p[1], p[2], it["davidson"]) for it, p in zip(its, ph))),
              file=sys.stderr, flush=True)
    return R


def batching_form(shci, args):
    """Batching form on the 6x6 Hubbard lattice (18+18 electrons, t = 1,
    U = 4, site basis, Neel reference), eps1 2e-4, four roots, three
    variational iterations (428,641 determinants): the reference's Davidson
    vs the block Davidson (one sigma per block of corrections,
    hipGraph-replayed iterations), both on the default half storage; and
    the block's sigma of four vectors alone, on half storage (one
    kSigmaTiledSym per vector) and on both triangles (one SpMM)."""
    from dice_amd import synth
    d = os.path.join(tempfile.gettempdir(), "shci_bench_hubbard6x6")
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "FCIDUMP")
    if not os.path.exists(p):
        synth.write_hubbard_fcidump(p, 6, 6, 1.0, 4.0, 18, 18)
    inp = os.path.join(d, "input.dat")
    synth.write_input(inp, p, synth.neel_occupation(6, 6), [(0, 2e-4)], 3, nroots=4, extra=["prefix " + d])
    log = os.path.join(d, "gpu.log")
    res = {}
    graphs = not args.no_graphs
    for name, method in (("reference_davidson", {}),
                         ("block_davidson", dict(davidson_block=True, autotune=True, hip_graphs=graphs))):
        # (per-launch sigma events would keep the block Davidson from replaying graphs)
        R = run_form(shci, inp, log, shci.PREC_FP64, 1, 1, tag="hubbard " + name, profile=not method, **method)
        res[name] = {
            "s_per_iteration": round(statistics.mean(R["iter_times"]), 6),
            "ms_per_step": round(1e3 * statistics.mean(R["step_times"]), 3),
            "davidson_s": round(sum(it["davidson_s"] for it in R["per_iter"]), 4),
            "energies": R["all_energies"],
            "sigma_GBps": round(R["nbytes"] / (R["ms"] * 1e-3) / 1e9, 1) if R["ms"] else None,
            "iterations": R["per_iter"],
        }
    # the block Davidson's sigma alone, k = 4 vectors on the final
    # Hamiltonian (wall time of shci_sigma_device, x permutation included):
    # on half storage four launches over the stored half, on both triangles
    # one SpMM of 4 interleaved vectors over the row-ordered relabelled
    # copy; algorithmic bytes per pass over the matrix as for the
    # one-vector sigma, with 2 k N vector values
    for key, half in (("sigma_k4_half_storage", None), ("spmm_k4_full_storage", False)):
        r = shci.do_variational(inp, log, davidson_block=True, autotune=True, hip_graphs=graphs, half_storage=half)
        n, nnz_lower = r.iterations()[-1]["ndets"], r.phases()[-1][3]
        t = time_sigma_device(shci, r._h, n, 4, 20)
        del r
        if half is None:
            stored = nnz_lower - n
            nbytes = 4 * (stored * 12 + (n + 1) * 8 + n * 8 + 2 * n * 8)
        else:
            stored = 2 * (nnz_lower - n)
            nbytes = stored * 12 + (n + 1) * 8 + n * 8 + 2 * 4 * n * 8
        res["block_davidson"][key] = {"ndets": n, "nnz_stored": stored, "ms": round(1e3 * t, 4), "bytes": nbytes,
                                      "GBps": round(nbytes / t / 1e9, 1),
                                      "frac_of_hbm_peak": round(nbytes / t / 1e9 / PEAK_HBM_GBS, 4)}
        print("# Hubbard %s: %.3f ms, %.1f GB/s" % (key, 1e3 * t, nbytes / t / 1e9), file=sys.stderr, flush=True)
    res["workload"] = ("6x6 periodic Hubbard, t=1, U=4, 18a+18b, eps1 2e-4, 4 roots, 3 iterations "
                       "(the fourth would select 12.6M more determinants)")
    res["davidson_speedup"] = round(res["reference_davidson"]["davidson_s"] / res["block_davidson"]["davidson_s"], 3)
    return res


def batching_fourth_iteration(shci, args):
    """The Hubbard batching workload's fourth iteration (eps1 2e-4 x 4):
    12.6M determinants join the 428,641 of the third, and the Hamiltonian
    grows to about 14.5e9 stored elements.  The block Davidson (four roots,
    one SpMM per block) on whatever storage fits: the full matrix without a
    relabelled copy, or the lower triangle alone (half storage, chosen by
    the library when the full one does not fit)."""
    from dice_amd import synth
    d = os.path.join(tempfile.gettempdir(), "shci_bench_hubbard6x6")
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "FCIDUMP")
    if not os.path.exists(p):
        synth.write_hubbard_fcidump(p, 6, 6, 1.0, 4.0, 18, 18)
    inp = os.path.join(d, "input4.dat")
    synth.write_input(inp, p, synth.neel_occupation(6, 6), [(0, 2e-4)], 4, nroots=4, extra=["prefix " + d])
    log = os.path.join(d, "gpu4.log")
    print("# Hubbard fourth iteration", file=sys.stderr, flush=True)
    t0 = time.time()
    r = shci.do_variational(inp, log, davidson_block=True, autotune=True, hip_graphs=not args.no_graphs)
    wall = time.time() - t0
    ph, its = r.phases(), r.iterations()
    res = {"workload": "6x6 periodic Hubbard, t=1, U=4, 18a+18b, eps1 2e-4, 4 roots, 4 iterations, block Davidson",
           "wall_s": round(wall, 3), "energies": [float(e) for e in r.energies()],
           "iterations": [{"ndets": it["ndets"], "nnz_lower": p_[3], "davidson_iters": it["davidson"],
                           "screen_s": round(p_[0], 4), "hamiltonian_s": round(p_[1], 4),
                           "davidson_s": round(p_[2], 4)} for it, p_ in zip(its, ph)],
           "final_davidson_s": round(r.final_davidson_seconds(), 4)}
    del r
    print("# Hubbard fourth iteration: %.1f s, %d dets" % (wall, res["iterations"][-1]["ndets"]),
          file=sys.stderr, flush=True)
    return res


C3_EPS1 = 1e-2   # the 40-orbital system's eps1 (see c3_form)
C3_NITER = 2
C4_EPS1 = 2e-2   # the learned form on the 40-orbital system (see c4_form)
C4_NITER = 3


def c3_form(shci, args):
    """The 40-orbital system (20 alpha + 20 beta electrons, 160 Cholesky
    vectors, random stream 1003, two-word determinants) with two roots: one
    variational step of C3_NITER iterations at eps1 C3_EPS1 from HF, in FP64
    and in the FP32 form (the reference's single-vector Davidson).  The
    configured schedule (down to 2e-5) selects far more than one GPU holds
    on this synthetic construction; C3_EPS1 is the largest-space schedule
    measured to fit and finish within the benchmark's time.  The FP64 run
    stores the lower triangle alone (the default); an FP64 run with both
    triangles stored follows."""
    from dice_amd import synth
    d = os.path.join(tempfile.gettempdir(), "shci_bench_c3")
    os.makedirs(d, exist_ok=True)
    fcidump = os.path.join(d, "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c3"]
    if not os.path.exists(fcidump):
        synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    inp = os.path.join(d, "input.dat")
    synth.write_input(inp, fcidump, synth.hf_occupation(na, nb), [(0, C3_EPS1)], C3_NITER, nroots=2,
                      extra=["prefix " + d])
    log = os.path.join(d, "gpu.log")
    res = {"workload": "40-orbital synthetic FCIDUMP (20a+20b, 160 Cholesky vectors, stream 1003), 2 roots, "
                       "heat-bath eps1 %.0e x %d iterations from HF" % (C3_EPS1, C3_NITER)}
    runs = (("fp64", shci.PREC_FP64, None), ("fp32", shci.PREC_FP32, None),
            ("fp64_full_storage", shci.PREC_FP64, False))
    for tag, prec, half in runs:
        print("# C3 %s" % tag, file=sys.stderr, flush=True)
        t0 = time.time()
        r = shci.do_variational(inp, log, precision=prec, profile_sigma=True, half_storage=half)
        wall = time.time() - t0
        ph, its = r.phases(), r.iterations()
        calls, ms, nbytes = r.sigma_profile()
        res[tag] = {
            "wall_s": round(wall, 3),
            "step_s": round(sum(a + b + c for a, b, c, _ in ph) + r.final_davidson_seconds(), 3),
            "energies": [float(e) for e in r.energies()],
            "final_ndets": its[-1]["ndets"], "nnz_lower": ph[-1][3],
            "sigma_launches": calls, "sigma_avg_ms": round(ms / calls, 4) if calls else None,
            "sigma_GBps": round(nbytes / (ms * 1e-3) / 1e9, 1) if ms else None,
            "sigma_frac_of_hbm_peak": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4) if ms else None,
            "iterations": [{"ndets": it["ndets"], "nnz_lower": p[3], "davidson_iters": it["davidson"],
                            "screen_s": round(p[0], 4), "hamiltonian_s": round(p[1], 4),
                            "davidson_s": round(p[2], 4)} for it, p in zip(its, ph)],
        }
        del r
        print("# C3 %s: %.1f s, %d dets, E %s" % (tag, wall, res[tag]["final_ndets"], res[tag]["energies"]),
              file=sys.stderr, flush=True)
    res["fp64"]["description"] = (
        "the lower triangle alone (the reference's layout, half the matrix bytes); sigma on its tiled copy "
        "(kSigmaTiledSym, the transposed half in 64-bit fixed point); sigma bytes count the stored lower triangle")
    res["fp64_full_storage"]["description"] = (
        "both triangles stored; sigma on their tiled copy (kSigmaTiled); sigma bytes count both triangles")
    res["fp64_full_storage"]["speedup_vs_fp64"] = round(
        res["fp64"]["step_s"] / res["fp64_full_storage"]["step_s"], 3)
    res["fp64_full_storage"]["energy_error_vs_fp64_Ha"] = [
        a - b for a, b in zip(res["fp64_full_storage"]["energies"], res["fp64"]["energies"])]
    res["fp32"]["energy_error_vs_fp64_Ha"] = [a - b for a, b in zip(res["fp32"]["energies"], res["fp64"]["energies"])]
    res["fp32"]["speedup_vs_fp64"] = round(res["fp64"]["step_s"] / res["fp32"]["step_s"], 3)
    return res


C0_EPS1 = 1e-3
C0_NITER = 4


def c0_form(shci, args):
    """The 16-orbital system (8 alpha + 8 beta electrons, 64 Cholesky
    vectors, random stream 1000) at eps1 C0_EPS1 for C0_NITER iterations
    from HF in FP64 (its configured second eps1, 1e-4, selects more than
    one GPU holds on this synthetic construction), and its first iteration
    (5,695 determinants, plus the final tight Davidson) on the GPU and on
    the CPU oracle with the host threads of the cpu_baseline leg."""
    from dice_amd import synth
    d = os.path.join(tempfile.gettempdir(), "shci_bench_c0")
    os.makedirs(d, exist_ok=True)
    fcidump = os.path.join(d, "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c0"]
    if not os.path.exists(fcidump):
        synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    inp = os.path.join(d, "input.dat")
    synth.write_input(inp, fcidump, synth.hf_occupation(na, nb), [(0, C0_EPS1)], C0_NITER, extra=["prefix " + d])
    first = os.path.join(d, "first.dat")
    synth.write_input(first, fcidump, synth.hf_occupation(na, nb), [(0, C0_EPS1)], 1, extra=["prefix " + d])
    log = os.path.join(d, "gpu.log")
    shci.do_variational(inp, log)  # warm-up
    t0 = time.time()
    r = shci.do_variational(inp, log)
    wall = time.time() - t0
    ph, its = r.phases(), r.iterations()
    res = {"workload": "16-orbital synthetic FCIDUMP (8a+8b, 64 Cholesky vectors, stream 1000), 1 root, heat-bath "
                       "eps1 %.0e x %d iterations from HF, FP64" % (C0_EPS1, C0_NITER),
           "wall_s": round(wall, 3),
           "iter_s": [round(a + b + c, 4) for a, b, c, _ in ph],
           "final_ndets": its[-1]["ndets"], "energy": float(r.energies()[0]),
           "iterations": [{"ndets": it["ndets"], "nnz_lower": p[3], "davidson_iters": it["davidson"]}
                          for it, p in zip(its, ph)]}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import OracleRun
        threads = host_threads()
        os.environ["ORACLE_THREADS"] = str(threads)
        t0 = time.time()
        OracleRun(first, os.path.join(d, "oracle.log"))
        cpu = time.time() - t0
        t0 = time.time()
        r1 = shci.do_variational(first, log)
        gpu = time.time() - t0
        res["first_iteration"] = {"ndets": r1.iterations()[-1]["ndets"], "cpu_oracle_s": round(cpu, 3),
                                  "cpu_threads": threads, "gpu_s": round(gpu, 4)}
    return res


def c4_form(shci, args):
    """The learned form on the 40-orbital system (two roots, FP32 values):
    C4_NITER iterations at eps1 C4_EPS1 from HF (the learned form logs two
    iterations and ranks from the third on), gated (the default) and forced
    on, beside the plain FP32 run of the same schedule.  Reported: the
    energy gap of each to the plain run, each ranked iteration's record,
    and the screen's share of the plain step, which caps what any faster
    screen can save."""
    from dice_amd import synth
    d = os.path.join(tempfile.gettempdir(), "shci_bench_c3")
    os.makedirs(d, exist_ok=True)
    fcidump = os.path.join(d, "FCIDUMP")
    c, n, na, nb, nchol = synth.MOLECULAR["c3"]
    if not os.path.exists(fcidump):
        synth.write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
    inp = os.path.join(d, "input_c4.dat")
    synth.write_input(inp, fcidump, synth.hf_occupation(na, nb), [(0, C4_EPS1)], C4_NITER, nroots=2,
                      extra=["prefix " + d])
    log = os.path.join(d, "gpu_c4.log")
    res = {"workload": "40-orbital synthetic FCIDUMP (as c3_large), 2 roots, FP32 values, heat-bath eps1 %.0e x %d "
                       "iterations from HF; learned form gated and forced" % (C4_EPS1, C4_NITER)}
    runs = {}
    for tag, learned in (("plain", False), ("gated", True), ("forced", "forced")):
        print("# C4 %s" % tag, file=sys.stderr, flush=True)
        t0 = time.time()
        r = shci.do_variational(inp, log, precision=shci.PREC_FP32, learned=learned)
        wall = time.time() - t0
        ph, its = r.phases(), r.iterations()
        runs[tag] = {
            "wall_s": round(wall, 3),
            "step_s": round(sum(a + b + c for a, b, c, _ in ph) + r.final_davidson_seconds(), 3),
            "screen_s": round(sum(p[0] for p in ph), 4),
            "energies": [float(e) for e in r.energies()],
            "final_ndets": its[-1]["ndets"],
            "iterations": [{"ndets": it["ndets"], "davidson_iters": it["davidson"], "screen_s": round(p[0], 4),
                            "hamiltonian_s": round(p[1], 4), "davidson_s": round(p[2], 4)}
                           for it, p in zip(its, ph)],
        }
        if learned:
            recs, train = r.learned()
            runs[tag]["ranked_iterations"] = recs
            runs[tag]["training"] = train
        del r
        print("# C4 %s: %.1f s, %d dets, E %s" % (tag, wall, runs[tag]["final_ndets"], runs[tag]["energies"]),
              file=sys.stderr, flush=True)
    plain = runs["plain"]
    for tag in ("gated", "forced"):
        runs[tag]["energy_gap_vs_plain_Ha"] = [a - b for a, b in zip(runs[tag]["energies"], plain["energies"])]
        runs[tag]["speedup_vs_plain"] = round(plain["step_s"] / runs[tag]["step_s"], 3)
    res.update(runs)
    res["screen_share_of_plain_step"] = round(plain["screen_s"] / plain["step_s"], 4)
    return res


def hip_runtime():
    import ctypes
    H = ctypes.CDLL("libamdhip64.so")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


def time_sigma_device(shci, handle, n, k, reps):
    """Mean wall time (s) of shci_sigma_device with k vectors on the handle's
    Hamiltonian (device vectors, synchronised around the repetitions; the x
    permutation kernel included)."""
    import ctypes
    import numpy as np
    L, H = shci.library(), hip_runtime()
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dx), n * k * 8) == 0 and H.hipMalloc(ctypes.byref(dy), n * k * 8) == 0
    try:
        x = np.random.default_rng(1).normal(size=n * k)
        H.hipMemcpy(dx, x.ctypes.data, n * k * 8, 1)
        assert L.shci_sigma_device(handle, dx, dy, k, n) == 0
        H.hipDeviceSynchronize()
        t0 = time.time()
        for _ in range(reps):
            assert L.shci_sigma_device(handle, dx, dy, k, n) == 0
        H.hipDeviceSynchronize()
        return (time.time() - t0) / reps
    finally:
        H.hipFree(dx)
        H.hipFree(dy)


def host_threads():
    """Host threads for the oracle's rank-parallel loops: the box's CPU share
    (OMP_NUM_THREADS, set per GPU), else the host's core count."""
    return max(1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))


def cpu_final_space(shci, full, args):
    """The reference's algorithm (the oracle, rank-parallel loops on the host
    threads) on the final space of the workload: its Hamiltonian build
    (MakeHfromSMHelpers2) and its sigma (Hmult2), against a from-scratch GPU
    build of the same space and the GPU sigma."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleSystem
    threads = host_threads()
    os.environ["ORACLE_THREADS"] = str(threads)
    r = shci.do_variational(full, os.path.join(os.path.dirname(full), "gpu.log"))
    dets = r.dets()
    n = len(dets)
    fcidump = os.path.join(os.path.dirname(full), "FCIDUMP")
    G = shci.System.from_fcidump(fcidump)
    G.set_dets(dets)
    t0 = time.time()
    nnz_lower = G.make_hamiltonian()
    gpu_build = time.time() - t0
    gpu_sigma = time_sigma_device(shci, G._h, n, 1, 20)
    x = np.random.default_rng(2).normal(size=n)
    y_gpu = G.hmult(x)
    G.close()
    del r
    o = OracleSystem(fcidump)
    t0 = time.time()
    o.csr_append(dets)
    cpu_build = time.time() - t0
    y_cpu = o.sigma(x)
    t0 = time.time()
    reps = 3
    for _ in range(reps):
        o.sigma(x)
    cpu_sigma = (time.time() - t0) / reps
    del o
    # the two products on the full space: a parity check at the workload's size
    sigma_rel_diff = float(np.max(np.abs(y_gpu - y_cpu)) / np.max(np.abs(y_cpu)))
    print("# final space %d dets: build CPU %.2f s / GPU %.3f s; sigma CPU %.3f s / GPU %.2f ms (%d threads)"
          % (n, cpu_build, gpu_build, cpu_sigma, 1e3 * gpu_sigma, threads), file=sys.stderr, flush=True)
    return {"what": "the reference's Hamiltonian build (MakeHfromSMHelpers2, rows round-robin) and sigma "
                    "(Hmult2, lower-triangle rows applied symmetrically, per-thread partial vectors summed as the "
                    "per-rank ones) restated in the oracle, on the workload's final space, against a from-scratch "
                    "GPU build of the same space and one GPU sigma (shci_sigma_device, x permutation included)",
            "ndets": n, "nnz_lower": nnz_lower, "cores": threads, "host_nproc": os.cpu_count(),
            "cpu_build_s": round(cpu_build, 3), "gpu_build_s": round(gpu_build, 4),
            "cpu_sigma_s": round(cpu_sigma, 4), "gpu_sigma_s": round(gpu_sigma, 6),
            "sigma_speedup": round(cpu_sigma / gpu_sigma, 1), "build_speedup": round(cpu_build / gpu_build, 1),
            "sigma_max_rel_diff_gpu_vs_cpu": sigma_rel_diff}


def pt_measure(shci, full, sample, log, args):
    """Deterministic EN-PT2 (the step after the variational one) on the final
    FP64 space of the workload at eps2 = PT_EPS2, and on the first-iteration
    sample at eps2 = PT_SAMPLE_EPS2 on the GPU and in the CPU oracle (16
    threads at most)."""
    r = shci.do_variational(full, log)
    dets, ci, e = r.dets(), r.ci()[:, 0], float(r.energies()[0])
    del r
    S = shci.System.from_fcidump(os.path.join(os.path.dirname(full), "FCIDUMP"))
    core = S.core_energy
    S.set_dets(dets)
    S.pt_deterministic(ci, e - core, PT_EPS2)  # warm-up
    t0 = time.time()
    ept, psi1, next_ = S.pt_deterministic(ci, e - core, PT_EPS2)
    gpu_s = time.time() - t0
    # semistochastic: the deterministic part at PT_EPS2, 50 sampled iterations
    # of 200 draws down to PT_STOCH_EPS2
    t0 = time.time()
    sval, per, its, serr = S.pt_semistochastic(ci, e - core, PT_STOCH_EPS2, PT_EPS2, 200, 50, -1.0, 1)
    stoch_s = time.time() - t0
    # spatial 1- and 2-RDMs of the final state (DoRDM): the first call builds
    # the connections of the held determinants, the second is timed
    S.rdm(ci)
    t0 = time.time()
    g1, g2 = S.rdm(ci)
    rdm_s = time.time() - t0
    S.close()
    res = {"eps2": PT_EPS2, "ndets": len(dets), "seconds": round(gpu_s, 4), "E_PT": ept, "E_var_plus_PT": e + ept,
           "external_dets": next_, "external_per_s": round(next_ / gpu_s, 1),
           "semistochastic": {"eps2": PT_STOCH_EPS2, "eps2_large": PT_EPS2, "sample_n": 200, "iterations": its,
                              "seconds": round(stoch_s, 3),
                              "seconds_per_iteration": round((stoch_s - gpu_s) / max(its, 1), 5),
                              "E_PT_mean": float(per.mean()) if len(per) else None, "std_error": serr},
           "rdm": {"what": "spatial 1- and 2-RDM of the final state (DoRDM), one call with the connections stored",
                   "seconds": round(rdm_s, 4), "trace_1rdm": float(g1.trace()),
                   "trace_2rdm": float(g2.reshape(g1.size, g1.size).trace())}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import OracleRun, OracleSystem
        threads = host_threads()
        os.environ["ORACLE_THREADS"] = str(threads)
        fcidump = os.path.join(os.path.dirname(full), "FCIDUMP")
        run = OracleRun(sample, None)
        o = OracleSystem(fcidump)
        sd, sc, se = run.dets(o.words), run.ci()[:, 0], float(run.energies()[0]) - core
        t0 = time.time()
        cpu_pt = o.pt_deterministic(sd, sc, se, PT_SAMPLE_EPS2)
        cpu_s = time.time() - t0
        G = shci.System.from_fcidump(fcidump)
        G.set_dets(sd)
        G.pt_deterministic(sc, se, PT_SAMPLE_EPS2)
        t0 = time.time()
        gpu_pt = G.pt_deterministic(sc, se, PT_SAMPLE_EPS2)
        gpu_sample_s = time.time() - t0
        # RDMs of the same sample: the reference's EvaluateOneRDM / EvaluateRDM
        # restated in C (one thread, as one reference rank) vs the GPU
        G.set_dets(sd)
        G.rdm(sc)
        t0 = time.time()
        gg1, gg2 = G.rdm(sc)
        gpu_rdm_s = time.time() - t0
        G.close()
        o.csr_append(sd)
        t0 = time.time()
        og1, og2 = o.rdm(sc)
        cpu_rdm_s = time.time() - t0
        res["cpu_sample"] = {"sample": "first variational iteration's space (%d determinants), eps2 %.0e"
                                       % (len(sd), PT_SAMPLE_EPS2),
                             "cpu_s": round(cpu_s, 3), "cores": threads, "kind": "port", "gpu_s": round(gpu_sample_s, 4),
                             "E_PT_cpu": cpu_pt[0], "E_PT_gpu": gpu_pt[0], "external_dets": cpu_pt[2]}
        res["rdm"]["cpu_sample"] = {"sample": "the same %d determinants" % len(sd), "cpu_s": round(cpu_rdm_s, 3),
                                    "cores": 1, "kind": "port", "gpu_s": round(gpu_rdm_s, 4),
                                    "max_abs_diff_2rdm": float(abs(gg2 - og2).max())}
    print("# PT eps2 %.0e: %.3f s, %d external determinants, E_PT %.10f" % (PT_EPS2, gpu_s, next_, ept),
          file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-forms", action="store_true", help="time only the --precision form")
    ap.add_argument("--davidson-block", action="store_true",
                    help="the --precision form with the block Davidson and hipGraph replay (batching form)")
    ap.add_argument("--post-variational", action="store_true",
                    help="only the steps after the variational one (PT, RDMs) on the final space, printed as "
                         "their own JSON object (for profiling them)")
    ap.add_argument("--form", choices=["c0", "c2", "c2big", "c3", "c4"],
                    help="only that configuration's line (c0: 16 orbitals; c2: the Hubbard batching form; c2big: its "
                         "fourth iteration; c3: 40 orbitals; c4: the "
                         "learned form on them), printed as its own JSON object (for profiling)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="block Davidson forms without hipGraph replay (rocprofv3's kernel tracer does not "
                         "survive graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the variational step on a single MI355X "
              "(the row-block sharding over RCCL is a design, not built); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from dice_amd import shci
    work = os.path.join(tempfile.gettempdir(), "shci_bench_c1")
    full, sample = workload_inputs(work)
    log = os.path.join(work, "gpu.log")
    primary = shci.PREC_FP32 if args.precision == "fp32" else shci.PREC_FP64
    if args.post_variational:
        print(json.dumps({"post_variational": pt_measure(shci, full, sample, log, args)}))
        return
    if args.form:
        fn = {"c0": c0_form, "c2": batching_form, "c2big": batching_fourth_iteration, "c3": c3_form,
              "c4": c4_form}[args.form]
        print(json.dumps({args.form: fn(shci, args)}))
        return
    results = {}
    forms = [primary] if args.no_forms else [primary] + [p for p in (shci.PREC_FP64, shci.PREC_FP32) if p != primary]
    block = dict(davidson_block=True, hip_graphs=not args.no_graphs) if args.davidson_block else {}
    # the headline form over the requested steps, the other forms over at
    # most 5 steps after one warm-up (they are reported beside it)
    fsteps, fwarm = min(args.steps, 5), min(args.warmup, 1)
    for prec in forms:
        results[prec] = run_form(shci, full, log, prec, args.steps if prec == primary else fsteps,
                                 args.warmup if prec == primary else fwarm, profile=not block,
                                 **(block if prec == primary else {}))
    prec = primary
    R = results[prec]
    iter_times, step_times, calls, ms, nbytes, energies, per_iter = (
        R["iter_times"], R["step_times"], R["calls"], R["ms"], R["nbytes"], R["energies"], R["per_iter"])
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    value = statistics.mean(iter_times)
    out = {
        "metric": "SHCI var-iter wall time + sigma=Hc GB/s vs plain FP64 GPU at <0.1 mHa error",
        "value": round(value, 6),
        "unit": "s per variational iteration",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * statistics.mean(step_times), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == shci.PREC_FP64 else "f32 values, f64 accumulation",
        "data": "synthetic",
        "config": {
            "workload": "24-orbital synthetic FCIDUMP (12a+12b, 96 Cholesky vectors, stream 1001), "
                        "heat-bath eps1 %.0e x %d iterations from HF, ~1.18M determinants" % (EPS1, NITER),
            "precision": args.precision,
            "final_ndets": per_iter[-1]["ndets"] if per_iter else None,
            "final_energy": energies[-1] if energies else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "kSigmaTiledSym (sigma = H x over the stored lower triangle: 256-row blocks of the "
                      "alpha-major relabelled rows, entries grouped by 1024-column tiles, each entry applied to its "
                      "row and, in 64-bit fixed point, to its column)",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic_bytes(args.precision),
            "launches": calls,
            "avg_launch_ms": round(ms / calls, 4) if calls else None,
            "bytes_per_launch": nbytes / calls if calls else None,
        },
        "iterations": per_iter,
    }

    other = [p for p in results if p != prec]
    if other:
        o = results[other[0]]
        fp64 = results[shci.PREC_FP64]
        mixed = results[shci.PREC_FP32]
        out["forms"] = {
            "mixed_fp32": {
                "description": "FP32 Hamiltonian values with FP64 accumulation in the Davidson iterations; "
                               "final tight Davidson on FP32 hi+lo values (FP64-accurate refinement)",
                "value": round(statistics.mean(mixed["iter_times"]), 6),
                "ms_per_step": round(1e3 * statistics.mean(mixed["step_times"]), 3),
                "final_energy": mixed["energies"][-1],
                "energy_error_vs_fp64_Ha": mixed["energies"][-1] - fp64["energies"][-1],
                "final_ndets": mixed["per_iter"][-1]["ndets"],
                "speedup_vs_fp64": round(statistics.mean(fp64["step_times"]) / statistics.mean(mixed["step_times"]),
                                         3),
                "sigma_GBps": round(mixed["nbytes"] / (mixed["ms"] * 1e-3) / 1e9, 1) if mixed["ms"] else None,
                "iterations": mixed["per_iter"],
            }
        }
        # both triangles stored (the FP64 default stores the lower one alone):
        # sigma on their tiled copy, kSigmaTiled
        F = run_form(shci, full, log, shci.PREC_FP64, fsteps, fwarm, tag="fp64 full storage", half_storage=False)
        fms = F["ms"] / F["calls"] if F["calls"] else None
        fgbs = F["nbytes"] / (F["ms"] * 1e-3) / 1e9 if F["ms"] else None
        out["forms"]["full_storage_fp64"] = {
            "description": "both triangles of H stored; sigma on their tiled copy (kSigmaTiled: the segmented "
                           "per-row scan, large segments gathering x from an LDS copy of their tile); sigma bytes "
                           "count both triangles",
            "value": round(statistics.mean(F["iter_times"]), 6),
            "ms_per_step": round(1e3 * statistics.mean(F["step_times"]), 3),
            "speedup_vs_fp64": round(statistics.mean(fp64["step_times"]) / statistics.mean(F["step_times"]), 3),
            "final_energy": F["energies"][-1],
            "energy_error_vs_fp64_Ha": F["energies"][-1] - fp64["energies"][-1],
            "final_ndets": F["per_iter"][-1]["ndets"],
            "sigma_avg_ms": round(fms, 4) if fms else None,
            "sigma_GBps": round(fgbs, 1) if fgbs else None,
            "sigma_frac_of_hbm_peak": round(fgbs / PEAK_HBM_GBS, 4) if fgbs else None,
            "sigma_traffic_bytes": traffic_bytes("fp64_full_storage"),
            "iterations": F["per_iter"],
        }
        # learned screening + mixed precision (the combined fast form)
        L = run_form(shci, full, log, shci.PREC_FP32, fsteps, fwarm, tag="learned+fp32", learned=True)
        recs, train = L["learned"]
        out["forms"]["learned_mixed_fp32"] = {
            "description": "learned candidate ranking (gated: after the first exact screen, and again after the "
                           "second logging iteration, the form turns itself off unless ranking the candidate "
                           "records at 2 ns each would cost less than the exact screen; when on, an MLP trained on "
                           "the first two iterations' screening values predicts log10(|c||v|) and candidates "
                           "predicted above log10(eps1) - 0.8 are tested exactly) with the FP32 form",
            "value": round(statistics.mean(L["iter_times"]), 6),
            "ms_per_step": round(1e3 * statistics.mean(L["step_times"]), 3),
            "speedup_vs_fp64": round(statistics.mean(fp64["step_times"]) / statistics.mean(L["step_times"]), 3),
            "final_energy": L["energies"][-1],
            "energy_error_vs_fp64_Ha": L["energies"][-1] - fp64["energies"][-1],
            "final_ndets": L["per_iter"][-1]["ndets"],
            "training": train,
            "ranked_iterations": recs,
            "iterations": L["per_iter"],
        }
        # batching + mixed precision on the same workload: block Davidson
        # (one root: 17-vector subspace, thick restart), hipGraph-replayed
        # iterations, FP32 values (half storage, as the FP64 path)
        B = run_form(shci, full, log, shci.PREC_FP32, fsteps, fwarm, tag="block+fp32", profile=False,
                     davidson_block=True, hip_graphs=not args.no_graphs)
        out["forms"]["batched_mixed_fp32"] = {
            "description": "block Davidson (thick restart, hipGraph-replayed iterations) with the FP32 form; its "
                           "Davidson vectors differ from the reference Davidson's within the 5e-5 residual "
                           "tolerance, so determinants at the eps1 threshold can be selected differently",
            "value": round(statistics.mean(B["iter_times"]), 6),
            "ms_per_step": round(1e3 * statistics.mean(B["step_times"]), 3),
            "speedup_vs_fp64": round(statistics.mean(fp64["step_times"]) / statistics.mean(B["step_times"]), 3),
            "final_energy": B["energies"][-1],
            "energy_error_vs_fp64_Ha": B["energies"][-1] - fp64["energies"][-1],
            "final_ndets": B["per_iter"][-1]["ndets"],
            "iterations": B["per_iter"],
        }
        # the three forms together: gated learned screening, block Davidson
        # with hipGraph replay, FP32 values
        LB = run_form(shci, full, log, shci.PREC_FP32, fsteps, fwarm, tag="learned+block+fp32", profile=False,
                      davidson_block=True, hip_graphs=not args.no_graphs, learned=True)
        out["forms"]["learned_batched_mixed_fp32"] = {
            "description": "gated learned screening + block Davidson (hipGraph-replayed) + FP32 values",
            "value": round(statistics.mean(LB["iter_times"]), 6),
            "ms_per_step": round(1e3 * statistics.mean(LB["step_times"]), 3),
            "speedup_vs_fp64": round(statistics.mean(fp64["step_times"]) / statistics.mean(LB["step_times"]), 3),
            "final_energy": LB["energies"][-1],
            "energy_error_vs_fp64_Ha": LB["energies"][-1] - fp64["energies"][-1],
            "final_ndets": LB["per_iter"][-1]["ndets"],
            "learned": LB["learned"][0],
            "iterations": LB["per_iter"],
        }
        out["forms"]["batched_block"] = batching_form(shci, args)

    if not args.no_forms:
        out["pt_deterministic"] = pt_measure(shci, full, sample, log, args)
        out["c0_small"] = c0_form(shci, args)
        out["c3_large"] = c3_form(shci, args)
        out["c4_learned_large"] = c4_form(shci, args)

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import OracleRun
        # host threads for the oracle's rank-parallel loops (rows of H,
        # screened determinants, sigma rows): the box's CPU share
        threads = host_threads()
        os.environ["ORACLE_THREADS"] = str(threads)
        cpu = []
        for _ in range(3):
            t0 = time.time()
            OracleRun(sample, os.path.join(work, "oracle.log"))
            cpu.append(time.time() - t0)
        t0 = time.time()
        shci.do_variational(sample, log, precision=prec)
        gpu_sample = time.time() - t0
        out["cpu_baseline"] = {
            "value": round(statistics.median(cpu), 3),
            "unit": "s per sample (first variational iteration, 27,685 determinants, + final tight Davidson)",
            "cores": threads,
            "kind": "port",
            "sample": "iteration 0 of the benchmark workload from HF at eps1 %.0e (27,685 determinants) and the "
                      "final tight Davidson, median of 3 runs; the oracle's rank-parallel loops on %d OpenMP "
                      "threads in place of MPI ranks; GPU time of the same sample: %.3f s"
                      % (EPS1, threads, gpu_sample),
            "gpu_same_sample_s": round(gpu_sample, 4),
            "host_nproc": os.cpu_count(),
            "final_space": cpu_final_space(shci, full, args),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
