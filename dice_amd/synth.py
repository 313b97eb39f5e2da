"""Synthetic FCIDUMP / input.dat writers for the benchmark systems.

Molecular-like systems: h_pp = linspace(-2, 1, n); h_pq (p < q) ~ N(0, 0.05)
drawn row-major and symmetrised; Cholesky-like vectors L^x (x < X), upper
triangle with diagonal ~ N(0, 0.3/(1+x)), mirrored; (pq|rs) = sum_x L^x_pq
L^x_rs.  Random stream numpy.random.default_rng(1000 + c), draw order h then
L^0..L^{X-1}.  Files use the reference writer's format (lib/shcitools.c:165-234:
header, 8-fold unique ERIs, one-electron terms, core energy; "%20.12e").

Hubbard: periodic Lx x Ly lattice, h_ij = -t for nearest neighbours and
(ii|ii) = U, site basis; the reference determinant is the Neel state (alpha
on even (x+y) sites, beta on odd ones).
"""
import numpy as np

# name: (config index, norb, nalpha, nbeta, ncholesky)
MOLECULAR = {
    "c0": (0, 16, 8, 8, 64),
    "c1": (1, 24, 12, 12, 96),
    "c3": (3, 40, 20, 20, 160),
}


def molecular_integrals(c, n, nchol):
    rng = np.random.default_rng(1000 + c)
    h = np.zeros((n, n))
    off = rng.normal(0.0, 0.05, size=n * (n - 1) // 2)
    iu = np.triu_indices(n, 1)
    h[iu] = off
    h = h + h.T
    h[np.diag_indices(n)] = np.linspace(-2.0, 1.0, n)
    iu0 = np.triu_indices(n)
    Ls = np.zeros((nchol, n, n))
    for x in range(nchol):
        L = np.zeros((n, n))
        L[iu0] = rng.normal(0.0, 0.3 / (1.0 + x), size=len(iu0[0]))
        L = L + L.T - np.diag(np.diag(L))
        Ls[x] = L
    return h, Ls


def hubbard_integrals(lx, ly, t, u):
    n = lx * ly
    h = np.zeros((n, n))
    for x in range(lx):
        for y in range(ly):
            s = x * ly + y
            for dx, dy in ((1, 0), (0, 1)):
                q = ((x + dx) % lx) * ly + (y + dy) % ly
                h[s, q] = -t
                h[q, s] = -t
    return h, u


def _write_header(f, norb, nelec, ms2):
    f.write(" &FCI NORB=%d ,NELEC=%d ,MS2=%d,\n" % (norb, nelec, ms2))
    f.write("  ORBSYM=" + "".join("1," for _ in range(norb)) + "\n")
    f.write("  ISYM=1,\n &END\n")


def write_molecular_fcidump(path, c, n, na, nb, nchol, tol=1e-12):
    h, Ls = molecular_integrals(c, n, nchol)
    npair = n * (n + 1) // 2
    pi, pj = [], []
    for i in range(n):
        for j in range(i + 1):
            pi.append(i)
            pj.append(j)
    pi, pj = np.array(pi), np.array(pj)
    V = Ls[:, pi, pj]  # nchol x npair
    eri = V.T @ V      # (ij|kl) over pairs
    with open(path, "w") as f:
        _write_header(f, n, na + nb, na - nb)
        lines = []
        for ij in range(npair):
            i, j = pi[ij], pj[ij]
            for kl in range(ij + 1):
                v = eri[ij, kl]
                if abs(v) > tol:
                    lines.append("%20.12e    %d  %d  %d  %d\n" % (v, i + 1, j + 1, pi[kl] + 1, pj[kl] + 1))
        f.write("".join(lines))
        for i in range(n):
            for j in range(i + 1):
                if abs(h[i, j]) > tol:
                    f.write("%20.12e    %d  %d  %d  %d\n" % (h[i, j], i + 1, j + 1, 0, 0))
        f.write("%20.12f   %d  %d  %d  %d\n" % (0.0, 0, 0, 0, 0))


def write_hubbard_fcidump(path, lx, ly, t, u, na, nb):
    h, U = hubbard_integrals(lx, ly, t, u)
    n = lx * ly
    with open(path, "w") as f:
        _write_header(f, n, na + nb, na - nb)
        for i in range(n):
            f.write("%20.12e    %d  %d  %d  %d\n" % (U, i + 1, i + 1, i + 1, i + 1))
        for i in range(n):
            for j in range(i + 1):
                if abs(h[i, j]) > 1e-12:
                    f.write("%20.12e    %d  %d  %d  %d\n" % (h[i, j], i + 1, j + 1, 0, 0))
        f.write("%20.12f   %d  %d  %d  %d\n" % (0.0, 0, 0, 0, 0))


def hf_occupation(na, nb):
    """Aufbau occupation in spin-orbital indices: alpha 0,2,..., beta 1,3,..."""
    return [2 * i for i in range(na)] + [2 * i + 1 for i in range(nb)]


def neel_occupation(lx, ly):
    occ = []
    for x in range(lx):
        for y in range(ly):
            s = x * ly + y
            occ.append(2 * s if (x + y) % 2 == 0 else 2 * s + 1)
    return sorted(occ)


def write_input(path, fcidump, occ, schedule, maxiter, nroots=1, seed=1, davidson_tol=5e-5, dE=1e-8,
                extra=()):
    with open(path, "w") as f:
        f.write("nocc %d\n" % len(occ))
        f.write(" ".join(str(o) for o in occ) + "\n")
        f.write("end\n")
        f.write("orbitals %s\n" % fcidump)
        f.write("nroots %d\n" % nroots)
        f.write("schedule\n")
        for it, eps in schedule:
            f.write("%d %g\n" % (it, eps))
        f.write("end\n")
        f.write("davidsonTol %g\ndE %g\nmaxiter %d\nseed %d\n" % (davidson_tol, dE, maxiter, seed))
        f.write("noio\nsampleN 200\n")
        if not any(l.lower().startswith("nptiter") for l in extra):
            f.write("nPTiter 0\n")
        for line in extra:
            f.write(line + "\n")


# the benchmark workloads: (fcidump writer args, schedule, maxiter, nroots)
def make_config(name, directory):
    """Write FCIDUMP + input.dat for a named workload into `directory`; returns the input path."""
    import os
    directory = os.path.abspath(directory)
    os.makedirs(directory, exist_ok=True)
    fcidump = os.path.join(directory, "FCIDUMP")
    inp = os.path.join(directory, "input.dat")
    if name in MOLECULAR:
        c, n, na, nb, nchol = MOLECULAR[name]
        if not os.path.exists(fcidump):
            write_molecular_fcidump(fcidump, c, n, na, nb, nchol)
        sched = {"c0": ([(0, 1e-3), (3, 1e-4)], 9, 1),
                 "c1": ([(0, 1e-4), (3, 5e-5)], 9, 1),
                 "c3": ([(0, 1e-4), (3, 5e-5), (6, 2e-5)], 12, 2)}[name]
        write_input(inp, fcidump, hf_occupation(na, nb), sched[0], sched[1], nroots=sched[2])
    elif name == "c2":
        if not os.path.exists(fcidump):
            write_hubbard_fcidump(fcidump, 6, 6, 1.0, 4.0, 18, 18)
        write_input(inp, fcidump, neel_occupation(6, 6), [(0, 2e-4)], 6, nroots=4)
    else:
        raise KeyError(name)
    return inp
