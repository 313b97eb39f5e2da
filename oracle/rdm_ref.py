"""CPU oracle of the spin-summed spatial 1- and 2-RDMs of a selected-CI
wavefunction (TEST INFRASTRUCTURE ONLY: the tests load it as the checker;
the product computes RDMs on the GPU).

It follows the quantities of SHCI/SHCIrdm.cpp (EvaluateOneRDM :996-1100,
EvaluateRDM :863-995, the text layout of save1RDM / saveRDM :354-400,
:477-526) but not their algorithm: every element is a second-quantised
expectation value, with the operators applied one by one to occupation
bitstrings and the fermionic signs counted explicitly (two ways: overlaps
of the states the annihilators leave, and a brute-force application of all
four operators to each determinant).

    gamma[P, R]       = sum_s   <Psi| a+_{P s} a_{R s} |Psi>
    Gamma[P, Q, R, S] = sum_s,t <Psi| a+_{P s} a+_{Q t} a_{S t} a_{R s} |Psi>

(the reference's s2RDM(P*n + Q, R*n + S), written to spatialRDM.i.i.txt as
"P Q R S value"), and the spin-orbital forms behind them.  Pure Python
loops: for spaces of up to a few thousand determinants.
"""
import numpy as np


def _as_int(words):
    v = 0
    for k, w in enumerate(words):
        v |= (int(w) & ((1 << 64) - 1)) << (64 * k)
    return v


def _annihilate(d, k):
    if not (d >> k) & 1:
        return None, 0
    sign = -1 if bin(d & ((1 << k) - 1)).count("1") & 1 else 1
    return d ^ (1 << k), sign


def _create(d, k):
    if (d >> k) & 1:
        return None, 0
    sign = -1 if bin(d & ((1 << k) - 1)).count("1") & 1 else 1
    return d | (1 << k), sign


def _coefficients(bra, ket):
    cb = np.asarray(bra, dtype=np.float64)
    return cb, (cb if ket is None else np.asarray(ket, dtype=np.float64))


def spin_one_rdm(dets, bra, norb, ket=None):
    """s1[p, r] = <bra| a+_p a_r |ket> = <a_p bra | a_r ket>, from the
    (N-1)-electron states the annihilators leave."""
    ints = [_as_int(d) for d in np.atleast_2d(dets)]
    cb, ck = _coefficients(bra, ket)
    nso = 2 * norb
    one = {}
    for k, d in enumerate(ints):
        for r in range(nso):
            if (d >> r) & 1:
                d1, s_1 = _annihilate(d, r)
                one.setdefault(d1, []).append((k, r, s_1))
    s1 = np.zeros((nso, nso))
    for lst in one.values():
        for i, p, si in lst:
            for j, r, sj in lst:
                s1[p, r] += cb[i] * ck[j] * si * sj
    return s1


def spin_rdms(dets, bra, norb, ket=None):
    """Spin-orbital (s1[p, r], G[p, q, r, s]) of <bra| ... |ket>:
    s1[p, r] = <a+_p a_r>, G[p, q, r, s] = <a+_p a+_q a_s a_r>, from the
    overlaps of the (N-1)- and (N-2)-electron states that the annihilators
    leave: <I| a+_p a_r |J> = <a_p I | a_r J> and
    <I| a+_p a+_q a_s a_r |J> = <a_q a_p I | a_s a_r J>."""
    ints = [_as_int(d) for d in np.atleast_2d(dets)]
    cb, ck = _coefficients(bra, ket)
    nso = 2 * norb
    two = {}
    for k, d in enumerate(ints):
        occ = [o for o in range(nso) if (d >> o) & 1]
        for r in occ:
            d1, s_1 = _annihilate(d, r)
            for s in occ:
                if s != r:
                    d2, s_2 = _annihilate(d1, s)
                    two.setdefault(d2, []).append((k, r, s, s_1 * s_2))
    G = np.zeros((nso,) * 4)
    for lst in two.values():
        for i, p, q, si in lst:
            for j, r, s, sj in lst:
                G[p, q, r, s] += cb[i] * ck[j] * si * sj
    return spin_one_rdm(dets, bra, norb, ket), G


def spin_rdms_brute(dets, bra, norb, ket=None):
    """spin_rdms by applying a+_p a+_q a_s a_r to every ket determinant and
    looking the result up among the bra determinants (for small spaces)."""
    ints = [_as_int(d) for d in np.atleast_2d(dets)]
    index = {d: k for k, d in enumerate(ints)}
    cb, ck = _coefficients(bra, ket)
    nso = 2 * norb
    s1 = np.zeros((nso, nso))
    G = np.zeros((nso,) * 4)
    for j, dj in enumerate(ints):
        cj = ck[j]
        if cj == 0.0:
            continue
        occ = [k for k in range(nso) if (dj >> k) & 1]
        for r in occ:  # a+_p a_r
            d1, s_1 = _annihilate(dj, r)
            for p in range(nso):
                d2, s_2 = _create(d1, p)
                if d2 is None or d2 not in index:
                    continue
                s1[p, r] += cb[index[d2]] * cj * s_1 * s_2
        for r in occ:  # a+_p a+_q a_s a_r
            d1, s_1 = _annihilate(dj, r)
            for s in occ:
                if s == r:
                    continue
                d2, s_2 = _annihilate(d1, s)
                for q in range(nso):
                    d3, s_3 = _create(d2, q)
                    if d3 is None:
                        continue
                    for p in range(nso):
                        d4, s_4 = _create(d3, p)
                        if d4 is None or d4 not in index:
                            continue
                        G[p, q, r, s] += cb[index[d4]] * cj * s_1 * s_2 * s_3 * s_4
    return s1, G


def spatial_from_spin(s1, G):
    """Spin sums: gamma[P, R] = sum_s s1[Ps, Rs], Gamma[P, Q, R, S] =
    sum_st G[Ps, Qt, Rs, St] (spin-orbital 2P + s)."""
    nso = s1.shape[0]
    n = nso // 2
    g1 = s1[0::2, 0::2] + s1[1::2, 1::2]
    g2 = np.zeros((n,) * 4)
    for a in (0, 1):
        for b in (0, 1):
            g2 += G[a::2, b::2, a::2, b::2]
    return g1, g2


def spin_pair_matrix(G):
    """The pair-indexed spin 2-RDM D[pair(p,q), pair(r,s)] = G[p,q,r,s] for
    p > q, r > s, pair(p,q) = p(p+1)/2 + q (zero elsewhere)."""
    nso = G.shape[0]
    npair = nso * (nso + 1) // 2
    p, q = np.tril_indices(nso, -1)
    idx = p * (p + 1) // 2 + q
    D = np.zeros((npair, npair))
    D[np.ix_(idx, idx)] = G[p[:, None], q[:, None], p[None, :], q[None, :]]
    return D


def rdms(dets, ci, norb, ket=None):
    """(gamma n x n, Gamma n x n x n x n) of sum_I ci[I] |dets[I]> (bra) and
    ket (default: the same state)."""
    return spatial_from_spin(*spin_rdms(dets, ci, norb, ket))


def energy(g1, g2, h1, eri, core):
    """E = core + sum h_PQ gamma_PQ + 1/2 sum Gamma[P,Q,R,S] (PR|QS), with
    gamma recovered from Gamma as the reference's ComputeEnergyFromSpatialRDM
    does when only Gamma is given (here the given gamma is used)."""
    return core + float(np.einsum("pq,pq->", h1, g1)) + 0.5 * float(np.einsum("pqrs,prqs->", g2, eri))


def read_dice_1rdm(path):
    with open(path) as f:
        lines = f.read().split("\n")
    n = int(lines[0].split()[0])
    g = np.zeros((n, n))
    for line in lines[1:]:
        t = line.split()
        if t:
            g[int(t[0]), int(t[1])] = float(t[2])
    return g


def read_dice_2rdm(path):
    with open(path) as f:
        lines = f.read().split("\n")
    n = int(lines[0].split()[0])
    g = np.zeros((n,) * 4)
    for line in lines[1:]:
        t = line.split()
        if t:
            g[int(t[0]), int(t[1]), int(t[2]), int(t[3])] = float(t[4])
    return g
