"""Determinant utilities and an independent second-quantisation evaluator
used to pin the oracle (no Slater-Condon formulas here: operators are applied
to occupation lists with explicit fermionic signs)."""
import itertools

import numpy as np


def occ_to_det(occ, words):
    w = [0] * words
    for o in occ:
        o = int(o)
        w[o // 64] |= 1 << (o % 64)
    return np.array([x - (1 << 64) if x >= 1 << 63 else x for x in w], dtype=np.int64)


def det_to_occ(det):
    occ = []
    for w, x in enumerate(det):
        u = int(x) & ((1 << 64) - 1)
        for b in range(64):
            if (u >> b) & 1:
                occ.append(64 * w + b)
    return occ


def words_for(norb):
    return (2 * norb) // 64 + 1


def dice_string_to_occ(tokens):
    """'2 2 a b 0 ...' (one token per spatial orbital) -> spin-orbital list."""
    occ = []
    for i, t in enumerate(tokens):
        if t in ("a", "2"):
            occ.append(2 * i)
        if t in ("b", "2"):
            occ.append(2 * i + 1)
    return sorted(occ)


def read_golden_dets(path, norb):
    coeffs, occs = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                continue
            coeffs.append(float(tok[0]))
            occs.append(dice_string_to_occ(tok[1:1 + norb]))
    return np.array(coeffs), occs


def random_dets(norb, na, nb, n, rng, words=None):
    words = words or words_for(norb)
    out = set()
    res = []
    while len(res) < n:
        a = sorted(rng.choice(norb, na, replace=False))
        b = sorted(rng.choice(norb, nb, replace=False))
        occ = tuple(sorted([2 * x for x in a] + [2 * x + 1 for x in b]))
        if occ not in out:
            out.add(occ)
            res.append(occ_to_det(occ, words))
    return np.array(res, dtype=np.int64)


def excite(det_occ, norb, rng, nexc):
    """Random determinant reached from det_occ by nexc spin-conserving excitations."""
    occ = set(det_occ)
    for _ in range(nexc):
        i = rng.choice(sorted(occ))
        empt = [a for a in range(2 * norb) if a not in occ and (a % 2) == (i % 2)]
        if not empt:
            continue
        a = rng.choice(empt)
        occ.remove(i)
        occ.add(a)
    return sorted(occ)


class Integrals:
    """Spatial integrals read straight from an FCIDUMP (independent parser)."""

    def __init__(self, path):
        with open(path) as f:
            lines = f.read().splitlines()
        head = []
        k = 0
        while True:
            s = lines[k].strip()
            k += 1
            head.append(s)
            if s.upper().startswith("&END") or s == "/":
                break
        import re
        txt = " ".join(head)
        self.norb = int(re.search(r"NORB\s*=\s*(\d+)", txt, re.I).group(1))
        self.nelec = int(re.search(r"NELEC\s*=\s*(\d+)", txt, re.I).group(1))
        n = self.norb
        self.h = np.zeros((n, n))
        self.eri = np.zeros((n, n, n, n))
        self.core = 0.0
        for line in lines[k:]:
            t = line.replace(",", " ").split()
            if len(t) != 5:
                continue
            v = float(t[0])
            p, q, r, s = (int(x) for x in t[1:])
            if p == q == r == s == 0:
                self.core = v
            elif q == r == s == 0:
                continue
            elif r == s == 0:
                self.h[p - 1, q - 1] = self.h[q - 1, p - 1] = v
            else:
                p, q, r, s = p - 1, q - 1, r - 1, s - 1
                for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                                   (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                    self.eri[a, b, c, d] = v


def _annihilate(occ, q):
    if q not in occ:
        return None, 0
    i = occ.index(q)
    return occ[:i] + occ[i + 1:], (-1) ** i


def _create(occ, p):
    if p in occ:
        return None, 0
    i = sum(1 for x in occ if x < p)
    return occ[:i] + [p] + occ[i:], (-1) ** i


def apply_h(ints, ket_occ, include_core=False):
    """H|ket> as {tuple(occ): amplitude}; H = sum h a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q."""
    out = {}
    ket = sorted(ket_occ)
    nso = 2 * ints.norb

    def add(occ, v):
        key = tuple(occ)
        out[key] = out.get(key, 0.0) + v

    for q in ket:
        o1, s1 = _annihilate(ket, q)
        for p in range(q % 2, nso, 2):
            hv = ints.h[p // 2, q // 2]
            if hv == 0.0:
                continue
            o2, s2 = _create(o1, p)
            if o2 is None:
                continue
            add(o2, s1 * s2 * hv)
    for q in ket:
        o1, s1 = _annihilate(ket, q)
        for s in o1:
            o2, s2 = _annihilate(o1, s)
            for r in range(s % 2, nso, 2):
                o3, s3 = _create(o2, r)
                if o3 is None:
                    continue
                for p in range(q % 2, nso, 2):
                    v = ints.eri[p // 2, q // 2, r // 2, s // 2]
                    if v == 0.0:
                        continue
                    o4, s4 = _create(o3, p)
                    if o4 is None:
                        continue
                    add(o4, 0.5 * s1 * s2 * s3 * s4 * v)
    if include_core:
        add(ket, ints.core)
    return out


def fci_space(norb, na, nb):
    res = []
    for a in itertools.combinations(range(norb), na):
        for b in itertools.combinations(range(norb), nb):
            res.append(sorted([2 * x for x in a] + [2 * x + 1 for x in b]))
    return res


def read_dets_bin(path):
    """dets.bin of writeBestDeterminants (SHCI/SHCI.cpp:464-505): int count,
    int spatial orbitals, then per determinant a double coefficient (times the
    Dice -> alpha-beta parity) and one char per spatial orbital (0, a, b, 2).
    Returns (norb, [(coefficient, string)])."""
    import struct
    with open(path, "rb") as f:
        b = f.read()
    num, norb = struct.unpack("<ii", b[:8])
    if len(b) != 8 + num * (8 + norb):
        raise ValueError("dets.bin: %d bytes, expected %d" % (len(b), 8 + num * (8 + norb)))
    out, off = [], 8
    for _ in range(num):
        c = struct.unpack("<d", b[off:off + 8])[0]
        s = b[off + 8:off + 8 + norb].decode("ascii")
        if set(s) - set("0ab2"):
            raise ValueError("dets.bin: bad occupation string %r" % s)
        out.append((c, s))
        off += 8 + norb
    return norb, out


def occ_string(occ, norb):
    """Spin-orbital list -> the 0/a/b/2 string of dets.bin."""
    o = set(occ)
    return "".join("2" if (2 * i in o and 2 * i + 1 in o) else "a" if 2 * i in o else "b" if 2 * i + 1 in o
                   else "0" for i in range(norb))


def read_fcidump_arrays(path):
    """(norb, nelec, h1 norb x norb, eri8 8-fold packed, core energy, orbsym)
    from a FCIDUMP, in the in-memory layout of shci_create_from_integrals
    (lower-triangle pair index A = p(p+1)/2 + q for p >= q, element
    A(A+1)/2 + B for A >= B)."""
    import re
    with open(path) as f:
        text = f.read()
    head, body = re.split(r"&END|/\s*\n", text, maxsplit=1, flags=re.I)
    norb = int(re.search(r"NORB\s*=\s*(\d+)", head, re.I).group(1))
    nelec = int(re.search(r"NELEC\s*=\s*(\d+)", head, re.I).group(1))
    m = re.search(r"ORBSYM\s*=\s*([\d,\s]+)", head, re.I)
    orbsym = [int(x) for x in re.split(r"[,\s]+", m.group(1).strip()) if x][:norb] if m else [1] * norb
    npair = norb * (norb + 1) // 2
    h1 = np.zeros((norb, norb))
    eri8 = np.zeros(npair * (npair + 1) // 2)
    ecore = 0.0

    def pair(p, q):
        return p * (p + 1) // 2 + q if p >= q else q * (q + 1) // 2 + p

    for line in body.splitlines():
        t = line.split()
        if len(t) != 5:
            continue
        v = float(t[0])
        p, q, r, s = (int(x) for x in t[1:])
        if p == q == r == s == 0:
            ecore = v
        elif r == 0 and s == 0:
            if q:
                h1[p - 1, q - 1] = h1[q - 1, p - 1] = v
        else:
            A, B = pair(p - 1, q - 1), pair(r - 1, s - 1)
            if A < B:
                A, B = B, A
            eri8[A * (A + 1) // 2 + B] = v
    return norb, nelec, h1, eri8, ecore, orbsym
