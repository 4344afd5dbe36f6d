"""Integral input: FCIDUMP reading/writing and the seeded synthetic generator.

FCIDUMP follows the reference parser (models/chem/parse_integrals.h:98-170):
a 4-line &FCI header, then `value i j k l` lines (1-based; one-electron
terms `p q 0 0`; core energy `0 0 0 0`).  Indices are shifted to 0-based
(-1 = absent) and brought to canonical order with `align`
(utils/align.h:18-28).  A second numeric column (imaginary part) is
accepted and must be zero.

The synthetic systems follow the benchmark configurations: one-electron
h_pq ~ N(0, 0.1) symmetric with diagonal -2 + 0.1 p; two-electron
(pq|rs) ~ N(0, 0.05) drawn once per canonical 8-fold-symmetric quadruple;
"positive Coulomb diagonal 0.5" is read as adding +0.5 to every (pp|pp)
and (pp|qq) element (the Coulomb integrals J_pq), which keeps the
two-electron operator closer to positive semi-definite.  Optional distance
decay multiplies h_pq by exp(-|p-q|/d) and (pq|rs) by
exp(-(|p-q|+|r-s|+|p-r|)/d).
"""
from __future__ import annotations

import re

import numpy as np

from .terms import align


def read_fcidump(path_or_text, is_text=False):
    text = path_or_text if is_text else open(path_or_text).read()
    lines = text.splitlines()
    header = []
    body_start = 0
    for n, ln in enumerate(lines):
        header.append(ln)
        if re.search(r"&END|/\s*$", ln, re.IGNORECASE):
            body_start = n + 1
            break
    hdr = " ".join(header)
    m = re.search(r"NORB\s*=\s*(\d+)", hdr, re.IGNORECASE)
    norb = int(m.group(1)) if m else None
    m = re.search(r"NELEC\s*=\s*(\d+)", hdr, re.IGNORECASE)
    nelec = int(m.group(1)) if m else None
    m = re.search(r"MS2\s*=\s*(-?\d+)", hdr, re.IGNORECASE)
    ms2 = int(m.group(1)) if m else 0
    ints = []
    for ln in lines[body_start:]:
        f = ln.split()
        if not f:
            continue
        if len(f) == 6:
            val, im = float(f[0]), float(f[1])
            if im != 0.0:
                raise ValueError("complex integrals are not supported")
            idx = [int(x) for x in f[2:6]]
        elif len(f) == 5:
            val = float(f[0])
            idx = [int(x) for x in f[1:5]]
        else:
            raise ValueError(f"bad FCIDUMP line: {ln!r}")
        ints.append((tuple(idx), val))
    return dict(norb=norb, nelec=nelec, ms2=ms2, integrals=ints)


def write_fcidump(path, norb, nelec, ms2, ints):
    with open(path, "w") as f:
        f.write(f" &FCI NORB={norb},NELEC={nelec},MS2={ms2},\n")
        f.write("  ORBSYM=" + ",".join(["1"] * norb) + ",\n")
        f.write("  ISYM=1,\n")
        f.write(" &END\n")
        for (i, j, k, l), v in ints:
            f.write(f"{v: .17e} {i:4d} {j:4d} {k:4d} {l:4d}\n")


def to_aligned(ints, cutoff=1e-300):
    """1-based FCIDUMP tuples -> 0-based aligned (i,j,k,l) (parse_integrals.h:150-166)."""
    out = []
    for (i, j, k, l), v in ints:
        if abs(v) > cutoff:
            out.append((align(i - 1, j - 1, k - 1, l - 1), v))
    return out


def synthetic(L, seed, decay=None):
    """Seeded integrals for the benchmark systems, as 1-based FCIDUMP tuples."""
    rng = np.random.default_rng(seed)
    h = rng.normal(0.0, 0.1, size=(L, L))
    h = 0.5 * (h + h.T)
    for p in range(L):
        h[p, p] = -2.0 + 0.1 * p
    if decay:
        for p in range(L):
            for q in range(L):
                if p != q:
                    h[p, q] *= np.exp(-abs(p - q) / decay)
    ints = []
    quads = []
    for p in range(L):
        for q in range(p + 1):
            for r in range(L):
                for s in range(r + 1):
                    if (p, q) >= (r, s):
                        quads.append((p, q, r, s))
    vals = rng.normal(0.0, 0.05, size=len(quads))
    for (p, q, r, s), v in zip(quads, vals):
        if p == q and r == s:
            v += 0.5
        if decay:
            v *= np.exp(-(abs(p - q) + abs(r - s) + abs(p - r)) / decay)
        ints.append(((p + 1, q + 1, r + 1, s + 1), float(v)))
    for p in range(L):
        for q in range(p + 1):
            ints.append(((p + 1, q + 1, 0, 0), float(h[p, q])))
    ints.append(((0, 0, 0, 0), 0.0))
    return ints


def dense_tensors(L, ints):
    """(h[L,L], eri[L,L,L,L] chemist order, ecore) from 1-based tuples, 8-fold expanded."""
    h = np.zeros((L, L))
    g = np.zeros((L, L, L, L))
    ecore = 0.0
    for (i, j, k, l), v in ints:
        if i == 0 and j == 0 and k == 0 and l == 0:
            ecore += v
        elif k == 0 and l == 0:
            h[i - 1, j - 1] = v
            h[j - 1, i - 1] = v
        else:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for (a, b, c, d) in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                 (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                g[a, b, c, d] = v
    return h, g, ecore
