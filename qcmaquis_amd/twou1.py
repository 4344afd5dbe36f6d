"""Spin-orbital (U(1) x U(1)) quantum-chemistry Hamiltonian terms.

    H = sum_pq h_pq sum_s c+_ps c_qs
      + 1/2 sum_pqrs (pq|rs) sum_st c+_ps c+_rt c_st c_qs  +  E_core

(chemist's notation (pq|rs)), the Hamiltonian of
models/chem/2u1/model.hpp:174-447, generated here directly from the second-
quantized form instead of the reference's case analysis: every operator
string is brought to site order (the sign of the fermionic permutation),
operators on one site are multiplied into one local operator, and the
Jordan-Wigner string is carried as in the reference's TermMaker: a local
operator is followed by the fill (parity) operator when an odd number of
fermionic operators acts on sites to its right, and the MPO maker inserts
fill operators on the sites between.  The one-site operators are the
reference's (models/chem/2u1/model.hpp:25-171: down-spin operators carry the
on-site fill, so up precedes down within a site).  Identical operator
strings are merged before the MPO is built.
"""
from __future__ import annotations

import numpy as np

from .integrals import dense_tensors
from .operators import SiteOp, TwoU1Ops, _block_gemm, proportional
from .terms import Term

CU, CD, AU, AD = 0, 1, 2, 3  # create up / down, destroy up / down


def _is_zero(op: SiteOp) -> bool:
    return all(not np.any(np.abs(m) > 1e-14) for m in op.blocks.values())


class TwoU1Terms:
    """Builds site-ordered MPO terms from fermionic operator strings."""

    def __init__(self, ops: TwoU1Ops):
        self.ops = ops
        self.table = ops.table
        self.base = {CU: ops.create_up, CD: ops.create_down, AU: ops.destroy_up, AD: ops.destroy_down}
        self.cache = {}

    def _register(self, op: SiteOp, fermionic: bool):
        """(tag, scale) with scale * table[tag] == op.  The MPO of these terms is
        built without hermitian-partner bonds, so no conjugate tags are needed."""
        tb = self.table
        for t, o in enumerate(tb.ops):
            if tb.fermionic[t] == fermionic and proportional(o, op):
                k = o.sorted_keys()[0]
                fo = o.blocks[k].ravel()
                i = np.flatnonzero(np.abs(fo) > 1e-50)[0]
                return t, float(op.blocks[k].ravel()[i] / fo[i])
        return tb.register(op, fermionic), 1.0

    def local(self, kinds, fill):
        """Tag and scale of the site operator O_1 O_2 ... O_n (O_1 leftmost, acting
        last), right-multiplied by the fill when `fill`; None if it vanishes.
        Block matrices hold <out|O|in> at (in, out), so the product reverses."""
        key = (tuple(kinds), fill)
        if key in self.cache:
            return self.cache[key]
        tb = self.table
        m = tb[self.base[kinds[-1]]]
        for k in reversed(kinds[:-1]):
            m = _block_gemm(m, tb[self.base[k]])
        if fill:
            m = _block_gemm(tb[self.ops.fill], m)
        m.remove_empty_blocks()
        res = None
        if m.blocks and not _is_zero(m):
            res = self._register(m, len(kinds) % 2 == 1)
        self.cache[key] = res
        return res

    def term(self, coeff, string):
        """Term for coeff * prod of (site, kind) in operator order, or None."""
        order = sorted(range(len(string)), key=lambda i: string[i][0])
        # parity of the permutation that sorts the fermionic string by site (stable)
        perm = list(order)
        sign = 1
        seen = [False] * len(perm)
        for i in range(len(perm)):
            if seen[i]:
                continue
            j, n = i, 0
            while not seen[j]:
                seen[j] = True
                j = perm[j]
                n += 1
            if n % 2 == 0:
                sign = -sign
        srt = [string[i] for i in order]
        sites = sorted(set(s for s, _ in srt))
        ops = []
        c = coeff * sign
        for site in sites:
            kinds = [k for s, k in srt if s == site]
            right = sum(1 for s, _ in srt if s > site)
            loc = self.local(kinds, right % 2 == 1)
            if loc is None:
                return None
            tag, sc = loc
            c *= sc
            ops.append((site, tag))
        return Term(c, True, ops)


def twou1_terms(integrals, L, ops: TwoU1Ops, cutoff=1e-14, progress=None):
    """Terms of the spin-orbital Hamiltonian from 1-based FCIDUMP tuples.
    progress(msg), when given, is called after each first orbital index of the
    two-body terms."""
    h, g, ecore = dense_tensors(L, integrals)
    tb = TwoU1Terms(ops)
    acc = {}

    def add(coeff, string):
        t = tb.term(coeff, string)
        if t is None:
            return
        key = tuple(t.ops)
        acc[key] = acc.get(key, 0.0) + t.coeff

    for p in range(L):
        for q in range(L):
            v = h[p, q]
            if abs(v) <= cutoff:
                continue
            add(v, [(p, CU), (q, AU)])
            add(v, [(p, CD), (q, AD)])
    spins = ((CU, AU), (CD, AD))
    for p in range(L):
        if progress and p:
            progress(f"two-body terms: orbital {p}/{L}, {len(acc)} strings")
        for q in range(L):
            for r in range(L):
                for s in range(L):
                    v = g[p, q, r, s]
                    if abs(v) <= cutoff:
                        continue
                    for cs, as_ in spins:
                        for ct, at in spins:
                            if p == r and cs == ct:
                                continue
                            if q == s and cs == ct:
                                continue
                            add(0.5 * v, [(p, cs), (r, ct), (s, at), (q, as_)])
    terms = [Term(c, True, list(k)) for k, c in acc.items() if abs(c) > cutoff]
    if ecore != 0.0:
        terms.append(Term(ecore, False, [(0, ops.ident)]))
    return terms
