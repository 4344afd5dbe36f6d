"""MPO construction and the MPO tensor data model.

`TaggedMPOMaker` restates models/generate_mpo/tagged_mpo_maker_optim.hpp:
terms are inserted as chains of (left key, right key) -> (tag, scale)
entries per site, with fork/merge keys sharing prefixes and suffixes,
fillings between operators, and hermitian-partner detection on the right
keys.  `MPOTensor` holds what mp_tensors/mpotensor.h:24-107 holds: a
column-compressed map (b1, b2) -> [(tag, scale)], bond spins, and the
hermitian partner info of mpotensor_detail.h:118-168.
`make_twosite_mpo` restates mp_tensors/ts_ops.h:36-126 with the SU(2)
Kronecker product of block_matrix/site_operator_algorithms.h:209-318.
"""
from __future__ import annotations

import math
from collections import defaultdict

import numpy as np

from .blocks import ProductBasis, fuse
from .operators import OpTable, SiteOp
from .wigner import mod_coupling, sixj

TRIVIAL_LEFT, BULK, BULK_NO_MERGE, TRIVIAL_RIGHT = 0, 1, 2, 3
KEY_TL = (TRIVIAL_LEFT, (), 0)
KEY_TR = (TRIVIAL_RIGHT, (), 0)


def _key(pos_op):
    return (BULK, tuple(pos_op), 0)


class MPOTensor:
    """One site (or two-site) MPO tensor."""

    def __init__(self, row_dim, col_dim, prempo, table: OpTable, left_herm, right_herm,
                 left_phase, right_phase, left_spins, right_spins):
        self.row_dim = row_dim
        self.col_dim = col_dim
        self.table = table
        # CSC order: sort by (b2, b1); within an element later tags are prepended
        # (mpotensor.hpp:18-34).
        ent = defaultdict(list)
        for (b1, b2, tag, scale) in sorted(prempo, key=lambda x: (x[1], x[0])):
            ent[(b1, b2)].insert(0, (tag, scale))
        self.entries = dict(ent)
        self.cols = [[] for _ in range(col_dim)]
        self.rows = [[] for _ in range(row_dim)]
        for (b1, b2) in sorted(self.entries.keys(), key=lambda x: (x[1], x[0])):
            self.cols[b2].append(b1)
            self.rows[b1].append(b2)
        for r in self.rows:
            r.sort()
        self.row_nnz = [len(r) for r in self.rows]
        self.col_nnz = [len(c) for c in self.cols]
        self.num_one_rows = sum(1 for x in self.row_nnz if x == 1)
        self.num_one_cols = sum(1 for x in self.col_nnz if x == 1)
        self.left_herm = list(left_herm)
        self.right_herm = list(right_herm)
        self.left_phase = list(left_phase)
        self.right_phase = list(right_phase)
        self.left_spins = list(left_spins)
        self.right_spins = list(right_spins)
        assert len(self.left_herm) == row_dim and len(self.right_herm) == col_dim
        assert len(self.left_spins) == row_dim and len(self.right_spins) == col_dim

    def has(self, b1, b2):
        return (b1, b2) in self.entries

    def at(self, b1, b2):
        return self.entries[(b1, b2)]

    def left_skip(self, b1):
        return self.left_herm[b1] < b1

    def right_skip(self, b2):
        return self.right_herm[b2] < b2

    def nnz(self):
        return len(self.entries)

    def used_tags(self):
        s = set()
        for v in self.entries.values():
            for t, _ in v:
                s.add(t)
        return sorted(s)


class TaggedMPOMaker:
    def __init__(self, L, table: OpTable, identities, fillings, identities_full=None, su2=True, herm=True):
        self.L = L
        self.herm = herm  # detect hermitian-partner MPO bonds (needs exact conjugate tags)
        self.table = table
        self.ident = identities
        self.ident_full = identities_full if identities_full is not None else identities
        self.fill = fillings
        self.su2 = su2
        self.prempo = [dict() for _ in range(L)]  # (k1,k2) -> list of (tag, scale) in insertion order
        self.order = [list() for _ in range(L)]   # insertion sequence of keys for stable multimap order
        self.site_terms = {}
        self.leftmost_right = L
        self.rightmost_left = 0
        self.core_energy = 0.0
        self.finalized = False

    # -- insertion helpers --------------------------------------------------
    def _insert(self, p, kk, val, detach):
        d = self.prempo[p]
        if detach or kk not in d:
            if kk not in d:
                d[kk] = []
                self.order[p].append(kk)
            d[kk].append(val)
        return kk[1]

    def _insert_filling(self, i, j, k, trivial_fill, spin_gt_one):
        for p in range(i, j):
            use_ident = self.ident_full if spin_gt_one else self.ident
            op = use_ident if trivial_fill else self.fill
            d = self.prempo[p]
            kk = (k, k)
            if kk not in d:
                d[kk] = [(op, 1.0)]
                self.order[p].append(kk)
            elif d[kk][0] != (op, 1.0):
                raise RuntimeError(f"Pre-existing term at site {p}")

    def add_term(self, term):
        ops = sorted(term.ops, key=lambda x: x[0])
        n = len(ops)
        if n == 1:
            self._add_1term(term.coeff, ops)
        elif n == 2:
            self._add_2term(term.coeff, ops)
        else:
            self._add_generic_term(term.coeff, ops)
        self.leftmost_right = min(self.leftmost_right, ops[-1][0])
        self.rightmost_left = max(self.rightmost_left, ops[0][0])

    def _add_1term(self, coeff, ops):
        p, tag = ops[0]
        if tag == self.ident:
            self.core_energy += coeff
        else:
            op = self.table[tag].scaled(coeff)
            if p in self.site_terms:
                self.site_terms[p].iadd(op)
            else:
                self.site_terms[p] = op

    def _add_2term(self, coeff, ops):
        tb = self.table
        spin = 0
        nferm = sum(1 for _, t in ops if tb.fermionic[t])
        spin += tb[ops[0][1]].action
        k2 = _key([ops[1]])
        k1 = self._insert(ops[0][0], (KEY_TL, k2), (ops[0][1], coeff), True)
        if tb.fermionic[ops[0][1]]:
            nferm -= 1
        trivial_fill = (nferm % 2 == 0)
        self._insert_filling(ops[0][0] + 1, ops[1][0], k1, trivial_fill, spin > 1)
        spin += tb[ops[1][1]].action
        self._insert(ops[1][0], (k1, KEY_TR), (ops[1][1], 1.0), False)
        assert spin == 0

    def _add_generic_term(self, coeff, ops):
        tb = self.table
        nops = len(ops)
        nferm = sum(1 for _, t in ops if tb.fermionic[t])
        spin = 0
        thresh = nops // 2
        k1 = KEY_TL
        ops_left = []
        for i in range(thresh):
            spin += tb[ops[i][1]].action
            ops_left.append(ops[i])
            k1 = self._insert(ops[i][0], (k1, _key(ops_left)), (ops[i][1], 1.0), False)
            if tb.fermionic[ops[i][1]]:
                nferm -= 1
            self._insert_filling(ops[i][0] + 1, ops[i + 1][0], k1, nferm % 2 == 0, spin > 1)
        k2 = _key(ops[thresh + 1:])
        spin += tb[ops[thresh][1]].action
        k1 = self._insert(ops[thresh][0], (k1, k2), (ops[thresh][1], coeff), True)
        if tb.fermionic[ops[thresh][1]]:
            nferm -= 1
        self._insert_filling(ops[thresh][0] + 1, ops[thresh + 1][0], k1, nferm % 2 == 0, spin > 1)
        for i in range(thresh + 1, nops):
            k2 = KEY_TR if i == nops - 1 else _key(ops[i + 1:])
            spin += tb[ops[i][1]].action
            k1 = self._insert(ops[i][0], (k1, k2), (ops[i][1], 1.0), False)
            if tb.fermionic[ops[i][1]]:
                nferm -= 1
            if i != nops - 1:
                self._insert_filling(ops[i][0] + 1, ops[i + 1][0], k1, nferm % 2 == 0, spin > 1)
        assert spin == 0

    def _finalize(self):
        kk = (KEY_TL, KEY_TR)
        for p in sorted(self.site_terms):
            tag = self.table.register(self.site_terms[p], False)
            d = self.prempo[p]
            if kk in d:
                raise RuntimeError("another site term already existing!")
            d[kk] = [(tag, 1.0)]
            self.order[p].append(kk)
        for p in range(0, self.rightmost_left):
            k = (KEY_TL, KEY_TL)
            self.prempo[p].setdefault(k, [])
            if not self.prempo[p][k]:
                self.order[p].append(k)
            self.prempo[p][k].append((self.ident, 1.0))
        for p in range(self.leftmost_right + 1, self.L):
            k = (KEY_TR, KEY_TR)
            self.prempo[p].setdefault(k, [])
            if not self.prempo[p][k]:
                self.order[p].append(k)
            self.prempo[p][k].append((self.ident, 1.0))
        self.finalized = True

    def _np(self, c):
        return c[0] if self.su2 else c[0] + c[1]

    def _conjugate_key(self, k, p):
        tb = self.table
        kind, pos_op, off = k
        conj = (kind, tuple((q, tb.herm_conj(t)) for q, t in pos_op), off)
        phase = (1, 1)
        if len(pos_op) == 1:
            op1 = tb[pos_op[0][1]]
            if p < pos_op[0][0]:
                cin, cout = op1.first_block()
                pdiff = self._np(cin) - self._np(cout)
                if pdiff == 1:
                    phase = (1, -1)
                elif pdiff == -1:
                    phase = (-1, 1)
            else:
                if op1.twoS == 1:
                    phase = (-1, 1)
        if len(pos_op) == 2:
            op1 = tb[pos_op[0][1]]
            op2 = tb[pos_op[1][1]]
            c1i, c1o = op1.first_block()
            c2i, c2o = op2.first_block()
            d1 = self._np(c1i) - self._np(c1o)
            d2 = self._np(c2i) - self._np(c2o)
            if op1.twoS == 1 and op2.twoS == 1 and op2.action == -1 and d1 == -d2:
                phase = (-1, -1)
            if op1.twoS == 1 and op2.twoS == 1 and op2.action == 1 and d1 == d2:
                phase = (-1, -1)
        return conj, phase

    def create_mpo(self):
        if not self.finalized:
            self._finalize()
        tb = self.table
        L = self.L
        mpo = []
        left = {KEY_TL: 0}
        left_spins = [0]
        left_herm = [0]
        left_phase = [1]
        for p in range(L):
            # multimap iteration order: sorted by (k2, k1), insertion order within a key
            keys = sorted(self.prempo[p].keys(), key=lambda kk: (kk[1], kk[0]))
            right = {}
            r = 2
            pre = []
            herm_pairs = {}
            herm_phases = {}
            for kk in keys:
                k1, k2 = kk
                if k1 not in left:
                    raise RuntimeError("k1 not found!")
                ll = left[k1]
                if k2 not in right:
                    if k2 == KEY_TL:
                        right[k2] = 0
                    elif k2 == KEY_TR:
                        right[k2] = 1
                    else:
                        right[k2] = r
                        r += 1
                rr = 0 if p == L - 1 else right[k2]
                for (tag, scale) in self.prempo[p][kk]:
                    pre.append((ll, rr, tag, scale))
                ck2, phase = self._conjugate_key(k2, p) if self.herm else (k2, None)
                if ck2 != k2:
                    herm_pairs[k2] = ck2
                    herm_phases[k2] = phase
            ldim = max(x[0] for x in pre) + 1
            rdim = max(x[1] for x in pre) + 1
            right_spins = [0] * rdim
            for (ll, rr, tag, scale) in pre:
                out = left_spins[ll] + tb[tag].action
                assert right_spins[rr] == 0 or right_spins[rr] == out
                right_spins[rr] = out
            right_herm = list(range(rdim))
            right_phase = [1] * rdim
            for hk in sorted(herm_pairs.keys()):
                romeo = right[hk]
                if herm_pairs[hk] not in right:
                    right[herm_pairs[hk]] = 0
                julia = right[herm_pairs[hk]]
                if romeo < julia:
                    right_herm[romeo], right_herm[julia] = right_herm[julia], right_herm[romeo]
                    right_phase[romeo] = herm_phases[hk][0]
                    right_phase[julia] = herm_phases[hk][1]
            rows = 1 if p == 0 else ldim
            cols = 1 if p == L - 1 else rdim
            lh = left_herm if len(left_herm) == rows else list(range(rows))
            lp = left_phase if len(left_phase) == rows else [1] * rows
            rh = right_herm if len(right_herm) == cols else list(range(cols))
            rp = right_phase if len(right_phase) == cols else [1] * cols
            ls = (left_spins + [0] * rows)[:rows]
            rs = (right_spins + [0] * cols)[:cols]
            mpo.append(MPOTensor(rows, cols, pre, tb, lh, rh, lp, rp, ls, rs))
            left = right
            left_spins = right_spins
            left_herm = right_herm
            left_phase = right_phase
        return mpo, self.core_energy


# ---------------------------------------------------------------------------
# two-site MPO
# ---------------------------------------------------------------------------

def _allowed_spins(left, right, k1, k2):
    out = []
    for s in range(abs(k1 - k2), abs(k1 + k2) + 1, 2):
        if right >= abs(s - left) and right <= abs(s + left):
            out.append(s)
    return out


def _product_spin(a, b):
    sa, sb = a[1], b[1]
    if sa == -1 and sb == 1:
        return 2
    return abs(sa + sb)


def op_kron_su2(phys_a, phys_b, Ao: SiteOp, Bo: SiteOp, lspin, mspin, rspin, target):
    """SU(2) Kronecker product of two site operators (site_operator_algorithms.h:209-318)."""
    pb = ProductBasis(phys_a, phys_b)
    A = Ao.copy()
    B = Bo.copy()
    if A.twoS > 0 and B.twoS == 0:
        cb, cc = phys_b[1][0], phys_b[2][0]
        if (cb, cc) not in B.blocks:
            B.blocks[(cb, cc)] = np.ones((1, 1))
            B.blocks[(cc, cb)] = np.ones((1, 1))
    if A.twoS == 0 and B.twoS > 0:
        cb, cc = phys_a[1][0], phys_a[2][0]
        if (cb, cc) not in A.blocks:
            A.blocks[(cb, cc)] = np.ones((1, 1))
            A.blocks[(cc, cb)] = np.ones((1, 1))
    k1, k2 = A.twoS, B.twoS
    j, jpp, jp = lspin, mspin, rspin
    k = target
    blocks = {}
    basis_spins = {}
    for (inA, outA), a in sorted(A.blocks.items(), key=lambda x: x[0], reverse=True):
        for (inB, outB), b in sorted(B.blocks.items(), key=lambda x: x[0], reverse=True):
            new_left = fuse(inA, inB)
            new_right = fuse(outA, outB)
            tmp = np.zeros((pb.size[new_left], pb.size[new_right]))
            yo = pb(inA, inB)
            xo = pb(outA, outB)
            la, ra = a.shape
            lb, rb = b.shape
            for l1 in range(la):
                for r1 in range(ra):
                    for l2 in range(lb):
                        for r2 in range(rb):
                            tmp[yo + l1 * lb + l2, xo + r1 * rb + r2] = b[l2, r2] * a[l1, r1]
            j1, j2, J = abs(inA[1]), abs(inB[1]), _product_spin(inA, inB)
            j1p, j2p, Jp = abs(outA[1]), abs(outB[1]), _product_spin(outA, outB)
            tmp *= mod_coupling(j1, j2, J, k1, k2, k, j1p, j2p, Jp)
            key = (new_left, new_right)
            if key in blocks:
                blocks[key] = blocks[key] + tmp
            else:
                blocks[key] = tmp
            rs, cs = basis_spins.get(key, ([], []))
            rs = list(rs) + [0] * max(0, tmp.shape[0] - len(rs))
            cs = list(cs) + [0] * max(0, tmp.shape[1] - len(cs))
            rs[yo] = J
            cs[xo] = Jp
            basis_spins[key] = (rs, cs)
    c = math.sqrt((jpp + 1) * (k + 1)) * sixj(j, jp, k, k2, k1, jpp)
    if ((j + jp + k1 + k2) // 2) % 2:
        c = -c
    for key in blocks:
        blocks[key] = blocks[key] * c
    return SiteOp(blocks, k, jp - j, basis_spins)


def op_kron_abelian(phys_a, phys_b, A: SiteOp, B: SiteOp):
    pb = ProductBasis(phys_a, phys_b)
    blocks = {}
    for (inA, outA), a in A.blocks.items():
        for (inB, outB), b in B.blocks.items():
            new_left = fuse(inA, inB)
            new_right = fuse(outA, outB)
            tmp = np.zeros((pb.size[new_left], pb.size[new_right]))
            yo = pb(inA, inB)
            xo = pb(outA, outB)
            la, ra = a.shape
            lb, rb = b.shape
            for l1 in range(la):
                for r1 in range(ra):
                    for l2 in range(lb):
                        for r2 in range(rb):
                            tmp[yo + l1 * lb + l2, xo + r1 * rb + r2] = b[l2, r2] * a[l1, r1]
            key = (new_left, new_right)
            blocks[key] = blocks[key] + tmp if key in blocks else tmp
    return SiteOp(blocks)


def make_twosite_mpo(m1: MPOTensor, m2: MPOTensor, phys1, phys2, su2=True, progress=None):
    """Two-site MPO tensor for sites (p, p+1) (ts_ops.h:77-126).  progress(msg),
    when given, is called after every tenth of the rows."""
    assert m1.col_dim == m2.row_dim
    table = OpTable()
    prempo = []
    cache = {}
    step = max(1, m1.row_dim // 10)
    for b1 in range(m1.row_dim):
        if progress and b1 and b1 % step == 0:
            progress(f"two-site MPO: row {b1}/{m1.row_dim}")
        row1 = m1.rows[b1]
        # the b3 reached from b1 through some b2, in ascending order; for each, the
        # b2 in the order of row b1
        reach = {}
        for b2 in row1:
            for b3 in m2.rows[b2]:
                reach.setdefault(b3, []).append(b2)
        for b3 in sorted(reach):
            summands = reach[b3]
            ret = {}
            for b2 in summands:
                p1 = m1.at(b1, b2)
                p2 = m2.at(b2, b3)
                for (t1, s1) in p1:
                    for (t2, s2) in p2:
                        o1, o2 = m1.table[t1], m2.table[t2]
                        if su2:
                            spins = _allowed_spins(m1.left_spins[b1], m2.right_spins[b3], o1.twoS, o2.twoS)
                        else:
                            spins = [0]
                        for s in spins:
                            ck = (id(m1.table), t1, id(m2.table), t2, m1.left_spins[b1],
                                  m1.right_spins[b2], m2.right_spins[b3], s)
                            prod = cache.get(ck)
                            if prod is None:
                                if su2:
                                    prod = op_kron_su2(phys1, phys2, o1, o2, m1.left_spins[b1],
                                                       m1.right_spins[b2], m2.right_spins[b3], s)
                                else:
                                    prod = op_kron_abelian(phys1, phys2, o1, o2)
                                prod.remove_empty_blocks()
                                cache[ck] = prod
                            sc = prod.scaled(s1 * s2)
                            if s in ret:
                                ret[s].iadd(sc)
                            else:
                                ret[s] = sc
            for s in sorted(ret):
                tag = table.register(ret[s])
                prempo.append((b1, b3, tag, 1.0))
    return MPOTensor(m1.row_dim, m2.col_dim, prempo, table,
                     m1.left_herm, m2.right_herm, m1.left_phase, m2.right_phase,
                     m1.left_spins, m2.right_spins)
