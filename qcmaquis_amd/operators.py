"""Site operators and operator tables for the quantum-chemistry Hamiltonian.

A `SiteOp` is a block matrix over the physical index whose block key is
(input charge, output charge): the W-application reads the ket physical
charge from the left charge of a block
(contractions/non-abelian/apply_op.hpp:67-70).  SU(2) operators carry a
spin descriptor (2S of the reduced tensor, action = change of the MPO bond
spin; block_matrix/symmetry/spin_descriptor.h:44-80) and a per-row/column
"spin basis" used for the sparse entries (sparse_operator.h:105-168,
site_operator.hpp:312-341).
"""
from __future__ import annotations

import numpy as np

from .blocks import Index


class SiteOp:
    __slots__ = ("blocks", "twoS", "action", "spin_basis", "_sparse")

    def __init__(self, blocks=None, twoS=0, action=0, spin_basis=None):
        self.blocks = {} if blocks is None else blocks  # (cin, cout) -> ndarray
        self.twoS = twoS
        self.action = action
        self.spin_basis = spin_basis  # (cin,cout) -> (row_spins, col_spins) or None
        self._sparse = None

    def copy(self):
        sb = None if self.spin_basis is None else {k: (list(v[0]), list(v[1])) for k, v in self.spin_basis.items()}
        return SiteOp({k: v.copy() for k, v in self.blocks.items()}, self.twoS, self.action, sb)

    def with_spin(self, twoS, action):
        r = self.copy()
        r.twoS, r.action = twoS, action
        return r

    def add_block(self, cin, cout, val):
        self.blocks[(cin, cout)] = np.array([[float(val)]])
        self._sparse = None

    def sorted_keys(self):
        return sorted(self.blocks.keys(), reverse=True)

    def scaled(self, s):
        # blocks and spin-basis lists are never changed in place (iadd replaces
        # them), so the scaled operator shares the lists and owns its dicts
        sb = None if self.spin_basis is None else dict(self.spin_basis)
        return SiteOp({k: v * s for k, v in self.blocks.items()}, self.twoS, self.action, sb)

    def iadd(self, other):
        """SiteOperator += (site_operator.hpp:92-103, extend_spin_basis)."""
        if not self.blocks:
            self.twoS, self.action = other.twoS, other.action
        for k, v in other.blocks.items():
            if k in self.blocks:
                a = self.blocks[k]
                if a.shape != v.shape:
                    r = max(a.shape[0], v.shape[0])
                    c = max(a.shape[1], v.shape[1])
                    na = np.zeros((r, c))
                    na[:a.shape[0], :a.shape[1]] = a
                    na[:v.shape[0], :v.shape[1]] += v
                    self.blocks[k] = na
                else:
                    self.blocks[k] = a + v
            else:
                self.blocks[k] = v.copy()
        if other.spin_basis is not None:
            if self.spin_basis is None:
                self.spin_basis = {}
            for k, (rs, cs) in other.spin_basis.items():
                sr, sc = self.spin_basis.get(k, ([], []))
                sr = list(sr) + [0] * max(0, len(rs) - len(sr))
                sc = list(sc) + [0] * max(0, len(cs) - len(sc))
                for i, v in enumerate(rs):
                    if v != 0:
                        sr[i] = v
                for i, v in enumerate(cs):
                    if v != 0:
                        sc[i] = v
                self.spin_basis[k] = (sr, sc)
        self._sparse = None
        return self

    def remove_empty_blocks(self):
        for k in list(self.blocks.keys()):
            if not np.any(self.blocks[k] != 0.0):
                del self.blocks[k]
        self._sparse = None

    def sparse(self):
        """Sparse entries per block in sorted-block order.

        Returns list of (cin, cout, rows, cols, [(row, col, row_spin, col_spin, coeff), ...]).
        Default spin basis = |2S| of the block charges (site_operator.hpp:321-332).
        """
        if self._sparse is not None:
            return self._sparse
        out = []
        for k in self.sorted_keys():
            m = self.blocks[k]
            cin, cout = k
            if self.spin_basis is not None and k in self.spin_basis:
                rs, cs = self.spin_basis[k]
            else:
                rs = [abs(cin[1])] * m.shape[0]
                cs = [abs(cout[1])] * m.shape[1]
            ents = []
            ncs = len(cs)
            for i, row in enumerate(m.tolist()):
                ri = rs[i] if i < len(rs) else 0
                for j, v in enumerate(row):
                    if v != 0.0:
                        ents.append((i, j, ri, cs[j] if j < ncs else 0, v))
            out.append((cin, cout, m.shape[0], m.shape[1], ents))
        self._sparse = out
        return out

    def first_block(self):
        return self.sorted_keys()[0]


def proportional(a: SiteOp, b: SiteOp):
    """Equality modulo a scale factor, including shape and spin (models/tag_detail.h:56-)."""
    if a.twoS != b.twoS or a.action != b.action:
        return False
    if sorted(a.blocks.keys()) != sorted(b.blocks.keys()):
        return False
    for k in a.blocks:
        if a.blocks[k].shape != b.blocks[k].shape:
            return False
    if not a.blocks:
        return True
    ka = a.sorted_keys()[0]
    fa = a.blocks[ka].ravel()
    fb = b.blocks[ka].ravel()
    ia = np.flatnonzero(np.abs(fa) > 1e-50)
    ib = np.flatnonzero(np.abs(fb) > 1e-50)
    if len(ia) == 0 or len(ib) == 0:
        return False
    s = fb[ib[0]] / fa[ia[0]]
    for k in a.blocks:
        if not np.allclose(a.blocks[k] * s, b.blocks[k], rtol=1e-12, atol=1e-14):
            return False
    return True


class OpTable:
    """Operator tag table with hermitian partners and fermion flags (TagHandler)."""

    def __init__(self):
        self.ops = []
        self.herm = []
        self.fermionic = []

    def register(self, op: SiteOp, fermionic=False):
        self.ops.append(op)
        self.herm.append(len(self.ops) - 1)
        self.fermionic.append(fermionic)
        return len(self.ops) - 1

    def hermitian_pair(self, t1, t2):
        if self.herm[t1] == t2 and self.herm[t2] == t1:
            return
        assert self.herm[t1] == t1 and self.herm[t2] == t2
        self.herm[t1], self.herm[t2] = t2, t1

    def herm_conj(self, t):
        return self.herm[t]

    def __getitem__(self, t):
        return self.ops[t]

    def __len__(self):
        return len(self.ops)


# ---------------------------------------------------------------------------
# SU(2) x U(1) one-site operators (models/chem/su2u1/term_maker.h:136-377)
# ---------------------------------------------------------------------------
A2, B2, C2, D2 = (2, 0), (1, 1), (1, -1), (0, 0)


def su2_phys_index():
    return Index([(A2, 1), (B2, 1), (C2, 1), (D2, 1)])


def _op(entries, twoS=0, action=0):
    o = SiteOp(twoS=twoS, action=action)
    for cin, cout, v in entries:
        o.add_block(cin, cout, v)
    return o


class SU2Ops:
    """Tags of the SU(2) operator set, registered in the reference's order."""

    def __init__(self, table: OpTable):
        s2 = np.sqrt(2.0)
        r32 = np.sqrt(3.0 / 2.0)
        half_up = (1, 1)      # SpinDescriptor(1,0,1): 2S=1, action +1
        half_down = (1, -1)   # SpinDescriptor(1,1,0)
        one_up = (2, 2)       # (2,0,2)
        one_flat = (2, 0)     # (2,1,1)
        one_down = (2, -2)    # (2,2,0)
        A, B, C, D = A2, B2, C2, D2
        ident = _op([(A, A, 1), (B, B, 1), (C, C, 1), (D, D, 1)])
        ident_full = _op([(A, A, 1), (D, D, 1), (B, B, 1), (C, C, 1), (B, C, 1), (C, B, 1)])
        fill = _op([(A, A, 1), (D, D, 1), (B, B, -1), (C, C, -1), (B, C, -1), (C, B, -1)])
        create_fill = _op([(B, A, s2), (C, A, s2), (D, B, 1), (D, C, 1)], *half_up)
        destroy = _op([(A, B, 1), (A, C, 1), (B, D, s2), (C, D, s2)], *half_down)
        destroy_fill = _op([(A, B, 1), (A, C, 1), (B, D, -s2), (C, D, -s2)], *half_up)
        create = _op([(B, A, s2), (C, A, s2), (D, B, -1), (D, C, -1)], *half_down)
        create_fill_couple_down = create_fill.with_spin(*half_down)
        destroy_fill_couple_down = destroy_fill.with_spin(*half_down)
        create_couple_up = create.with_spin(*half_up)
        destroy_couple_up = destroy.with_spin(*half_up)
        create_fill_count = _op([(B, A, s2), (C, A, s2)], *half_up)
        destroy_count = _op([(A, B, 1), (A, C, 1)], *half_down)
        destroy_fill_count = _op([(A, B, 1), (A, C, 1)], *half_up)
        create_count = _op([(B, A, s2), (C, A, s2)], *half_down)
        count = _op([(A, A, 2), (B, B, 1), (C, C, 1)])
        docc = _op([(A, A, 1)])
        e2d = _op([(D, A, 1)])
        d2e = _op([(A, D, 1)])
        count_fill = _op([(A, A, 2), (B, B, -1), (C, C, -1), (B, C, -1), (C, B, -1)])
        flip_to_S2 = _op([(B, B, r32), (C, C, r32), (B, C, r32), (C, B, r32)], *one_up)
        flip_to_S0 = flip_to_S2.with_spin(*one_down)
        flip_S0 = flip_to_S2.with_spin(*one_flat)

        reg = table.register
        self.ident = reg(ident)
        self.ident_full = reg(ident_full)
        self.fill = reg(fill)
        self.create_fill = reg(create_fill, True)
        self.create = reg(create, True)
        self.destroy_fill = reg(destroy_fill, True)
        self.destroy = reg(destroy, True)
        self.create_fill_couple_down = reg(create_fill_couple_down, True)
        self.destroy_fill_couple_down = reg(destroy_fill_couple_down, True)
        self.create_couple_up = reg(create_couple_up, True)
        self.destroy_couple_up = reg(destroy_couple_up, True)
        self.create_fill_count = reg(create_fill_count, True)
        self.create_count = reg(create_count, True)
        self.destroy_fill_count = reg(destroy_fill_count, True)
        self.destroy_count = reg(destroy_count, True)
        self.count = reg(count)
        self.docc = reg(docc)
        self.e2d = reg(e2d)
        self.d2e = reg(d2e)
        self.flip_S0 = reg(flip_S0)
        self.flip_to_S2 = reg(flip_to_S2)
        self.flip_to_S0 = reg(flip_to_S0)
        self.count_fill = reg(count_fill)

        hp = table.hermitian_pair
        hp(self.create_fill, self.destroy_fill)
        hp(self.create, self.destroy)
        hp(self.e2d, self.d2e)
        hp(self.create_fill_count, self.destroy_fill_count)
        hp(self.create_count, self.destroy_count)
        hp(self.create_fill_couple_down, self.destroy_fill_couple_down)
        hp(self.create_couple_up, self.destroy_couple_up)

        # operator bundles (term_maker.h:321-377)
        self.b_create = dict(couple_up=self.create_couple_up, couple_down=self.create,
                             fill_couple_up=self.create_fill, fill_couple_down=self.create_fill_couple_down)
        self.b_destroy = dict(couple_up=self.destroy_couple_up, couple_down=self.destroy,
                              fill_couple_up=self.destroy_fill, fill_couple_down=self.destroy_fill_couple_down)
        self.b_create_count = dict(couple_down=self.create_count, fill_couple_up=self.create_fill_count)
        self.b_destroy_count = dict(couple_down=self.destroy_count, fill_couple_up=self.destroy_fill_count)
        self.b_count = dict(no_couple=self.count, fill_no_couple=self.count_fill)
        self.b_flip = dict(no_couple=self.flip_S0, couple_up=self.flip_to_S2, couple_down=self.flip_to_S0)


# ---------------------------------------------------------------------------
# U(1) x U(1) one-site operators (models/chem/2u1/model.hpp:25-171)
# ---------------------------------------------------------------------------
A1, B1, C1, D1 = (0, 0), (1, 0), (0, 1), (1, 1)


def twou1_phys_index():
    return Index([(A1, 1), (B1, 1), (C1, 1), (D1, 1)])


def _block_gemm(x: SiteOp, y: SiteOp) -> SiteOp:
    """Operator block product x*y on (in,out) keyed blocks, as block_matrix gemm."""
    r = SiteOp()
    for (li, lo), a in x.blocks.items():
        for (ri, ro), b in y.blocks.items():
            if lo == ri:
                k = (li, ro)
                r.blocks[k] = r.blocks.get(k, 0) + a @ b
    return r


class TwoU1Ops:
    def __init__(self, table: OpTable):
        A, B, C, D = A1, B1, C1, D1
        ident = _op([(A, A, 1), (B, B, 1), (C, C, 1), (D, D, 1)])
        create_up = _op([(A, B, 1), (C, D, 1)])
        create_down = _op([(A, C, 1), (B, D, 1)])
        destroy_up = _op([(B, A, 1), (D, C, 1)])
        destroy_down = _op([(C, A, 1), (D, B, 1)])
        count_up = _op([(B, B, 1), (D, D, 1)])
        count_down = _op([(C, C, 1), (D, D, 1)])
        count_up_down = _op([(B, B, 1), (C, C, 1), (D, D, 2)])
        docc = _op([(D, D, 1)])
        e2d = _op([(A, D, 1)])
        d2e = _op([(D, A, 1)])
        fill = _op([(A, A, 1), (B, B, -1), (C, C, -1), (D, D, 1)])
        create_down = _block_gemm(fill, create_down)
        destroy_down = _block_gemm(destroy_down, fill)
        self.table = table
        reg = table.register
        self.ident = reg(ident)
        self.fill = reg(fill)
        self.create_up = reg(create_up, True)
        self.create_down = reg(create_down, True)
        self.destroy_up = reg(destroy_up, True)
        self.destroy_down = reg(destroy_down, True)
        self.count_up = reg(count_up)
        self.count_down = reg(count_down)
        self.e2d = reg(e2d)
        self.d2e = reg(d2e)
        self.docc = reg(docc)
        self.count_up_down = reg(count_up_down)
        table.hermitian_pair(self.create_up, self.destroy_up)
        table.hermitian_pair(self.create_down, self.destroy_down)
        table.hermitian_pair(self.e2d, self.d2e)
        self._products = {}

    def product(self, t1, t2):
        """Tag of op(t1)*op(t2), registered once (TagHandler::get_product_tag)."""
        key = (t1, t2)
        if key in self._products:
            return self._products[key]
        tb = self.table
        p = _block_gemm(tb[t1], tb[t2])
        p.remove_empty_blocks()
        ferm = tb.fermionic[t1] != tb.fermionic[t2]
        for t, o in enumerate(tb.ops):
            if proportional(o, p):
                # scale relative to the stored op
                k = o.sorted_keys()[0]
                s = p.blocks[k].ravel()[np.flatnonzero(o.blocks[k].ravel())[0]] / \
                    o.blocks[k].ravel()[np.flatnonzero(o.blocks[k].ravel())[0]]
                self._products[key] = (t, float(s))
                return self._products[key]
        t = tb.register(p, ferm)
        self._products[key] = (t, 1.0)
        return self._products[key]
