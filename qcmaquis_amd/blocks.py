"""Block-sparse containers of the symmetry-adapted DMRG data model.

Charges are tuples of two ints.  SU(2)xU(1) uses (N, 2S); U(1)xU(1) uses
(N_alpha, N_beta).  Both fuse by component-wise addition
(dmrg/framework/dmrg/block_matrix/symmetry/su2u1.h:37-40,
 symmetry/nu1_tpl.h fuse) and compare lexicographically.

`Index` keeps (charge, size) pairs sorted by *descending* charge
(indexing_stable.hpp:82-130, index_detail::gt); `BlockMatrix` keeps its
blocks sorted by descending (left charge, right charge)
(dual_index.h:39-52 gt, block_matrix.hpp:136-145 insert_block).
Dense blocks are numpy arrays interpreted column-major when flattened,
as the reference's alps::numeric::matrix is.
"""
from __future__ import annotations

import numpy as np

SU2U1 = "su2u1"
TWOU1 = "2u1"


def fuse(a, b):
    return (a[0] + b[0], a[1] + b[1])


def neg(a):
    return (-a[0], -a[1])


def spin(c):
    """SU(2) doubled spin of an SU2U1 charge (su2u1.h:30-31)."""
    return c[1]


def triangle(a, b, c):
    """Triangle rule on doubled spins (gsl_coupling.h:25-28)."""
    return ((a + b + c) % 2 == 0) and abs(a - b) <= c <= a + b


class Index:
    """Sorted list of (charge, size), descending in charge."""

    __slots__ = ("_c", "_s", "_pos")

    def __init__(self, pairs=()):
        d = {}
        for c, s in pairs:
            c = (int(c[0]), int(c[1]))
            d[c] = d.get(c, 0) + int(s) if c in d else int(s)
        self._c = sorted(d.keys(), reverse=True)
        self._s = [d[c] for c in self._c]
        self._pos = {c: i for i, c in enumerate(self._c)}

    @staticmethod
    def from_dict(d):
        ix = Index()
        ix._c = sorted(d.keys(), reverse=True)
        ix._s = [int(d[c]) for c in ix._c]
        ix._pos = {c: i for i, c in enumerate(ix._c)}
        return ix

    def __len__(self):
        return len(self._c)

    def __iter__(self):
        return iter(zip(self._c, self._s))

    def __getitem__(self, i):
        return (self._c[i], self._s[i])

    def __eq__(self, other):
        return isinstance(other, Index) and self._c == other._c and self._s == other._s

    def __repr__(self):
        return "Index(" + ", ".join(f"{c}:{s}" for c, s in self) + ")"

    def charges(self):
        return list(self._c)

    def sizes(self):
        return list(self._s)

    def has(self, c):
        return c in self._pos

    def position(self, c):
        return self._pos.get(c, len(self._c))

    def size_of_block(self, c, default=None):
        p = self._pos.get(c)
        if p is None:
            if default is not None:
                return default
            raise KeyError(c)
        return self._s[p]

    def sum_of_sizes(self):
        return sum(self._s)

    def as_dict(self):
        return dict(zip(self._c, self._s))

    def __mul__(self, other):
        """Fused product index (indexing_stable.hpp operator*)."""
        d = {}
        for c1, s1 in self:
            for c2, s2 in other:
                pc = fuse(c1, c2)
                d[pc] = d.get(pc, 0) + s1 * s2
        return Index.from_dict(d)

    def adjoin(self):
        return Index.from_dict({neg(c): s for c, s in self})

    def common_subset(self, other):
        """Charges present in both (indexing_stable.hpp common_subset)."""
        return Index.from_dict({c: s for c, s in self if other.has(c)})


class ProductBasis:
    """Offsets of (a_charge, b_charge) sub-blocks inside the fused index.

    Restates indexing_stable.hpp:305-375: iterate a (outer) then b (inner) in
    Index order, offset = running size of the fused charge.  `right=True`
    uses the fusion f(a, b) = fuse(-a, b) used for right-paired tensors.
    """

    __slots__ = ("off", "size")

    def __init__(self, a: Index, b: Index, right=False):
        self.off = {}
        self.size = {}
        for c1, s1 in a:
            for c2, s2 in b:
                pc = fuse(neg(c1), c2) if right else fuse(c1, c2)
                cur = self.size.get(pc, 0)
                self.off[(c1, c2)] = cur
                self.size[pc] = cur + s1 * s2

    def __call__(self, a, b):
        return self.off[(a, b)]

    def get(self, a, b):
        return self.off.get((a, b))


class BlockMatrix:
    """Block-diagonal-in-charge matrix: {(lc, rc): ndarray}."""

    __slots__ = ("blocks",)

    def __init__(self, blocks=None):
        self.blocks = {} if blocks is None else blocks

    def keys_sorted(self):
        return sorted(self.blocks.keys(), reverse=True)

    def n_blocks(self):
        return len(self.blocks)

    def copy(self):
        return BlockMatrix({k: v.copy() for k, v in self.blocks.items()})

    def has(self, lc, rc):
        return (lc, rc) in self.blocks

    def get(self, lc, rc):
        return self.blocks.get((lc, rc))

    def insert(self, lc, rc, m):
        self.blocks[(lc, rc)] = m

    def match_and_add(self, lc, rc, m):
        k = (lc, rc)
        if k in self.blocks:
            self.blocks[k] = self.blocks[k] + m
        else:
            self.blocks[k] = m.copy()

    def left_basis(self):
        return Index.from_dict({lc: m.shape[0] for (lc, rc), m in self.blocks.items()})

    def right_basis(self):
        return Index.from_dict({rc: m.shape[1] for (lc, rc), m in self.blocks.items()})

    def transpose(self):
        return BlockMatrix({(rc, lc): m.T.copy() for (lc, rc), m in self.blocks.items()})

    def scale(self, a):
        return BlockMatrix({k: v * a for k, v in self.blocks.items()})

    def norm(self):
        return float(np.sqrt(sum(float(np.sum(v * v)) for v in self.blocks.values())))

    def dot(self, other):
        """Sum over common blocks (block_matrix scalar_overlap)."""
        s = 0.0
        for k, v in self.blocks.items():
            w = other.blocks.get(k)
            if w is not None:
                s += float(np.sum(v * w))
        return s

    def num_elements(self):
        return sum(v.size for v in self.blocks.values())

    def __add__(self, other):
        r = self.copy()
        for k, v in other.blocks.items():
            r.match_and_add(k[0], k[1], v)
        return r

    def __sub__(self, other):
        return self + other.scale(-1.0)


def gemm(A: BlockMatrix, B: BlockMatrix) -> BlockMatrix:
    """Charge-matched block product C(lc, rc) += A(lc, x) B(x, rc)."""
    byleft = {}
    for (lc, rc), m in B.blocks.items():
        byleft.setdefault(lc, []).append((rc, m))
    C = BlockMatrix()
    for (lc, x), a in A.blocks.items():
        for rc, b in byleft.get(x, ()):
            C.match_and_add(lc, rc, a @ b)
    return C
