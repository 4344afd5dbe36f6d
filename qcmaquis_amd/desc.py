"""ctypes mirror of include/mqgpu.h and builders for its flat descriptors."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .blocks import Index

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)

MQ_SYMM_TWOU1 = 0
MQ_SYMM_SU2U1 = 1
MQ_PREC_FP64 = 0
MQ_PREC_FP32 = 1
MQ_PREC_BF16X3 = 2


class mq_index(C.Structure):
    _fields_ = [("n", C.c_int32), ("charges", i32p), ("sizes", i32p)]


class mq_blocks(C.Structure):
    _fields_ = [("n", C.c_int32), ("lc", i32p), ("rc", i32p), ("rows", i32p), ("cols", i32p), ("offset", i64p)]


class mq_boundary_desc(C.Structure):
    _fields_ = [("aux_dim", C.c_int32), ("bptr", i32p), ("blocks", mq_blocks), ("total", C.c_int64)]


class mq_mpo_desc(C.Structure):
    _fields_ = [("row_dim", C.c_int32), ("col_dim", C.c_int32),
                ("col_ptr", i32p), ("row_idx", i32p), ("term_ptr", i32p), ("term_op", i32p), ("term_scale", f64p),
                ("n_ops", C.c_int32), ("op_spin", i32p), ("op_bptr", i32p), ("opb_in", i32p), ("opb_out", i32p),
                ("opb_eptr", i32p), ("ent_row", i32p), ("ent_col", i32p), ("ent_rspin", i32p), ("ent_cspin", i32p),
                ("ent_coeff", f64p), ("left_spin", i32p), ("right_spin", i32p), ("left_herm", i32p),
                ("right_herm", i32p), ("left_phase", i32p), ("right_phase", i32p)]


class mq_site_desc(C.Structure):
    _fields_ = [("symm", C.c_int32), ("phys", mq_index), ("left", mq_index), ("right", mq_index),
                ("mpo", mq_mpo_desc), ("L", mq_boundary_desc), ("R", mq_boundary_desc)]


def _p(a, t):
    return a.ctypes.data_as(t)


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))


class IndexArrays:
    __slots__ = ("charges", "sizes", "n")

    def __init__(self, ix: Index):
        self.n = len(ix)
        self.charges = _i32([c for cc in ix.charges() for c in cc]) if self.n else np.zeros(2, np.int32)
        self.sizes = _i32(ix.sizes()) if self.n else np.zeros(1, np.int32)

    def c(self):
        return mq_index(self.n, _p(self.charges, i32p), _p(self.sizes, i32p))


class BoundaryLayout:
    """Block structure of a boundary: per bond b, blocks [bptr[b], bptr[b+1])."""

    __slots__ = ("aux", "bptr", "lc", "rc", "rows", "cols", "offset", "total", "_key")

    def __init__(self, aux, bptr, lc, rc, rows, cols, offset, total):
        self.aux = int(aux)
        self.bptr = _i32(bptr)
        n = int(self.bptr[-1])
        self.lc = _i32(lc) if n else np.zeros(2, np.int32)
        self.rc = _i32(rc) if n else np.zeros(2, np.int32)
        self.rows = _i32(rows) if n else np.zeros(1, np.int32)
        self.cols = _i32(cols) if n else np.zeros(1, np.int32)
        self.offset = np.ascontiguousarray(np.asarray(offset, dtype=np.int64).reshape(-1)) if n else np.zeros(1, np.int64)
        self.total = int(total)
        self._key = None

    @property
    def nblocks(self):
        return int(self.bptr[-1])

    def c(self):
        return mq_boundary_desc(self.aux, _p(self.bptr, i32p),
                                mq_blocks(self.nblocks, _p(self.lc, i32p), _p(self.rc, i32p), _p(self.rows, i32p),
                                          _p(self.cols, i32p), _p(self.offset, i64p)), self.total)

    @staticmethod
    def from_c(d: mq_boundary_desc):
        aux = d.aux_dim
        bptr = np.ctypeslib.as_array(d.bptr, shape=(aux + 1,)).copy()
        n = int(bptr[-1])
        if n == 0:
            return BoundaryLayout(aux, bptr, [], [], [], [], [], d.total)
        b = d.blocks
        return BoundaryLayout(aux, bptr,
                              np.ctypeslib.as_array(b.lc, shape=(2 * n,)).copy(),
                              np.ctypeslib.as_array(b.rc, shape=(2 * n,)).copy(),
                              np.ctypeslib.as_array(b.rows, shape=(n,)).copy(),
                              np.ctypeslib.as_array(b.cols, shape=(n,)).copy(),
                              np.ctypeslib.as_array(b.offset, shape=(n,)).copy(), d.total)

    @staticmethod
    def seed(charge):
        """Left/right boundary seed: one 1x1 block of ones (mps.hpp:244-266)."""
        return BoundaryLayout(1, [0, 1], [charge[0], charge[1]], [charge[0], charge[1]], [1], [1], [0], 1)

    def blocks_of(self, b):
        for k in range(int(self.bptr[b]), int(self.bptr[b + 1])):
            yield ((int(self.lc[2 * k]), int(self.lc[2 * k + 1])), (int(self.rc[2 * k]), int(self.rc[2 * k + 1])),
                   int(self.rows[k]), int(self.cols[k]), int(self.offset[k]))

    def to_dicts(self, data):
        """[{(lc, rc): ndarray}] per bond, from a host arena."""
        out = []
        for b in range(self.aux):
            d = {}
            for lc, rc, r, c, off in self.blocks_of(b):
                d[(lc, rc)] = np.asarray(data[off:off + r * c]).reshape((r, c), order="F")
            out.append(d)
        return out


class MPOArrays:
    """Flattened MPO tensor (one per site or site pair), built once and cached."""

    def __init__(self, w):
        tb = w.table
        used = w.used_tags()
        opid = {t: i for i, t in enumerate(used)}
        col_ptr = [0]
        row_idx, term_ptr, term_op, term_scale = [], [0], [], []
        for b2 in range(w.col_dim):
            for b1 in w.cols[b2]:
                row_idx.append(b1)
                for tag, sc in w.entries[(b1, b2)]:
                    term_op.append(opid[tag])
                    term_scale.append(sc)
                term_ptr.append(len(term_op))
            col_ptr.append(len(row_idx))
        op_spin, op_bptr = [], [0]
        opb_in, opb_out, opb_eptr = [], [], [0]
        er, ec, ers, ecs, eco = [], [], [], [], []
        for t in used:
            op = tb[t]
            op_spin.append(op.twoS)
            for cin, cout, rows, cols, ents in op.sparse():
                opb_in += [cin[0], cin[1]]
                opb_out += [cout[0], cout[1]]
                for (r, c, rs, cs, v) in ents:
                    er.append(r)
                    ec.append(c)
                    ers.append(rs)
                    ecs.append(cs)
                    eco.append(v)
                opb_eptr.append(len(er))
            op_bptr.append(len(opb_eptr) - 1)
        self.row_dim, self.col_dim = w.row_dim, w.col_dim
        self.col_ptr = _i32(col_ptr)
        self.row_idx = _i32(row_idx if row_idx else [0])
        self.term_ptr = _i32(term_ptr)
        self.term_op = _i32(term_op if term_op else [0])
        self.term_scale = np.ascontiguousarray(np.asarray(term_scale if term_scale else [0.0], dtype=np.float64))
        self.n_ops = len(used)
        self.op_spin = _i32(op_spin if op_spin else [0])
        self.op_bptr = _i32(op_bptr)
        self.opb_in = _i32(opb_in if opb_in else [0, 0])
        self.opb_out = _i32(opb_out if opb_out else [0, 0])
        self.opb_eptr = _i32(opb_eptr)
        self.ent_row = _i32(er if er else [0])
        self.ent_col = _i32(ec if ec else [0])
        self.ent_rspin = _i32(ers if ers else [0])
        self.ent_cspin = _i32(ecs if ecs else [0])
        self.ent_coeff = np.ascontiguousarray(np.asarray(eco if eco else [0.0], dtype=np.float64))
        self.left_spin = _i32(w.left_spins)
        self.right_spin = _i32(w.right_spins)
        self.left_herm = _i32(w.left_herm)
        self.right_herm = _i32(w.right_herm)
        self.left_phase = _i32(w.left_phase)
        self.right_phase = _i32(w.right_phase)

    def c(self):
        P = lambda a: _p(a, i32p)
        return mq_mpo_desc(self.row_dim, self.col_dim, P(self.col_ptr), P(self.row_idx), P(self.term_ptr),
                           P(self.term_op), _p(self.term_scale, f64p), self.n_ops, P(self.op_spin), P(self.op_bptr),
                           P(self.opb_in), P(self.opb_out), P(self.opb_eptr), P(self.ent_row), P(self.ent_col),
                           P(self.ent_rspin), P(self.ent_cspin), _p(self.ent_coeff, f64p), P(self.left_spin),
                           P(self.right_spin), P(self.left_herm), P(self.right_herm), P(self.left_phase),
                           P(self.right_phase))


def mpo_arrays(w):
    a = getattr(w, "_mq_arrays", None)
    if a is None:
        a = MPOArrays(w)
        w._mq_arrays = a
    return a


class mq_split_desc(C.Structure):
    _fields_ = [("symm", C.c_int32), ("phys1", mq_index), ("phys2", mq_index), ("left", mq_index),
                ("right", mq_index)]


class SplitDesc:
    """Owns the numpy arrays behind one mq_split_desc (two-site split)."""

    def __init__(self, symm, phys1, phys2, left, right):
        self.keep = [IndexArrays(phys1), IndexArrays(phys2), IndexArrays(left), IndexArrays(right)]
        self.desc = mq_split_desc(symm, *(k.c() for k in self.keep))

    def ptr(self):
        return C.byref(self.desc)


class SiteDesc:
    """Owns the numpy arrays behind one mq_site_desc."""

    def __init__(self, symm, phys, left, right, w, Llay: BoundaryLayout | None, Rlay: BoundaryLayout | None):
        self.keep = [IndexArrays(phys), IndexArrays(left), IndexArrays(right), mpo_arrays(w), Llay, Rlay]
        empty = BoundaryLayout(0, [0], [], [], [], [], [], 0)
        self.keep.append(empty)
        self.desc = mq_site_desc(symm, self.keep[0].c(), self.keep[1].c(), self.keep[2].c(), self.keep[3].c(),
                                 (Llay or empty).c(), (Rlay or empty).c())

    def ptr(self):
        return C.byref(self.desc)
