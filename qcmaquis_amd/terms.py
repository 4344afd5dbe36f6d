"""Hamiltonian term generation for the spin-adapted (SU2U1) QC model.

Restates models/chem/su2u1/model.hpp:47-298 (create_terms), the SpinSumSU2
and TermMakerSU2 helpers of models/chem/su2u1/term_maker.h:378-791 and the
term merging of models/chem/su2u1/chem_helper.h:47-95.  A term is
(coeff, is_fermionic, [(pos, tag), ...]).
"""
from __future__ import annotations

import math


class Term:
    __slots__ = ("coeff", "fermionic", "ops")

    def __init__(self, coeff=1.0, fermionic=False, ops=None):
        self.coeff = coeff
        self.fermionic = fermionic
        self.ops = [] if ops is None else ops

    def positions(self):
        return [p for p, _ in self.ops]

    def tags(self):
        return [t for _, t in self.ops]

    def canonical_order(self):
        self.ops.sort(key=lambda x: x[0])

    def __repr__(self):
        return f"Term({self.coeff:+.6g}, {self.ops})"


def align(i, j, k, l):
    """Canonical integral index order i>=j, k>=l, (i,j)>=(k,l) (utils/align.h:18-28)."""
    if i < j:
        i, j = j, i
    if k < l:
        k, l = l, k
    if i < k:
        i, k = k, i
        j, l = l, j
    if i == k and j < l:
        j, l = l, j
    return (i, j, k, l)


def _sgn(i, j, k, l):
    idx = [i, j, k, l]
    inv = 0
    for a in range(3):
        for b in range(a + 1, 4):
            if idx[a] > idx[b]:
                inv += 1
    return inv % 2 != 0


# ----------------------------- TermMakerSU2 --------------------------------

def tm_two_term(sign, scale, i, j, op1, op2):
    return Term(scale, sign, [(i, op1), (j, op2)])


def tm_positional_two_term(sign, scale, i, j, op1, op1_fill, op2, op2_fill):
    t = Term(scale, sign)
    op1_use = op1_fill if i < j else op2_fill
    op2_use = op2 if i < j else op1
    if j < i and sign:
        t.coeff = -t.coeff
    start, end = min(i, j), max(i, j)
    t.ops = [(start, op1_use), (end, op2_use)]
    return t


def tm_three_term(scale, pb, p1, p2, boson_op, boson_op_fill, op1, op1_fill, op2, op2_fill):
    t = Term(scale, True)
    op1_use = op1_fill if p1 < p2 else op2_fill
    op2_use = op2 if p1 < p2 else op1
    if (p1 < pb < p2) or (p2 < pb < p1):
        boson_use = boson_op_fill
    else:
        boson_use = boson_op
    if p2 < p1:
        t.coeff = -t.coeff
    start, end = min(p1, p2), max(p1, p2)
    t.ops = [(pb, boson_use), (start, op1_use), (end, op2_use)]
    t.canonical_order()
    return t


def tm_four_term(max_two_S, scale, i, j, k, l, op_i, op_k):
    t = Term(scale, True)
    if _sgn(i, j, k, l):
        t.coeff = -t.coeff
    sterm = sorted([(i, op_i), (j, op_i), (k, op_k), (l, op_k)], key=lambda x: x[0])
    if max_two_S == 2:
        t.ops = [(sterm[0][0], sterm[0][1]["fill_couple_up"]),
                 (sterm[1][0], sterm[1][1]["couple_up"]),
                 (sterm[2][0], sterm[2][1]["fill_couple_down"]),
                 (sterm[3][0], sterm[3][1]["couple_down"])]
    else:
        t.ops = [(sterm[0][0], sterm[0][1]["fill_couple_up"]),
                 (sterm[1][0], sterm[1][1]["couple_down"]),
                 (sterm[2][0], sterm[2][1]["fill_couple_up"]),
                 (sterm[3][0], sterm[3][1]["couple_down"])]
    return t


# ------------------------------- SpinSumSU2 --------------------------------

class SpinSumSU2:
    def __init__(self, ops):
        self.o = ops

    def V_term(self, v, i, k, l, j):
        n_unique = len(set([i, k, l, j]))
        if n_unique == 4:
            return self.four_term(v, i, k, l, j)
        if n_unique == 3:
            return self.three_term(v, i, k, l, j)
        if n_unique == 2:
            return self.two_term(v, i, k, l, j)
        return [Term(2.0 * v, False, [(i, self.o.docc)])]

    def two_term(self, v, i, k, l, j):
        if i == j and k == l and j != k:
            return [tm_two_term(False, v, i, k, self.o.count, self.o.count)]
        if i == k and j == l and j != k:
            return [tm_two_term(False, 2.0 * v, i, j, self.o.e2d, self.o.d2e)]
        if i == l and j == k and i != j:
            o = self.o
            return [tm_positional_two_term(False, math.sqrt(3.0) * v, i, j,
                                           o.b_flip["couple_down"], o.b_flip["couple_up"],
                                           o.b_flip["couple_down"], o.b_flip["couple_up"]),
                    tm_two_term(False, -0.5 * v, i, j, o.count, o.count)]
        if (i == k == l) or (k == l == j) or (i == l == j) or (i == k == j):
            return self._two_termC(v, i, k, l, j)
        raise RuntimeError("Unexpected index arrangement for V_ijjj term")

    def _two_termC(self, v, i, k, l, j):
        o = self.o
        if i == k == l:
            s, p = i, j
        elif k == l == j:
            s, p = j, i
        elif i == l == j:
            s, p = i, k
        elif i == k == j:
            s, p = i, l
        else:
            raise RuntimeError("Term generation logic has failed for V_ijjj term")
        if i == k:
            return [tm_positional_two_term(True, math.sqrt(2.0) * v, s, p,
                                           o.b_create_count["couple_down"], o.b_create_count["fill_couple_up"],
                                           o.b_destroy["couple_down"], o.b_destroy["fill_couple_up"])]
        return [tm_positional_two_term(True, -math.sqrt(2.0) * v, s, p,
                                       o.b_destroy_count["couple_down"], o.b_destroy_count["fill_couple_up"],
                                       o.b_create["couple_down"], o.b_create["fill_couple_up"])]

    def three_term(self, v, i, k, l, j):
        if i == j or k == l:
            return self._three_termA(v, i, k, l, j)
        return self._three_termB(v, i, k, l, j)

    def _three_termA(self, v, i, k, l, j):
        o = self.o
        if i == j:
            same = i
        if k == l:
            same = k
            k, l = i, j
        return [tm_three_term(math.sqrt(2.0) * v, same, k, l, o.count, o.count_fill,
                              o.b_create["couple_down"], o.b_create["fill_couple_up"],
                              o.b_destroy["couple_down"], o.b_destroy["fill_couple_up"])]

    def _three_termB(self, v, i, k, l, j):
        o = self.o
        ret = []
        cr, de, fl = o.b_create, o.b_destroy, o.b_flip
        if i == k:
            same, p1, p2 = i, min(l, j), max(l, j)
            ret.append(tm_three_term(-math.sqrt(2.0) * v, same, p1, p2, o.e2d, o.e2d,
                                     de["couple_down"], de["fill_couple_up"], de["couple_down"], de["fill_couple_up"]))
        if j == l:
            same, p1, p2 = j, min(i, k), max(i, k)
            ret.append(tm_three_term(-math.sqrt(2.0) * v, same, p1, p2, o.d2e, o.d2e,
                                     cr["couple_down"], cr["fill_couple_up"], cr["couple_down"], cr["fill_couple_up"]))
        if j == k or i == l:
            if j == k:
                same, p1, p2 = j, l, i
            else:
                same, p1, p2 = i, j, k
            phase = -1.0 if _sgn(i, k, l, j) else 1.0
            r3 = math.sqrt(3.0)
            if same < min(p1, p2):
                ret.append(tm_three_term(phase * r3 * v, same, p1, p2, fl["couple_up"], fl["couple_up"],
                                         cr["couple_down"], cr["fill_couple_down"], de["couple_down"], de["fill_couple_down"]))
                ret.append(tm_three_term(-0.5 * math.sqrt(2.0) * v, same, p1, p2, o.count, o.count,
                                         cr["couple_down"], cr["fill_couple_up"], de["couple_down"], de["fill_couple_up"]))
            elif same > max(p1, p2):
                ret.append(tm_three_term(phase * r3 * v, same, p1, p2, fl["couple_down"], fl["couple_down"],
                                         cr["couple_up"], cr["fill_couple_up"], de["couple_up"], de["fill_couple_up"]))
                ret.append(tm_three_term(-0.5 * math.sqrt(2.0) * v, same, p1, p2, o.count, o.count,
                                         cr["couple_down"], cr["fill_couple_up"], de["couple_down"], de["fill_couple_up"]))
            else:
                ret.append(tm_three_term(phase * r3 * v, same, p1, p2, fl["no_couple"], fl["no_couple"],
                                         cr["couple_down"], cr["fill_couple_up"], de["couple_down"], de["fill_couple_up"]))
                ret.append(tm_three_term(-0.5 * math.sqrt(2.0) * v, same, p1, p2, o.count_fill, o.count_fill,
                                         cr["couple_down"], cr["fill_couple_up"], de["couple_down"], de["fill_couple_up"]))
        return ret

    def four_term(self, v, i, k, l, j):
        o = self.o
        i_, j_, k_, l_ = align(i, j, k, l)
        ret = []
        if k_ > l_ > j_:
            ret.append(tm_four_term(2, -math.sqrt(3.0) * v, i, k, l, j, o.b_create, o.b_destroy))
            ret.append(tm_four_term(1, v, i, k, l, j, o.b_create, o.b_destroy))
        elif k_ > j_ > l_:
            le = -v if _sgn(i, k, l, j) else v
            ret.append(tm_four_term(2, math.sqrt(3.0) * le, i, k, l, j, o.b_create, o.b_destroy))
            ret.append(tm_four_term(1, le, i, k, l, j, o.b_create, o.b_destroy))
        elif j_ > k_ > l_:
            ret.append(tm_four_term(1, 2.0 * v, i, k, l, j, o.b_create, o.b_destroy))
        else:
            raise RuntimeError("unexpected index arrangment in V_ijkl term")
        return ret


class _Merger:
    """ChemHelperSU2: collapse terms with identical operators (chem_helper.h:47-95)."""

    def __init__(self):
        self.two, self.three, self.four = {}, {}, {}

    def _add(self, d, key, term):
        if key in d:
            d[key].coeff += term.coeff
        else:
            d[key] = term

    def add_2term(self, t):
        p, g = t.positions(), t.tags()
        self._add(self.two, (p[0], p[1], g[0], g[1]), t)

    def add_3term(self, t):
        p, g = t.positions(), t.tags()
        self._add(self.three, (p[0], p[1], p[2], g[0], g[1], g[2]), t)

    def add_4term(self, t):
        p, g = t.positions(), t.tags()
        self._add(self.four, tuple(p) + tuple(g), t)

    def commit(self, out):
        for k in sorted(self.two):
            out.append(self.two[k])
        for k in sorted(self.three):
            out.append(self.three[k])
        for k in sorted(self.four):
            out.append(self.four[k])


def su2_terms(integrals, L, N, ops):
    """Hamiltonian terms from aligned 0-based integrals [((i,j,k,l), v)], -1 = absent.

    Follows qc_su2::create_terms case by case (model.hpp:60-295).
    """
    terms = []
    mg = _Merger()
    ss = SpinSumSU2(ops)
    for (i, j, k, l), v in integrals:
        if i == -1 and j == -1 and k == -1 and l == -1:
            terms.append(Term(v, False, [(0, ops.ident)]))
        elif i == j and k == -1 and l == -1:
            terms.append(Term(v, False, [(i, ops.count)]))
        elif k == -1 and l == -1:
            if N == 1:
                r2 = math.sqrt(2.0) * v
                terms.append(tm_positional_two_term(True, r2, j, i, ops.create, ops.create_fill,
                                                    ops.destroy, ops.destroy_fill))
                terms.append(tm_positional_two_term(True, r2, i, j, ops.create, ops.create_fill,
                                                    ops.destroy, ops.destroy_fill))
            else:
                f = v * (1.0 / (N - 1))
                tt = []
                for kk in range(L):
                    if kk == j or kk == i:
                        continue
                    tt += ss.three_term(f, i, kk, kk, j)
                    tt += ss.three_term(f, j, kk, kk, i)
                for t in tt:
                    mg.add_3term(t)
                tt = []
                tt += ss.V_term(f, i, i, i, j)
                tt += ss.V_term(f, j, i, i, i)
                tt += ss.V_term(f, i, j, j, j)
                tt += ss.V_term(f, j, j, j, i)
                for t in tt:
                    mg.add_2term(t)
        elif i == j == k == l:
            terms.append(Term(v, False, [(i, ops.docc)]))
        elif (i == j == k and k != l) or (i != j and j == k == l):
            if i == j:
                s, p = i, l
            else:
                s, p = l, i
            for t in ss.two_term(v, s, s, s, p) + ss.two_term(v, s, p, s, s):
                mg.add_2term(t)
        elif i == j and k == l and j != k:
            for t in ss.two_term(v, i, k, k, i):
                mg.add_2term(t)
        elif i == k and j == l and i != j:
            tt = ss.two_term(0.5 * v, i, i, j, j) + ss.two_term(0.5 * v, j, j, i, i) + ss.two_term(v, i, j, i, j)
            for t in tt:
                mg.add_2term(t)
        elif (i == j and j != k and k != l) or (k == l and i != j and j != k):
            if i == j:
                tt = ss.three_term(v, i, k, l, i) + ss.three_term(v, i, l, k, i)
            else:
                tt = ss.three_term(v, i, k, k, j) + ss.three_term(v, j, k, k, i)
            for t in tt:
                mg.add_3term(t)
        elif ((i == k and j != l) or j == k or (j == l and i != k)) and (i != j and k != l):
            tt = (ss.three_term(v, i, k, l, j) + ss.three_term(v, i, l, k, j)
                  + ss.three_term(v, j, k, l, i) + ss.three_term(v, j, l, k, i))
            for t in tt:
                mg.add_3term(t)
        elif i != j and j != k and k != l and i != k and j != l:
            tt = (ss.four_term(v, i, k, l, j) + ss.four_term(v, i, l, k, j)
                  + ss.four_term(v, j, k, l, i) + ss.four_term(v, j, l, k, i))
            for t in tt:
                mg.add_4term(t)
    mg.commit(terms)
    return terms
