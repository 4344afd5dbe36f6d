"""MPS site tensors, pairing reshapes and the two-site glue of a TS sweep.

Layouts follow the reference exactly:
  left-paired  block (c, c): rows (s, l) at ProductBasis(phys, left)(s, l) + ss*ldim + ll
  right-paired block (l, l): cols (s, r) at ProductBasis(phys, right, -)(s, r) + ss*rdim + rr
(mp_tensors/reshapes.h:174-336, block_matrix/detail/alps_detail.hpp:129-152).
The canonical flat vector of a tensor lists its left-paired blocks for every
charge c of the right index present in phys*left, descending in c.

Two-site helpers restate mp_tensors/twositetensor.hpp:23-181,
mp_tensors/ts_reshape.h and the SU(2) recoupling mp_tensors/ts_reduction.h.
Truncation restates block_matrix/block_matrix_algorithms.h:211-333.
"""
from __future__ import annotations

import math

import numpy as np

from .blocks import BlockMatrix, Index, ProductBasis, fuse, neg, gemm
from .wigner import sixj


class MPSTensor:
    __slots__ = ("phys", "left", "right", "data", "pairing")

    def __init__(self, phys: Index, left: Index, right: Index, data: BlockMatrix, pairing="L"):
        self.phys, self.left, self.right = phys, left, right
        self.data = data
        self.pairing = pairing

    def copy(self):
        return MPSTensor(self.phys, self.left, self.right, self.data.copy(), self.pairing)

    def make_left_paired(self):
        if self.pairing == "R":
            self.data = reshape_r2l(self.phys, self.left, self.right, self.data)
            self.pairing = "L"
        return self

    def make_right_paired(self):
        if self.pairing == "L":
            self.data = reshape_l2r(self.phys, self.left, self.right, self.data)
            self.pairing = "R"
        return self

    def norm(self):
        return self.data.norm()

    def scale_(self, a):
        for k in self.data.blocks:
            self.data.blocks[k] *= a
        return self

    # canonical flat layout -------------------------------------------------
    def layout(self):
        return canonical_layout(self.phys, self.left, self.right)

    def to_flat(self):
        self.make_left_paired()
        lay, n = self.layout()
        x = np.zeros(n)
        for c, rows, cols, off in lay:
            m = self.data.blocks.get((c, c))
            if m is not None:
                if m.shape != (rows, cols):
                    raise ValueError(f"block {c} shape {m.shape} != {(rows, cols)}")
                x[off:off + rows * cols] = m.ravel(order="F")
        return x

    @staticmethod
    def from_flat(phys, left, right, x):
        lay, n = canonical_layout(phys, left, right)
        if len(x) != n:
            raise ValueError("flat vector length mismatch")
        bm = BlockMatrix()
        for c, rows, cols, off in lay:
            bm.blocks[(c, c)] = np.array(x[off:off + rows * cols]).reshape((rows, cols), order="F")
        return MPSTensor(phys, left, right, bm, "L")


def canonical_layout(phys, left, right):
    pb = ProductBasis(phys, left)
    lay = []
    off = 0
    for c, cols in right:
        rows = pb.size.get(c)
        if rows is None:
            continue
        lay.append((c, rows, cols, off))
        off += rows * cols
    return lay, off


def reshape_l2r(phys, left, right, m1: BlockMatrix) -> BlockMatrix:
    """[(phys,left), right] -> [left, (-phys,right)] (reshapes.h:174-230)."""
    in_left = ProductBasis(phys, left)
    out_right = ProductBasis(phys, right, right=True)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        if not right.has(rc):
            continue
        rdim = right.size_of_block(rc)
        for s, sdim in phys:
            lch = fuse(lc, neg(s))
            if not left.has(lch):
                continue
            ldim = left.size_of_block(lch)
            key = (lch, fuse(neg(s), rc))
            out = m2.blocks.get(key)
            if out is None:
                out = np.zeros((ldim, out_right.size[key[1]]))
                m2.blocks[key] = out
            ilo = in_left(s, lch)
            oro = out_right(s, rc)
            for ss in range(sdim):
                out[:, oro + ss * rdim: oro + (ss + 1) * rdim] = blk[ilo + ss * ldim: ilo + (ss + 1) * ldim, :]
    return m2


def reshape_r2l(phys, left, right, m1: BlockMatrix) -> BlockMatrix:
    """[left, (-phys,right)] -> [(phys,left), right] (reshapes.h:290-336)."""
    in_right = ProductBasis(phys, right, right=True)
    out_left = ProductBasis(phys, left)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        if not left.has(lc):
            continue
        ldim = left.size_of_block(lc)
        for s, sdim in phys:
            rch = fuse(rc, s)
            if not right.has(rch):
                continue
            rdim = right.size_of_block(rch)
            key = (fuse(s, lc), rch)
            out = m2.blocks.get(key)
            if out is None:
                out = np.zeros((out_left.size[key[0]], rdim))
                m2.blocks[key] = out
            iro = in_right(s, rch)
            olo = out_left(s, lc)
            for ss in range(sdim):
                out[olo + ss * ldim: olo + (ss + 1) * ldim, :] = blk[:, iro + ss * rdim: iro + (ss + 1) * rdim]
    return m2


# ---------------------------------------------------------------------------
# initialisation (mp_tensors/mps_sectors.h:24-112, mps_initializers.h:35-81)
# ---------------------------------------------------------------------------

def allowed_sectors(L, phys, right_end, Mmax, su2=True):
    phys_c = sorted(phys.charges())
    cmax_site, cmin_site = phys_c[-1], phys_c[0]
    max_tot = (cmax_site[0] * L, cmax_site[1] * L)
    min_tot = (cmin_site[0] * L, cmin_site[1] * L)
    left = [None] * (L + 1)
    right = [None] * (L + 1)
    left[0] = Index([((0, 0), 1)])
    right[L] = Index([(right_end, 1)])
    cmaxi, cmini = max_tot, min_tot
    for i in range(1, L + 1):
        li = phys * left[i - 1]
        cmaxi = fuse(cmaxi, neg(cmax_site))
        cmini = fuse(cmini, neg(cmin_site))
        d = {}
        for c, s in li:
            if fuse(c, cmaxi) < right_end:
                continue
            if fuse(c, cmini) > right_end:
                continue
            if su2 and c[1] < 0:
                continue
            d[c] = min(Mmax, s)
        left[i] = Index.from_dict(d)
    cmaxi, cmini = max_tot, min_tot
    for i in range(L - 1, -1, -1):
        ri = phys.adjoin() * right[i + 1]
        cmaxi = fuse(cmaxi, neg(cmax_site))
        cmini = fuse(cmini, neg(cmin_site))
        d = {}
        for c, s in ri:
            if fuse(c, neg(cmaxi)) > (0, 0):
                continue
            if fuse(c, neg(cmini)) < (0, 0):
                continue
            if su2 and c[1] < 0:
                continue
            d[c] = min(Mmax, s)
        right[i] = Index.from_dict(d)
        # extract_common_subset(left[i], right[i])
        lset = {c: s for c, s in left[i] if right[i].has(c)}
        rset = {c: s for c, s in right[i] if c in lset}
        left[i] = Index.from_dict(lset)
        right[i] = Index.from_dict(rset)
    allowed = []
    for i in range(L + 1):
        d = {}
        for c, s in left[i]:
            if right[i].has(c):
                d[c] = min(Mmax, s, right[i].size_of_block(c))
        allowed.append(Index.from_dict(d))
    return allowed


def bounded_sectors(L, phys, right_end, m_total, su2=True):
    """Allowed sectors with the total bond dimension of every bond capped at m_total.

    Sector sizes are taken proportional to the number of states the sector
    holds (its combinatorial degeneracy), each at least 1, so a random state
    has the block structure of an m_total state rather than the per-sector
    cap the reference's default initialiser applies."""
    full = allowed_sectors(L, phys, right_end, 1 << 30, su2)
    out = []
    for ix in full:
        tot = ix.sum_of_sizes()
        if tot <= m_total:
            out.append(ix)
            continue
        d = {}
        for c, s in ix:
            d[c] = max(1, int(round(s * m_total / tot)))
        out.append(Index.from_dict(d))
    return out


def random_mps_on(sectors, phys, seed=42):
    rng = np.random.default_rng(seed)
    mps = []
    for i in range(len(sectors) - 1):
        lay, n = canonical_layout(phys, sectors[i], sectors[i + 1])
        t = MPSTensor.from_flat(phys, sectors[i], sectors[i + 1], rng.uniform(0.0, 1.0, size=n))
        nrm = t.norm()
        if nrm > 0:
            t.scale_(1.0 / nrm)
        mps.append(t)
    return mps


def random_mps(L, phys, right_end, Mmax, seed=42, su2=True):
    """Random MPS over the allowed sectors, each site scaled to unit norm.

    The reference draws from boost::mt19937 (utils/random.hpp); this build
    uses numpy's PCG64 seeded with the same integer, so sectors and sizes
    match but the numbers do not."""
    rng = np.random.default_rng(seed)
    allowed = allowed_sectors(L, phys, right_end, Mmax, su2)
    mps = []
    for i in range(L):
        lay, n = canonical_layout(phys, allowed[i], allowed[i + 1])
        x = rng.uniform(0.0, 1.0, size=n)
        t = MPSTensor.from_flat(phys, allowed[i], allowed[i + 1], x)
        nrm = t.norm()
        if nrm > 0:
            t.scale_(1.0 / nrm)
        mps.append(t)
    return mps


# ---------------------------------------------------------------------------
# normalisation (mps.hpp:103-160, mpstensor.hpp:219-300, block_matrix_algorithms.h:443-490)
# ---------------------------------------------------------------------------

def _sign_fix_qr(q, r):
    # LAPACK's Householder QR leaves R's diagonal signs free; keep them as numpy returns.
    return q, r


def left_normalize_and_return(t: MPSTensor):
    """Thin QR of the left-paired tensor; returns R (block matrix)."""
    t.make_left_paired()
    Q = BlockMatrix()
    R = BlockMatrix()
    newright = {}
    for (lc, rc), m in t.data.blocks.items():
        q, r = np.linalg.qr(m, mode="reduced")
        Q.blocks[(lc, rc)] = q
        R.blocks[(rc, rc)] = r
        newright[rc] = q.shape[1]
    t.data = Q
    t.right = Index.from_dict(newright)
    return R


def right_normalize_and_return(t: MPSTensor):
    """Thin LQ of the right-paired tensor; returns L (block matrix)."""
    t.make_right_paired()
    Qb = BlockMatrix()
    Lb = BlockMatrix()
    newleft = {}
    for (lc, rc), m in t.data.blocks.items():
        q, r = np.linalg.qr(m.T, mode="reduced")
        Qb.blocks[(lc, rc)] = q.T
        Lb.blocks[(lc, lc)] = r.T
        newleft[lc] = q.shape[1]
    t.data = Qb
    t.left = Index.from_dict(newleft)
    return Lb


def multiply_from_left(t: MPSTensor, N: BlockMatrix):
    t.make_right_paired()
    t.data = gemm(N, t.data)
    t.left = N.left_basis()


def multiply_from_right(t: MPSTensor, N: BlockMatrix):
    t.make_left_paired()
    t.data = gemm(t.data, N)
    t.right = N.right_basis()


def normalize_right(mps):
    L = len(mps)
    for i in range(L - 1, 0, -1):
        f = right_normalize_and_return(mps[i])
        multiply_from_right(mps[i - 1], f)
    f = right_normalize_and_return(mps[0])
    tr = sum(float(np.trace(m)) for m in f.blocks.values())
    if tr < 0:
        mps[0].scale_(-1.0)


def normalize_left(mps):
    L = len(mps)
    for i in range(0, L - 1):
        f = left_normalize_and_return(mps[i])
        multiply_from_left(mps[i + 1], f)
    f = left_normalize_and_return(mps[L - 1])
    tr = sum(float(np.trace(m)) for m in f.blocks.values())
    if tr < 0:
        mps[0].scale_(-1.0)


# ---------------------------------------------------------------------------
# two-site tensor
# ---------------------------------------------------------------------------

def ts_both(m1: MPSTensor, m2: MPSTensor) -> BlockMatrix:
    """Both-paired two-site data (twositetensor.hpp:23-30)."""
    m1.make_left_paired()
    m2.make_right_paired()
    return gemm(m1.data, m2.data)


def ts_both_to_right(p1, p2, left, right, m1: BlockMatrix) -> BlockMatrix:
    """ts_reshape.h:230-285 reshape_both_to_right."""
    phys2 = p1 * p2
    phys_pb = ProductBasis(p1, p2)
    in_left = ProductBasis(p1, left)
    in_right = ProductBasis(p2, right, right=True)
    out_right = ProductBasis(phys2, right, right=True)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        for s1, sd1 in p1:
            lch = fuse(lc, neg(s1))
            if not left.has(lch):
                continue
            ldim = left.size_of_block(lch)
            for s2, sd2 in p2:
                rch = fuse(rc, s2)
                if not right.has(rch):
                    continue
                rdim = right.size_of_block(rch)
                sc = fuse(s1, s2)
                key = (lch, fuse(neg(sc), rch))
                out = m2.blocks.get(key)
                if out is None:
                    out = np.zeros((ldim, out_right.size[key[1]]))
                    m2.blocks[key] = out
                ilo, iro = in_left(s1, lch), in_right(s2, rch)
                oro, opo = out_right(sc, rch), phys_pb(s1, s2)
                for ss1 in range(sd1):
                    for ss2 in range(sd2):
                        so = opo + ss1 * sd2 + ss2
                        out[:, oro + so * rdim: oro + (so + 1) * rdim] = \
                            blk[ilo + ss1 * ldim: ilo + (ss1 + 1) * ldim, iro + ss2 * rdim: iro + (ss2 + 1) * rdim]
    return m2


def ts_left_to_both(p1, p2, left, right, m1: BlockMatrix) -> BlockMatrix:
    """ts_reshape.h:88-145 reshape_left_to_both."""
    phys2 = p1 * p2
    phys_pb = ProductBasis(p1, p2)
    in_left = ProductBasis(phys2, left)
    out_right = ProductBasis(p2, right, right=True)
    out_left = ProductBasis(p1, left)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        if not right.has(rc):
            raise RuntimeError("m1 matrix inconsistent with right_i.")
        rdim = right.size_of_block(rc)
        for s1, sd1 in p1:
            for s2, sd2 in p2:
                sc = fuse(s1, s2)
                lch = fuse(lc, neg(sc))
                if not left.has(lch):
                    continue
                ldim = left.size_of_block(lch)
                key = (fuse(s1, lch), fuse(neg(s2), rc))
                out = m2.blocks.get(key)
                if out is None:
                    out = np.zeros((out_left.size[key[0]], out_right.size[key[1]]))
                    m2.blocks[key] = out
                ilo, ipo = in_left(sc, lch), phys_pb(s1, s2)
                olo, oro = out_left(s1, lch), out_right(s2, rc)
                for ss1 in range(sd1):
                    for ss2 in range(sd2):
                        so = ipo + ss1 * sd2 + ss2
                        out[olo + ss1 * ldim: olo + (ss1 + 1) * ldim, oro + ss2 * rdim: oro + (ss2 + 1) * rdim] = \
                            blk[ilo + so * ldim: ilo + (so + 1) * ldim, :]
    return m2


def _recouple_coeff(jl, jr, j, S1, S2, jm):
    c = math.sqrt((j + 1) * (jm + 1)) * sixj(jl, jr, j, S2, S1, jm)
    if ((jl + jr + S1 + S2) // 2) % 2:
        c = -c
    return c


def reduce_right(p1, p2, left, right, m1: BlockMatrix) -> BlockMatrix:
    """SU(2) two-site reduction of a right-paired tensor (ts_reduction.h:42-149)."""
    phys2 = p1 * p2
    phys_pb = ProductBasis(p1, p2)
    in_right = ProductBasis(phys2, right, right=True)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        out = np.zeros_like(blk)
        m2.blocks[(lc, rc)] = out
        ldim = blk.shape[0]
        for s, _ in phys2:
            rch = fuse(rc, s)
            if not right.has(rch):
                continue
            iro = in_right(s, rch)
            rdim = right.size_of_block(rch)
            for s1, sd1 in p1:
                for s2, sd2 in p2:
                    if s != fuse(s1, s2):
                        continue
                    ipo = phys_pb(s1, s2)
                    S1, S2 = abs(s1[1]), abs(s2[1])
                    jl = lc[1]
                    jm = lc[1] + s1[1]
                    jr = rch[1]
                    if jm < 0:
                        continue
                    if jl == jr and jl > 0 and S1 == 1 and S2 == 1:
                        base = ipo if s1[1] == 1 else ipo - 1
                        for j in range(abs(S1 - S2), abs(S1 + S2) + 1, 2):
                            opo = base + j // 2
                            cc = _recouple_coeff(jl, jr, j, S1, S2, jm)
                            _reduce_r(out, blk, cc, iro, ipo, opo, sd1, sd2, ldim, rdim)
                    else:
                        j = abs(s1[1] + s2[1])
                        cc = _recouple_coeff(jl, jr, j, S1, S2, jm)
                        _reduce_r(out, blk, cc, iro, ipo, ipo, sd1, sd2, ldim, rdim)
    return m2


def _reduce_r(out, inp, scale, iro, ipo, opo, sd1, sd2, ldim, rdim):
    for ss1 in range(sd1):
        for ss2 in range(sd2):
            si = ipo + ss1 * sd2 + ss2
            so = opo + ss1 * sd2 + ss2
            out[:, iro + so * rdim: iro + (so + 1) * rdim] += scale * inp[:, iro + si * rdim: iro + (si + 1) * rdim]


def unreduce_left(p1, p2, left, right, m1: BlockMatrix) -> BlockMatrix:
    """Inverse SU(2) recoupling on a left-paired tensor (ts_reduction.h:151-251)."""
    phys2 = p1 * p2
    phys_pb = ProductBasis(p1, p2)
    in_left = ProductBasis(phys2, left)
    m2 = BlockMatrix()
    for (lc, rc), blk in m1.blocks.items():
        out = np.zeros_like(blk)
        m2.blocks[(lc, rc)] = out
        rdim = blk.shape[1]
        for s, _ in phys2:
            lch = fuse(lc, neg(s))
            if not left.has(lch):
                continue
            ilo = in_left(s, lch)
            ldim = left.size_of_block(lch)
            for s1, sd1 in p1:
                for s2, sd2 in p2:
                    if s != fuse(s1, s2):
                        continue
                    ipo = phys_pb(s1, s2)
                    S1, S2 = abs(s1[1]), abs(s2[1])
                    jl = lch[1]
                    jr = rc[1]
                    if jl == jr and jl > 0 and S1 == 1 and S2 == 1:
                        j = 0 if s1[1] == 1 else 2
                        base = ipo if j == 0 else ipo - 1
                        for jm in (jl - 1, jl + 1):
                            opo = base + 1 if jm == jl - 1 else base
                            cc = _recouple_coeff(jl, jr, j, S1, S2, jm)
                            _reduce_l(out, blk, cc, ilo, ipo, opo, sd1, sd2, ldim)
                    else:
                        j = abs(s1[1] + s2[1])
                        jm = jl + s1[1]
                        if jm < 0:
                            continue
                        cc = _recouple_coeff(jl, jr, j, S1, S2, jm)
                        _reduce_l(out, blk, cc, ilo, ipo, ipo, sd1, sd2, ldim)
    return m2


def _reduce_l(out, inp, scale, ilo, ipo, opo, sd1, sd2, ldim):
    for ss1 in range(sd1):
        for ss2 in range(sd2):
            si = ipo + ss1 * sd2 + ss2
            so = opo + ss1 * sd2 + ss2
            out[ilo + so * ldim: ilo + (so + 1) * ldim, :] += scale * inp[ilo + si * ldim: ilo + (si + 1) * ldim, :]


def make_twosite_mps(m1: MPSTensor, m2: MPSTensor, su2=True) -> MPSTensor:
    """TwoSiteTensor(m1, m2).make_mps() (twositetensor.hpp:126-139), left-paired."""
    p1, p2 = m1.phys, m2.phys
    left, right = m1.left, m2.right
    both = ts_both(m1, m2)
    rp = ts_both_to_right(p1, p2, left, right, both)
    if su2:
        rp = reduce_right(p1, p2, left, right, rp)
    t = MPSTensor(p1 * p2, left, right, rp, "R")
    return t.make_left_paired()


def truncation_keep(svals, Mmax, cutoff):
    """Global truncation rule of estimate_truncation (block_matrix_algorithms.h:
    211-262): the threshold is cutoff * the largest singular value, raised to the
    (Mmax+1)-th largest when there are more than Mmax, and every value at or
    above it is kept (so a bond can keep Mmax + 1 states, as the reference does).
    svals: per block, singular values in descending order.  Returns the
    per-block kept counts and the truncation record."""
    allv = np.sort(np.concatenate(list(svals)))[::-1] if len(svals) else np.zeros(0)
    if allv.size == 0 or allv[0] == 0.0:
        return [0] * len(svals), dict(bond_dimension=0, truncated_weight=0.0, smallest_ev=0.0)
    evalscut = cutoff * allv[0]
    if len(allv) > Mmax:
        evalscut = max(evalscut, allv[Mmax])
    smallest_ev = evalscut / allv[0]
    trunc = allv[allv < evalscut]
    tw = float(np.sum(trunc ** 2) / np.sum(allv ** 2))
    keep = [_keep_count(s, evalscut) for s in svals]
    return keep, dict(bond_dimension=int(sum(keep)), truncated_weight=tw, smallest_ev=smallest_ev)


def svd_truncate(M: BlockMatrix, Mmax, cutoff):
    """Block SVD with global truncation (block_matrix_algorithms.h:211-333)."""
    U, S, V = {}, {}, {}
    for k, m in M.blocks.items():
        u, s, vt = np.linalg.svd(m, full_matrices=False)
        U[k], S[k], V[k] = u, s, vt
    keys = list(M.blocks)
    keep, info = truncation_keep([S[k] for k in keys], Mmax, cutoff)
    Ub, Sb, Vb = BlockMatrix(), {}, BlockMatrix()
    for k, kk in zip(keys, keep):
        if kk == 0:
            continue
        lc, rc = k
        Ub.blocks[(lc, rc)] = U[k][:, :kk]
        Sb[rc] = S[k][:kk]
        Vb.blocks[(lc, rc)] = V[k][:kk, :]
    return Ub, Sb, Vb, info


def _keep_count(s, cut):
    for i, v in enumerate(s):
        if v < cut:
            return i
    return len(s)


def split_twosite(u: MPSTensor, p1, p2, left, right, Mmax, cutoff, l2r=True, su2=True):
    """tst << u; split_mps_l2r / split_mps_r2l (twositetensor.hpp:141-181)."""
    u.make_left_paired()
    lp = u.data
    if su2:
        lp = unreduce_left(p1, p2, left, right, lp)
    both = ts_left_to_both(p1, p2, left, right, lp)
    Ub, Sb, Vb, info = svd_truncate(both, Mmax, cutoff)
    mid = Index.from_dict({rc: len(s) for rc, s in Sb.items()})
    if l2r:
        A = MPSTensor(p1, left, mid, Ub, "L")
        SV = BlockMatrix({(lc, rc): Sb[rc][:, None] * v for (lc, rc), v in Vb.blocks.items()})
        B = MPSTensor(p2, mid, right, SV, "R")
    else:
        B = MPSTensor(p2, mid, right, Vb, "R")
        US = BlockMatrix({(lc, rc): u_ * Sb[rc][None, :] for (lc, rc), u_ in Ub.blocks.items()})
        A = MPSTensor(p1, left, mid, US, "L")
    return A, B, info


def split_twosite_sa(us, p1, p2, left, right, Mmax, cutoff, l2r=True, su2=True):
    """State-averaged split of several two-site tensors sharing one basis.

    The two-site tensors of the roots are laid side by side (l2r: columns, r2l:
    rows) with equal weights 1/sqrt(k); the block SVD of that matrix with the
    global truncation of svd_truncate gives the shared tensor (the leading
    eigenvectors of the averaged density matrix), and each root keeps its own
    projection onto it: l2r returns (A, [B_r]), r2l returns ([A_r], B)."""
    boths = []
    for u in us:
        u.make_left_paired()
        lp = u.data
        if su2:
            lp = unreduce_left(p1, p2, left, right, lp)
        boths.append(ts_left_to_both(p1, p2, left, right, lp))
    keys = list(boths[0].blocks.keys())
    for b in boths[1:]:
        if set(b.blocks.keys()) != set(keys):
            raise ValueError("state-averaged split: the roots have different block structures")
    w = 1.0 / np.sqrt(len(us))
    join = np.hstack if l2r else np.vstack
    stacked = BlockMatrix({k: w * join([b.blocks[k] for b in boths]) for k in keys})
    Ub, Sb, Vb, info = svd_truncate(stacked, Mmax, cutoff)
    mid = Index.from_dict({rc: len(s_) for rc, s_ in Sb.items()})
    if l2r:
        A = MPSTensor(p1, left, mid, Ub, "L")
        Bs = [MPSTensor(p2, mid, right, BlockMatrix({k: Ub.blocks[k].T @ b.blocks[k] for k in Ub.blocks}), "R")
              for b in boths]
        return A, Bs, info
    B = MPSTensor(p2, mid, right, Vb, "R")
    As = [MPSTensor(p1, left, mid, BlockMatrix({k: b.blocks[k] @ Vb.blocks[k].T for k in Vb.blocks}), "L")
          for b in boths]
    return As, B, info
