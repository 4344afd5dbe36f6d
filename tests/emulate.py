"""numpy executor of the device task lists (test infrastructure).

Runs a HostPlan's copy / GEMM / W-application tasks on host arrays exactly
as the HIP kernels define them, so the plan logic (offsets, transposes,
coefficients, tile and split-K coverage) is checked against the oracle on
machines without a GPU.
"""
import numpy as np
from numpy.lib.stride_tricks import as_strided


def _mat(buf, off, rows, cols, ld):
    if rows == 0 or cols == 0:
        return np.zeros((rows, cols))
    need = off + (rows - 1) + (cols - 1) * ld + 1
    assert off >= 0 and need <= buf.size, (off, rows, cols, ld, buf.size)
    return as_strided(buf[off:], shape=(rows, cols), strides=(8, 8 * ld))


def copy2d(tasks, src, dst):
    for t in tasks:
        r, c = int(t["rows"]), int(t["cols"])
        if t["trans"]:
            s = _mat(src, int(t["src"]), c, r, int(t["lds"])).T
        else:
            s = _mat(src, int(t["src"]), r, c, int(t["lds"]))
        d = _mat(dst, int(t["dst"]), r, c, int(t["ldd"]))
        d[...] = t["scale"] * s


def wapply(strips, srcs, T, Cbuf):
    for st in strips:
        r, c = int(st["rows"]), int(st["cols"])
        acc = np.zeros((r, c))
        for k in range(int(st["s0"]), int(st["s1"])):
            s = srcs[k]
            acc += s["coef"] * _mat(T, int(s["off"]), r, c, int(s["ld"]))
        _mat(Cbuf, int(st["dst"]), r, c, int(st["ldd"]))[...] = acc


def gemm(groups, tiles, A, B, C, rowmap=None):
    # tile coverage: every (m, n, k-split) tile exactly once (g < 0: padding tiles)
    seen = {}
    for t in tiles:
        if t["g"] < 0:
            continue
        key = (int(t["g"]), int(t["mt"]), int(t["nt"]), int(t["ks"]))
        assert key not in seen
        seen[key] = 1
    assert len(seen) == sum(1 for t in tiles if t["g"] >= 0)
    for gi, g in enumerate(groups):
        M, N, K = int(g["M"]), int(g["N"]), int(g["K"])
        tm, tn = int(g["tm"]), int(g["tn"])
        assert tm in (64, 128) and tn in range(16, 129, 16)
        mt, nt = -(-M // tm), -(-N // tn)
        for ks in range(int(g["split"])):
            for n in range(nt):
                for m in range(mt):
                    assert (gi, m, n, ks) in seen, ("tile not in the list", gi, m, n, ks, M, N, K, tm, tn)
        assert int(g["split"]) * int(g["kchunk"]) >= K
        if g["transA"]:
            a = _mat(A, int(g["a"]), K, M, int(g["lda"])).T
        else:
            a = _mat(A, int(g["a"]), M, K, int(g["lda"]))
        if g["transB"]:
            b = _mat(B, int(g["b"]), N, K, int(g["ldb"])).T
        else:
            b = _mat(B, int(g["b"]), K, N, int(g["ldb"]))
        prod = a @ b if K else np.zeros((M, N))
        if int(g["rowmap"]) > 0:
            # row r at c + map[rowmap - 1 + r].off + col * map[...].ld; the rows must not overlap
            assert int(g["split"]) == 1
            rm = rowmap[int(g["rowmap"]) - 1:int(g["rowmap"]) - 1 + M]
            assert len(rm) == M
            used = set()
            for r in range(M):
                addr = int(g["c"]) + int(rm["off"][r]) + np.arange(N) * int(rm["ld"][r])
                assert addr.min() >= 0 and addr.max() < C.size
                assert not used.intersection(addr.tolist())
                used.update(addr.tolist())
                C[addr] = prod[r]
        else:
            _mat(C, int(g["c"]), M, N, int(g["ldc"]))[...] = prod


def run_sigma(hp, Ldata, Rdata, x):
    s = hp.sizes
    P = hp.parts
    la, rb, prp = np.zeros(max(s["la"], 1)), np.zeros(max(s["rb"], 1)), np.zeros(max(s["prp"], 1))
    T, Cb = np.zeros(max(s["t"], 1)), np.zeros(max(s["c"], 1))
    y = np.full(s["vec"], np.nan)
    if getattr(hp, "rbtm", False):
        # rbtm: t = psi RC, BUF = W applied to t, S = LA^T BUF, y = S right -> left paired
        tmp = np.full(max(s["tmp"], 1), np.nan)
        copy2d(P["arrL"], Ldata, la)
        copy2d(P["arrR"], Rdata, rb)
        copy2d(P["reshape"], x, prp)
        gemm(P["g1_groups"], P["g1_tiles"], prp, rb, T, P["g1_rowmap"])
        wapply(P["strips"], P["srcs"], T, Cb)
        gemm(P["g2_groups"], P["g2_tiles"], la, Cb, tmp)
        copy2d(P["scatter"], tmp, y)
        assert not np.isnan(y).any(), "sigma element not written"
        return y
    copy2d(P["arrL"], Ldata, la)
    wapply(P["comb_strips"], P["comb_srcs"], la, la)
    copy2d(P["arrR"], Rdata, rb)
    copy2d(P["reshape"], x, prp)
    gemm(P["g1_groups"], P["g1_tiles"], la, prp, T, P["g1_rowmap"])
    wapply(P["strips"], P["srcs"], T, Cb)
    gemm(P["g2_groups"], P["g2_tiles"], Cb, rb, y)
    return y


def run_step(hp, kind, Bdata, A):
    s = hp.sizes
    P = hp.parts
    la, rb, prp = np.zeros(max(s["la"], 1)), np.zeros(max(s["rb"], 1)), np.zeros(max(s["prp"], 1))
    T, Cb, tmp = np.zeros(max(s["t"], 1)), np.zeros(max(s["c"], 1)), np.zeros(max(s["tmp"], 1))
    out = np.full(max(s["out"], 1), np.nan)
    copy2d(P["reshape"], A, prp)
    if kind == "lstep":
        copy2d(P["arrL"], Bdata, la)
        wapply(P["comb_strips"], P["comb_srcs"], la, la)
        gemm(P["g1_groups"], P["g1_tiles"], la, prp, T, P["g1_rowmap"])
        wapply(P["strips"], P["srcs"], T, Cb)
        gemm(P["g2_groups"], P["g2_tiles"], Cb, A, tmp)
    else:
        copy2d(P["arrR"], Bdata, rb)
        gemm(P["g1_groups"], P["g1_tiles"], A, rb, T, P["g1_rowmap"])
        wapply(P["strips"], P["srcs"], T, Cb)
        gemm(P["g2_groups"], P["g2_tiles"], Cb, prp, tmp)
    copy2d(P["scatter"], tmp, out)
    return out[:s["out"]]


def run_diag(hp, Ldata, Rdata):
    """Diagonal plan: L/R diagonals, strips into T (zero where no strip writes),
    then diag slice = T_slice * RB_rc; every element of the result must be written."""
    s = hp.sizes
    P = hp.parts
    la, rb, T = np.zeros(max(s["la"], 1)), np.zeros(max(s["rb"], 1)), np.zeros(max(s["t"], 1))
    out = np.full(s["vec"], np.nan)
    copy2d(P["arrL"], Ldata, la)
    copy2d(P["arrR"], Rdata, rb)
    wapply(P["strips"], P["srcs"], la, T)
    gemm(P["g2_groups"], P["g2_tiles"], T, rb, out)
    assert not np.isnan(out).any(), "diagonal element not written"
    return out


def jacobi_svd(M, max_sweeps=60):
    """numpy executor of the device split's block SVD (k_jacobi_init,
    k_jacobi_round, k_jacobi_sweep_end, k_jacobi_finish and k_split_assemble): the working matrix X is M (rows >= cols) or M^T, its
    columns are rotated pairwise in the kernel's round-robin order until a sweep
    rotates nothing, and the triplets come out sorted by descending singular
    value with ties broken by column.  Returns (U, s, Vt, sweeps) with
    M = U diag(s) Vt."""
    rows, cols = M.shape
    trans = cols > rows
    X = (M.T if trans else M).astype(np.float64).copy()
    R, C = X.shape
    V = np.eye(C)
    Ce = C + (C & 1)
    tol = 2.220446049250313e-16 * np.sqrt(R)
    sweeps, done = 0, C < 2
    while not done and sweeps < max_sweeps:
        rotated = False
        for t in range(Ce - 1):
            for k in range(Ce // 2):
                p, q = (t, Ce - 1) if k == 0 else ((t + k) % (Ce - 1), (t - k + Ce - 1) % (Ce - 1))
                p, q = min(p, q), max(p, q)
                if q >= C:
                    continue
                a, b, g = X[:, p] @ X[:, p], X[:, q] @ X[:, q], X[:, p] @ X[:, q]
                if abs(g) > tol * np.sqrt(a * b):
                    zeta = (b - a) / (2.0 * g)
                    tt = np.copysign(1.0, zeta) / (abs(zeta) + np.sqrt(1.0 + zeta * zeta))
                    if tt != 0.0:
                        cs = 1.0 / np.sqrt(1.0 + tt * tt)
                        sn = cs * tt
                        X[:, [p, q]] = X[:, [p, q]] @ np.array([[cs, sn], [-sn, cs]])
                        V[:, [p, q]] = V[:, [p, q]] @ np.array([[cs, sn], [-sn, cs]])
                        rotated = True
        sweeps += 1
        done = not rotated
    sig = np.sqrt(np.sum(X * X, axis=0))
    order = sorted(range(C), key=lambda j: (-sig[j], j))
    s = sig[order]
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    if trans:
        U, Vt = V[:, order], (X[:, order] * inv).T
    else:
        U, Vt = X[:, order] * inv, V[:, order].T
    return U, s, Vt, (sweeps if done else -1)
