"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py, never by the package.  `OracleEngine` implements the same engine
interface as the GPU engine (qcmaquis_amd.engine.GpuEngine) so the sweep
driver can run on either, and tests compare them call by call.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

from qcmaquis_amd.desc import BoundaryLayout, SiteDesc, f64p, i32p, i64p  # noqa: E402

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_last_error.restype = C.c_char_p
        L.orc_blas_name.restype = C.c_char_p
        L.orc_wigner9j.restype = C.c_double
        L.orc_wigner9j.argtypes = [C.c_int] * 9
        L.orc_wigner6j.restype = C.c_double
        L.orc_wigner6j.argtypes = [C.c_int] * 6
        L.orc_set_coupling.argtypes = [C.c_int] * 9 + [C.c_double, f64p]
        L.orc_vec_len.restype = C.c_int64
        L.orc_sigma.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, C.c_int, f64p]
        L.orc_left_step.argtypes = [C.c_void_p, f64p, f64p, C.POINTER(C.c_void_p), f64p]
        L.orc_right_step.argtypes = [C.c_void_p, f64p, f64p, C.POINTER(C.c_void_p), f64p]
        L.orc_result_info.argtypes = [C.c_void_p, i32p, i32p, i64p]
        L.orc_result_copy.argtypes = [C.c_void_p, i32p, i32p, i32p, i32p, i32p, i64p, f64p]
        L.orc_result_free.argtypes = [C.c_void_p]
        L.orc_jacobi_davidson.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, f64p, C.c_int, C.c_double,
                                          C.POINTER(f64p), C.c_int, C.POINTER(C.c_int), f64p]
        L.orc_davidson.argtypes = L.orc_jacobi_davidson.argtypes
        L.orc_diagonal.argtypes = [C.c_void_p, f64p, f64p, f64p]
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(f64p)


class HostBoundary:
    __slots__ = ("layout", "data")

    def __init__(self, layout: BoundaryLayout, data: np.ndarray):
        self.layout = layout
        self.data = np.ascontiguousarray(data, dtype=np.float64)

    def dicts(self):
        return self.layout.to_dicts(self.data)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


class OracleEngine:
    """Engine interface on the CPU oracle (FP64, reference schedule)."""

    name = "oracle"

    def __init__(self, threads=None):
        if threads:
            lib().orc_set_num_threads(int(threads))
        self.flops = 0.0

    def seed_boundary(self, charge):
        return HostBoundary(BoundaryLayout.seed(charge), np.ones(1))

    def _result(self, h):
        L = lib()
        aux, nb, tot = C.c_int32(), C.c_int32(), C.c_int64()
        L.orc_result_info(h, C.byref(aux), C.byref(nb), C.byref(tot))
        n = max(nb.value, 1)
        bptr = np.zeros(aux.value + 1, np.int32)
        lc = np.zeros(2 * n, np.int32)
        rc = np.zeros(2 * n, np.int32)
        rows = np.zeros(n, np.int32)
        cols = np.zeros(n, np.int32)
        off = np.zeros(n, np.int64)
        data = np.zeros(max(tot.value, 1))
        L.orc_result_copy(h, bptr.ctypes.data_as(i32p), lc.ctypes.data_as(i32p), rc.ctypes.data_as(i32p),
                          rows.ctypes.data_as(i32p), cols.ctypes.data_as(i32p), off.ctypes.data_as(i64p), _fp(data))
        L.orc_result_free(h)
        nbv = nb.value
        lay = BoundaryLayout(aux.value, bptr, lc[:2 * nbv], rc[:2 * nbv], rows[:nbv], cols[:nbv], off[:nbv], tot.value)
        return HostBoundary(lay, data[:max(tot.value, 1)])

    def left_step(self, symm, A, L: HostBoundary, W):
        x = A.to_flat()
        sd = SiteDesc(symm, A.phys, A.left, A.right, W, L.layout, None)
        h = C.c_void_p()
        fl = C.c_double()
        _check(lib().orc_left_step(sd.ptr(), _fp(L.data), _fp(x), C.byref(h), C.byref(fl)))
        self.flops += fl.value
        return self._result(h)

    def right_step(self, symm, A, R: HostBoundary, W):
        x = A.to_flat()
        sd = SiteDesc(symm, A.phys, A.left, A.right, W, None, R.layout)
        h = C.c_void_p()
        fl = C.c_double()
        _check(lib().orc_right_step(sd.ptr(), _fp(R.data), _fp(x), C.byref(h), C.byref(fl)))
        self.flops += fl.value
        return self._result(h)

    def sigma(self, symm, psi, L, R, W, x=None, schedule=-1):
        sd = SiteDesc(symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout)
        if x is None:
            x = psi.to_flat()
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        fl = C.c_double()
        _check(lib().orc_sigma(sd.ptr(), _fp(L.data), _fp(R.data), _fp(x), _fp(y), schedule, C.byref(fl)))
        self.flops += fl.value
        return y, fl.value

    def _solve(self, fn, symm, psi, L, R, W, max_iter, tol, ortho, x0):
        sd = SiteDesc(symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout)
        x0 = psi.to_flat() if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        u = np.zeros_like(x0)
        th = C.c_double()
        its = C.c_int()
        fl = C.c_double()
        orth = [np.ascontiguousarray(o, dtype=np.float64) for o in ortho]
        arr = (f64p * max(1, len(orth)))(*[_fp(o) for o in orth])
        _check(fn(sd.ptr(), _fp(L.data), _fp(R.data), _fp(x0), _fp(u), C.byref(th),
                  max_iter, tol, arr, len(orth), C.byref(its), C.byref(fl)))
        self.flops += fl.value
        return th.value, u, its.value

    def jd(self, symm, psi, L, R, W, max_iter=10, tol=1e-8, ortho=(), x0=None):
        return self._solve(lib().orc_jacobi_davidson, symm, psi, L, R, W, max_iter, tol, ortho, x0)

    def davidson(self, symm, psi, L, R, W, max_iter=10, tol=1e-8, ortho=(), x0=None):
        """Davidson with the diagonal preconditioner (SU2 only), optimize/davidson.h:64-161."""
        return self._solve(lib().orc_davidson, symm, psi, L, R, W, max_iter, tol, ortho, x0)

    def block_solve(self, symm, psis, L, R, W, max_iter=10, tol=1e-8):
        """The len(psis) lowest eigenpairs as the reference finds excited states: one
        Jacobi-Davidson solve per root, each constrained orthogonal to the roots
        before it (SingleSiteVS ortho vectors)."""
        thetas, us, its = [], [], 0
        for psi in psis:
            th, u, it = self.jd(symm, psi, L, R, W, max_iter, tol, ortho=us)
            thetas.append(th)
            us.append(u)
            its += it
        return thetas, us, its

    def diagonal(self, symm, psi, L, R, W):
        """Diagonal of H_eff, non-abelian/h_diag.hpp:120-152, flat left-paired layout."""
        sd = SiteDesc(symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout)
        out = np.zeros(psi.to_flat().size)
        _check(lib().orc_diagonal(sd.ptr(), _fp(L.data), _fp(R.data), _fp(out)))
        return out


def wigner9j(*a):
    return lib().orc_wigner9j(*a)


def wigner6j(*a):
    return lib().orc_wigner6j(*a)


def blas_name():
    return lib().orc_blas_name().decode()
