"""Python binding of libmqgpu (the MI355X hot path) and the GPU engine.

Device memory is held in torch tensors on the current HIP device; the
library launches on its own stream, and this module synchronises that
stream before torch reads results.  There is no CPU fallback: if the
library or a GPU is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .desc import (BoundaryLayout, SiteDesc, SplitDesc, MQ_PREC_FP64, MQ_SYMM_SU2U1, MQ_SYMM_TWOU1, f64p, i32p, i64p,
                   mq_boundary_desc)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_lib", "libmqgpu.so")

_lib = None


class MQError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise MQError(f"libmqgpu.so not built ({LIBPATH}); run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.mq_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mq_ctx_destroy.argtypes = [vp]
        L.mq_last_error.argtypes = [vp]
        L.mq_last_error.restype = C.c_char_p
        L.mq_ctx_stream.argtypes = [vp]
        L.mq_ctx_stream.restype = vp
        L.mq_ctx_synchronize.argtypes = [vp]
        L.mq_ctx_set_option.argtypes = [vp, C.c_int, C.c_int]
        for f in ("mq_sigma_plan_create", "mq_lstep_plan_create", "mq_rstep_plan_create", "mq_diag_plan_create"):
            getattr(L, f).argtypes = [vp, vp, C.POINTER(vp)]
        L.mq_plan_destroy.argtypes = [vp]
        L.mq_plan_vec_len.argtypes = [vp]
        L.mq_plan_vec_len.restype = C.c_int64
        L.mq_plan_flops.argtypes = [vp, f64p, f64p]
        L.mq_plan_workspace_bytes.argtypes = [vp]
        L.mq_plan_workspace_bytes.restype = C.c_int64
        L.mq_plan_exec_flops.argtypes = [vp]
        L.mq_plan_exec_flops.restype = C.c_double
        L.mq_sigma.argtypes = [vp, vp, vp, vp, vp, C.c_int]
        L.mq_sigma_timed.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.POINTER(C.c_float), f64p]
        L.mq_sigma_graph.argtypes = [vp, vp, vp, vp, vp, C.c_int]
        L.mq_sigma_batch.argtypes = [vp, vp, vp, vp, vp, C.c_int, C.c_int]
        L.mq_block_davidson.argtypes = [vp, vp, vp, vp, vp, vp, f64p, C.c_int, C.c_int, C.c_double, C.c_int,
                                        C.POINTER(vp), C.c_int, C.POINTER(C.c_int)]
        L.mq_plan_rebind.argtypes = [vp]
        L.mq_plan_bind.argtypes = [vp, vp, vp]
        L.mq_plan_live_mask.argtypes = [vp, vp]
        L.mq_plan_solve_info.argtypes = [vp, f64p]
        L.mq_plan_release.argtypes = [vp]
        L.mq_gemm_f64.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp,
                                  C.c_int, C.c_int]
        L.mq_gemm_f32.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp,
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.mq_step_out_desc.argtypes = [vp, C.POINTER(mq_boundary_desc)]
        L.mq_boundary_step.argtypes = [vp, vp, vp, vp]
        L.mq_jacobi_davidson.argtypes = [vp, vp, vp, vp, vp, f64p, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(vp), C.c_int, C.POINTER(C.c_int)]
        L.mq_diagonal.argtypes = [vp, vp, vp, vp]
        L.mq_davidson.argtypes = [vp, vp, vp, vp, vp, vp, f64p, C.c_int, C.c_double, C.c_int,
                                  C.POINTER(vp), C.c_int, C.POINTER(C.c_int)]
        L.mq_host_plan_create.argtypes = [vp, C.c_int, C.POINTER(vp)]
        L.mq_host_plan_part.argtypes = [vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.mq_host_plan_sizes.argtypes = [vp, i64p]
        L.mq_host_plan_flops.argtypes = [vp, f64p, f64p, f64p]
        L.mq_host_plan_out_desc.argtypes = [vp, C.POINTER(mq_boundary_desc)]
        L.mq_host_plan_destroy.argtypes = [vp]
        L.mq_host_plan_schedule.argtypes = [vp]
        L.mq_plan_schedule.argtypes = [vp]
        L.mq_split_create.argtypes = [vp, vp, C.POINTER(vp)]
        L.mq_split_destroy.argtypes = [vp]
        L.mq_split_blocks.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(i32p), C.POINTER(i32p), C.POINTER(i32p),
                                      C.POINTER(C.c_int64)]
        L.mq_split_decompose.argtypes = [vp, vp, f64p]
        L.mq_split_assemble.argtypes = [vp, i32p, i64p, i64p, C.c_int, vp, vp]
        L.mq_split_sweeps.argtypes = [vp, i32p]
        L.mq_host_split_both.argtypes = [vp, vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), vp]
        L.mq_host_last_error.restype = C.c_char_p
        _lib = L
    return _lib


# --------------------------------------------------------------------------- host plans
COPY2D = np.dtype([("dst", "<i8"), ("src", "<i8"), ("rows", "<i4"), ("cols", "<i4"), ("ldd", "<i4"),
                   ("lds", "<i4"), ("trans", "<i4"), ("scale", "<f8")], align=True)
GEMMGROUP = np.dtype([("a", "<i8"), ("b", "<i8"), ("c", "<i8"), ("part", "<i8"), ("M", "<i4"), ("N", "<i4"),
                      ("K", "<i4"), ("lda", "<i4"), ("ldb", "<i4"), ("ldc", "<i4"), ("transA", "<i4"),
                      ("transB", "<i4"), ("split", "<i4"), ("kchunk", "<i4"), ("tm", "<i4"), ("tn", "<i4"),
                      ("rowmap", "<i8")], align=True)
ROWMAP = np.dtype([("off", "<i8"), ("ld", "<i4"), ("pad", "<i4")], align=True)
GEMMTILE = np.dtype([("g", "<i4"), ("mt", "<i4"), ("nt", "<i4"), ("ks", "<i4")], align=True)
STRIP = np.dtype([("dst", "<i8"), ("rows", "<i4"), ("cols", "<i4"), ("ldd", "<i4"), ("s0", "<i4"),
                  ("s1", "<i4")], align=True)
SRC = np.dtype([("off", "<i8"), ("ld", "<i4"), ("coef", "<f8")], align=True)
PART_DTYPES = [COPY2D, COPY2D, COPY2D, GEMMGROUP, GEMMTILE, STRIP, SRC, GEMMGROUP, GEMMTILE, COPY2D, STRIP, SRC,
               ROWMAP, ROWMAP, np.dtype(np.uint8)]
PART_NAMES = ["arrL", "arrR", "reshape", "g1_groups", "g1_tiles", "strips", "srcs", "g2_groups", "g2_tiles",
              "scatter", "comb_strips", "comb_srcs", "g1_rowmap", "g2_rowmap", "live"]
KIND = {"sigma": 0, "lstep": 1, "rstep": 2, "diag": 3}


class HostPlan:
    """Task lists of a plan, built on the host (no GPU needed)."""

    def __init__(self, site: SiteDesc, kind: str, fuse=1, fuse_ratio=0, fuse_min=0, schedule=0):
        """schedule 1: sigma plans take the reference's per-site choice of lbtm or
        rbtm (MQ_HOST_SCHED_REFERENCE); self.rbtm tells which was built."""
        L = lib()
        h = C.c_void_p()
        code = KIND[kind] | ((fuse + 1) << 8) | (fuse_ratio << 10) | (fuse_min << 18) | ((schedule & 1) << 30)
        if L.mq_host_plan_create(site.ptr(), code, C.byref(h)) != 0:
            raise MQError(L.mq_host_last_error().decode())
        try:
            self.parts = {}
            for i, (name, dt) in enumerate(zip(PART_NAMES, PART_DTYPES)):
                ptr, cnt, el = C.c_void_p(), C.c_int64(), C.c_int64()
                L.mq_host_plan_part(h, i, C.byref(ptr), C.byref(cnt), C.byref(el))
                if el.value != dt.itemsize:
                    raise MQError(f"task record size mismatch for {name}: {el.value} != {dt.itemsize}")
                if cnt.value:
                    buf = (C.c_char * (cnt.value * el.value)).from_address(ptr.value)
                    self.parts[name] = np.frombuffer(bytes(buf), dtype=dt).copy()
                else:
                    self.parts[name] = np.zeros(0, dtype=dt)
            s = np.zeros(10, np.int64)
            L.mq_host_plan_sizes(h, s.ctypes.data_as(i64p))
            self.sizes = dict(zip(["vec", "la", "rb", "prp", "t", "c", "tmp", "out", "ws1", "ws2"], s.tolist()))
            a, b, c = C.c_double(), C.c_double(), C.c_double()
            L.mq_host_plan_flops(h, C.byref(a), C.byref(b), C.byref(c))
            self.ref_flops, self.exec_flops, self.min_bytes = a.value, b.value, c.value
            self.rbtm = L.mq_host_plan_schedule(h) == 1
            self.out_layout = None
            if kind in ("lstep", "rstep"):
                d = mq_boundary_desc()
                L.mq_host_plan_out_desc(h, C.byref(d))
                self.out_layout = BoundaryLayout.from_c(d)
        finally:
            L.mq_host_plan_destroy(h)


def host_split_both(p1, p2, left, right, psi, su2=True):
    """The both-paired blocks of the two-site tensor psi (host, canonical
    left-paired flat) from mq_split's gather map applied on the host
    (mq_host_split_both, a check of the map; no GPU).  Returns a list of
    row-major (rows, cols) arrays in mq_split_blocks order."""
    L = lib()
    d = SplitDesc(MQ_SYMM_SU2U1 if su2 else MQ_SYMM_TWOU1, p1, p2, left, right)
    nb, n = C.c_int32(), C.c_int64()
    if L.mq_host_split_both(d.ptr(), None, C.byref(nb), C.byref(n), None) != 0:
        raise MQError(L.mq_host_last_error().decode())
    x = np.ascontiguousarray(psi, dtype=np.float64)
    out = np.zeros(max(n.value, 1))
    if L.mq_host_split_both(d.ptr(), x.ctypes.data_as(C.c_void_p), C.byref(nb), C.byref(n),
                            out.ctypes.data_as(C.c_void_p)) != 0:
        raise MQError(L.mq_host_last_error().decode())
    return out[:n.value]


# --------------------------------------------------------------------------- GPU engine
class GpuBoundary:
    __slots__ = ("layout", "data")

    def __init__(self, layout: BoundaryLayout, data):
        self.layout = layout
        self.data = data  # torch tensor (float64, cuda)

    def dicts(self):
        return self.layout.to_dicts(self.data.cpu().numpy())


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class GpuEngine:
    """Engine interface on libmqgpu: sigma / Jacobi-Davidson / boundary steps on one GPU."""

    name = "gpu"

    def __init__(self, device=0, prec=MQ_PREC_FP64, fuse=1):
        import torch
        if not torch.cuda.is_available():
            raise MQError("no GPU visible: the hot path has no CPU fallback")
        self.torch = torch
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        self.ctx = C.c_void_p()
        if lib().mq_ctx_create(device, C.byref(self.ctx)) != 0:
            raise MQError("mq_ctx_create failed")
        # the library's stream as a torch stream: work on it is ordered against
        # torch's current stream with events (enter: the library waits for the
        # inputs torch produced; leave: torch waits for the library's outputs)
        self.stream = torch.cuda.ExternalStream(lib().mq_ctx_stream(self.ctx), device=self.device)
        from .autotune import Schedule
        self.set_schedule(Schedule(fuse=fuse))
        self.prec = prec
        self.flops = 0.0
        self.stats = []

    def set_fuse(self, mode=1, ratio=4, min_entries=64):
        """Fused-column rule of the plans built afterwards (MQ_OPT_FUSE*)."""
        for opt, v in ((0, mode), (1, ratio), (2, min_entries)):
            self.check(lib().mq_ctx_set_option(self.ctx, opt, v), "mq_ctx_set_option")

    def enter(self):
        self.stream.wait_stream(self.torch.cuda.current_stream(self.device))

    def leave(self):
        self.torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def set_ref_flops(self, on):
        """Count the reference schedule's FLOPs in the plans built afterwards
        (MQ_OPT_REF_FLOPS); without the count a plan's ref_flops is -1 and the
        engine's running total uses the executed FLOPs."""
        self.check(lib().mq_ctx_set_option(self.ctx, 6, 1 if on else 0), "mq_ctx_set_option")

    def set_graph(self, on):
        """Davidson drivers run their sigma vectors through captured HIP graphs (MQ_OPT_GRAPH)."""
        self.check(lib().mq_ctx_set_option(self.ctx, 4, 1 if on else 0), "mq_ctx_set_option")

    def set_schedule(self, rule):
        """Plan-construction schedule of the plans built afterwards: a
        Schedule (fused-column rule, split-K chunk, GEMM row and column tiles)."""
        self.set_fuse(rule.fuse, rule.ratio, rule.min_entries)
        self.check(lib().mq_ctx_set_option(self.ctx, 3, rule.kchunk), "mq_ctx_set_option")
        self.check(lib().mq_ctx_set_option(self.ctx, 5, rule.tile_m), "mq_ctx_set_option")
        self.check(lib().mq_ctx_set_option(self.ctx, 7, rule.sched), "mq_ctx_set_option")
        self.check(lib().mq_ctx_set_option(self.ctx, 8, rule.tile_n), "mq_ctx_set_option")
        self.schedule = rule

    def close(self):
        if self.ctx:
            lib().mq_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def err(self):
        return lib().mq_last_error(self.ctx).decode()

    def check(self, rc, what):
        if rc != 0:
            raise MQError(f"{what}: {self.err()} (code {rc})")

    def sync(self):
        self.check(lib().mq_ctx_synchronize(self.ctx), "synchronize")

    def to_dev(self, x):
        return self.torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)

    def gemm(self, a, b, transA=False, transB=False, kchunk=2048):
        """a @ b (device float64 tensors, M x K and K x N) with the library's
        FP64 GEMM kernel; transA / transB select the kernel path that reads
        the operand stored transposed."""
        t = self.torch
        M, K = a.shape
        N = b.shape[1]
        A = a.contiguous() if transA else a.t().contiguous()  # column-major K x M if transA else M x K
        B = b.contiguous() if transB else b.t().contiguous()  # column-major N x K if transB else K x N
        Cm = t.empty((N, M), dtype=t.float64, device=self.device)
        t.cuda.synchronize()
        lda = K if transA else M
        ldb = N if transB else K
        self.check(lib().mq_gemm_f64(self.ctx, int(transA), int(transB), M, N, K, _ptr(A), max(lda, 1), _ptr(B),
                                     max(ldb, 1), _ptr(Cm), max(M, 1), kchunk), "mq_gemm_f64")
        return Cm.t()

    def gemm32(self, a, b, transA=False, transB=False, kchunk=2048, tm=64, tn=64, bf16x3=False):
        """a @ b (device float32 tensors) with the FP32-operand kernel of the mixed
        forms, tile tm x tn, in FP32 or BF16x3 products."""
        t = self.torch
        M, K = a.shape
        N = b.shape[1]
        A = a.contiguous() if transA else a.t().contiguous()
        B = b.contiguous() if transB else b.t().contiguous()
        Cm = t.empty((N, M), dtype=t.float32, device=self.device)
        t.cuda.synchronize()
        lda = K if transA else M
        ldb = N if transB else K
        self.check(lib().mq_gemm_f32(self.ctx, int(transA), int(transB), M, N, K, _ptr(A), max(lda, 1), _ptr(B),
                                     max(ldb, 1), _ptr(Cm), max(M, 1), kchunk, tm, tn, int(bf16x3)), "mq_gemm_f32")
        return Cm.t()

    def seed_boundary(self, charge):
        return GpuBoundary(BoundaryLayout.seed(charge), self.to_dev(np.ones(1)))

    def upload_boundary(self, layout, host_data):
        return GpuBoundary(layout, self.to_dev(host_data))

    def _plan(self, kind, site):
        p = C.c_void_p()
        fn = {"sigma": lib().mq_sigma_plan_create, "lstep": lib().mq_lstep_plan_create,
              "rstep": lib().mq_rstep_plan_create, "diag": lib().mq_diag_plan_create}[kind]
        self.check(fn(self.ctx, site.ptr(), C.byref(p)), f"{kind} plan")
        return p

    def _step(self, kind, symm, A, B: GpuBoundary, W):
        site = SiteDesc(symm, A.phys, A.left, A.right, W, B.layout if kind == "lstep" else None,
                        B.layout if kind == "rstep" else None)
        p = self._plan(kind, site)
        try:
            d = mq_boundary_desc()
            self.check(lib().mq_step_out_desc(p, C.byref(d)), "out desc")
            lay = BoundaryLayout.from_c(d)
            out = self.torch.empty(max(lay.total, 1), dtype=self.torch.float64, device=self.device)
            a = self.to_dev(A.to_flat())
            self.torch.cuda.synchronize()
            self.check(lib().mq_boundary_step(p, _ptr(B.data), _ptr(a), _ptr(out)), kind)
            self.sync()
            fl = C.c_double()
            lib().mq_plan_flops(p, C.byref(fl), None)
            self.flops += fl.value if fl.value >= 0 else lib().mq_plan_exec_flops(p)
        finally:
            lib().mq_plan_destroy(p)
        return GpuBoundary(lay, out)

    def left_step(self, symm, A, L, W):
        return self._step("lstep", symm, A, L, W)

    def right_step(self, symm, A, R, W):
        return self._step("rstep", symm, A, R, W)

    def sigma_plan(self, symm, psi, L, R, W):
        site = SiteDesc(symm, psi.phys, psi.left, psi.right, W, L.layout, R.layout)
        return SigmaPlan(self, site, L, R)

    def sigma(self, symm, psi, L, R, W, x=None):
        sp = self.sigma_plan(symm, psi, L, R, W)
        try:
            xd = self.to_dev(psi.to_flat() if x is None else x)
            self.torch.cuda.synchronize()
            y = sp.apply(xd)
            self.sync()
            return y.cpu().numpy(), sp.ref_flops
        finally:
            sp.close()

    def jd(self, symm, psi, L, R, W, max_iter=10, tol=1e-8, ortho=(), x0=None):
        """Jacobi-Davidson on one site problem; last_timing records the host plan
        build (with the boundary binding), the vector upload, the solve and the
        download, and the solve's two phases (mq_plan_solve_info)."""
        import time
        t0 = time.perf_counter()
        sp = self.sigma_plan(symm, psi, L, R, W)
        try:
            x = self.to_dev(psi.to_flat() if x0 is None else x0)
            self.sync()
            t1 = time.perf_counter()
            theta, u, its = sp.jd(x, max_iter, tol, ortho)
            t2 = time.perf_counter()
            uh = u.cpu().numpy()
            self._last_u = (uh, u)  # split_twosite of this vector uses the device copy
            t3 = time.perf_counter()
            self.flops += its * sp.work_flops
            info = sp.solve_info()
            self.last_timing = dict(t_plan=t1 - t0, t_iter=t2 - t1, t_down=t3 - t2,
                                    form_iterations=info["form_iterations"], fp64_iterations=info["fp64_iterations"])
            return theta, uh, its
        finally:
            sp.close()

    def split_twosite(self, u, p1, p2, left, right, Mmax, cutoff, l2r=True, su2=True):
        """TwoSiteTensor::split_mps_l2r / split_mps_r2l (twositetensor.hpp:141-181)
        on the device: mq_split's map to the both-paired blocks and block Jacobi
        SVD, the reference's global truncation rule on the singular values
        (mps.truncation_keep), and the two site tensors assembled on the device.
        u: the two-site tensor as a canonical left-paired flat vector -- the
        vector the last solve returned (its device copy is used) or any host
        array.  Returns (A, B, info) as mps.split_twosite: A left-paired site
        tensor of (p1, left, kept bond), B right-paired of (p2, kept bond, right);
        l2r puts the singular values on B, r2l on A.  last_split records the
        host-side times."""
        import time
        from .blocks import BlockMatrix, Index
        from .mps import MPSTensor, canonical_layout, truncation_keep
        t = self.torch
        L = lib()
        t0 = time.perf_counter()
        d = SplitDesc(MQ_SYMM_SU2U1 if su2 else MQ_SYMM_TWOU1, p1, p2, left, right)
        sp = C.c_void_p()
        self.check(L.mq_split_create(self.ctx, d.ptr(), C.byref(sp)), "mq_split_create")
        try:
            nb, n = C.c_int32(), C.c_int64()
            ch, rw, cl = i32p(), i32p(), i32p()
            self.check(L.mq_split_blocks(sp, C.byref(nb), C.byref(ch), C.byref(rw), C.byref(cl), C.byref(n)),
                       "mq_split_blocks")
            nb = nb.value
            charges = [(ch[2 * b], ch[2 * b + 1]) for b in range(nb)]
            rows = [rw[b] for b in range(nb)]
            cols = [cl[b] for b in range(nb)]
            last = getattr(self, "_last_u", None)
            if last is not None and last[0] is u:
                ud = last[1]
            else:
                ud = self.to_dev(u)
            if ud.numel() != n.value:
                raise MQError(f"split: the two-site vector has {ud.numel()} elements, the layout {n.value}")
            nsv = [min(r, c) for r, c in zip(rows, cols)]
            s = np.zeros(max(sum(nsv), 1))
            t1 = time.perf_counter()
            t.cuda.synchronize()
            self.check(L.mq_split_decompose(sp, _ptr(ud), s.ctypes.data_as(f64p)), "mq_split_decompose")
            t2 = time.perf_counter()
            sw = np.zeros(max(nb, 1), np.int32)
            L.mq_split_sweeps(sp, sw.ctypes.data_as(i32p))
            self.last_split_blocks = dict(rows=rows, cols=cols, sweeps=sw[:nb].tolist())
            svals, o = [], 0
            for k in nsv:
                svals.append(s[o:o + k])
                o += k
            keep, info = truncation_keep(svals, Mmax, cutoff)
            mid = Index.from_dict({q: k for q, k in zip(charges, keep) if k > 0})
            lay, na = canonical_layout(p1, left, mid)
            aoff = {c: off for c, _, _, off in lay}
            a_off = np.array([aoff.get(q, 0) if k else 0 for q, k in zip(charges, keep)], np.int64)
            b_off = np.zeros(nb, np.int64)
            nbv = 0
            for b in range(nb):
                b_off[b] = nbv
                nbv += keep[b] * cols[b]
            Ad = t.empty(max(na, 1), dtype=t.float64, device=self.device)
            Bd = t.empty(max(nbv, 1), dtype=t.float64, device=self.device)
            kp = np.array(keep, np.int32)
            self.check(L.mq_split_assemble(sp, kp.ctypes.data_as(i32p), a_off.ctypes.data_as(i64p),
                                           b_off.ctypes.data_as(i64p), 1 if l2r else 0, _ptr(Ad), _ptr(Bd)),
                       "mq_split_assemble")
            Ah = Ad.cpu().numpy()[:na]
            Bh = Bd.cpu().numpy()[:nbv]
            A = MPSTensor.from_flat(p1, left, mid, Ah)
            Bb = BlockMatrix()
            for b in range(nb):
                if keep[b]:
                    Bb.blocks[(charges[b], charges[b])] = Bh[b_off[b]:b_off[b] + keep[b] * cols[b]].reshape(
                        keep[b], cols[b])
            B = MPSTensor(p2, mid, right, Bb, "R")
            t3 = time.perf_counter()
            self.last_split = dict(t_map=t1 - t0, t_decompose=t2 - t1, t_assemble=t3 - t2)
            return A, B, info
        finally:
            L.mq_split_destroy(sp)

    def davidson(self, symm, psi, L, R, W, max_iter=10, tol=1e-8, ortho=(), x0=None):
        """Davidson with the diagonal preconditioner (SU2U1), solve_ietl_davidson."""
        sp = self.sigma_plan(symm, psi, L, R, W)
        try:
            theta, u, its = sp.davidson(self.to_dev(psi.to_flat() if x0 is None else x0), max_iter, tol, ortho)
            self.flops += its * sp.work_flops
            return theta, u.cpu().numpy(), its
        finally:
            sp.close()

    def block_solve(self, symm, psis, L, R, W, max_iter=10, tol=1e-8):
        """The len(psis) lowest eigenpairs together (mq_block_davidson, one batched
        sigma per block iteration): (thetas, [u_r], block iterations).  SU2U1
        problems precondition the corrections with the diagonal of H_eff."""
        sp = self.sigma_plan(symm, psis[0], L, R, W)
        try:
            X0 = self.torch.as_tensor(independent_starts([p.to_flat() for p in psis],
                                                         live=sp.live_mask())).to(self.device)
            diag = sp.diagonal() if symm == "su2u1" else None
            thetas, U, its = sp.block_davidson(X0, max_iter, tol, diag=diag)
            self.flops += its * len(psis) * sp.work_flops
            Uh = U.cpu().numpy()
            return thetas, [Uh[r] for r in range(len(psis))], its
        finally:
            sp.close()

    def diagonal(self, symm, psi, L, R, W):
        """Diagonal of H_eff (SU2U1), contraction::SU2::diagonal_hamiltonian."""
        sp = self.sigma_plan(symm, psi, L, R, W)
        try:
            d = sp.diagonal()
            self.sync()
            return d.cpu().numpy()
        finally:
            sp.close()


def independent_starts(vecs, seed=0, live=None):
    """Orthonormal start vectors for a block solve: the given vectors by
    Gram-Schmidt, each one that is (numerically) in the span of those before it
    replaced by a random vector.  The roots' two-site tensors at the start of a
    state-averaged sweep can be dependent (all built on one shared neighbour).
    live (the plan's live mask, SigmaPlan.live_mask()): every vector is
    restricted to the coordinates H_eff couples, so that the structurally zero
    columns of H_eff, whose eigenvalue 0 may lie below every physical root,
    cannot enter the block."""
    rng = np.random.default_rng(seed)
    keep = None if live is None else np.asarray(live, dtype=bool)
    out = []
    for v in vecs:
        v = np.array(v, dtype=np.float64)
        for _ in range(8):
            if keep is not None:
                v = np.where(keep, v, 0.0)
            nv = np.linalg.norm(v)
            w = v.copy()
            for q in out:
                w -= (q @ w) * q
            for q in out:
                w -= (q @ w) * q
            if nv > 0 and np.linalg.norm(w) > 1e-8 * nv:
                break
            v = rng.standard_normal(v.size)
        else:
            raise MQError(f"block solve: cannot find {len(vecs)} independent start vectors of length {v.size}")
        out.append(w / np.linalg.norm(w))
    return np.stack(out)


class SigmaPlan:
    def __init__(self, eng: GpuEngine, site, L: GpuBoundary, R: GpuBoundary):
        self.eng = eng
        self.site = site
        self.L, self.R = L, R
        self.p = eng._plan("sigma", site)
        fl, by = C.c_double(), C.c_double()
        lib().mq_plan_flops(self.p, C.byref(fl), C.byref(by))
        self.ref_flops = fl.value
        self.min_bytes = by.value
        self.exec_flops = lib().mq_plan_exec_flops(self.p)
        # FLOPs per sigma for running totals: the reference count, or the executed one
        # when the plan was built without it (MQ_OPT_REF_FLOPS = 0)
        self.work_flops = self.ref_flops if self.ref_flops >= 0 else self.exec_flops
        self.n = lib().mq_plan_vec_len(self.p)
        self.workspace_bytes = lib().mq_plan_workspace_bytes(self.p)
        self.rbtm = lib().mq_plan_schedule(self.p) == 1
        self.bind()

    def bind(self, L: GpuBoundary = None, R: GpuBoundary = None):
        """Arrange the boundaries (by default the plan's own) into the plan's
        panels now (mq_plan_bind); later calls use them.  Call it after
        writing new values into a boundary tensor, or to switch boundaries."""
        if L is not None:
            self.L = L
        if R is not None:
            self.R = R
        self.eng.enter()
        self.eng.check(lib().mq_plan_bind(self.p, _ptr(self.L.data), _ptr(self.R.data)), "mq_plan_bind")

    def solve_info(self):
        """The last solver call: dict(form_iterations, fp64_iterations,
        theta_form, theta_rayleigh_fp64) (mq_plan_solve_info)."""
        info = (C.c_double * 4)()
        self.eng.check(lib().mq_plan_solve_info(self.p, info), "mq_plan_solve_info")
        return dict(form_iterations=int(info[0]), fp64_iterations=int(info[1]), theta_form=info[2],
                    theta_rayleigh_fp64=info[3])

    def live_mask(self):
        """Host bool array: the vector coordinates H_eff couples (mq_plan_live_mask)."""
        m = np.zeros(self.n, dtype=np.uint8)
        self.eng.check(lib().mq_plan_live_mask(self.p, m.ctypes.data_as(C.c_void_p)), "mq_plan_live_mask")
        return m.astype(bool)

    def release(self):
        """Free the plan's cached work arrays (batched lists and arenas, Davidson
        vectors, captured graph); the next call that needs them rebuilds them."""
        self.eng.check(lib().mq_plan_release(self.p), "mq_plan_release")

    def close(self):
        if getattr(self, "_dp", None):
            lib().mq_plan_destroy(self._dp)
            self._dp = None
        if self.p:
            lib().mq_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, x, y=None, prec=None):
        t = self.eng.torch
        if y is None:
            y = t.empty(self.n, dtype=t.float64, device=self.eng.device)
        prec = self.eng.prec if prec is None else prec
        self.eng.enter()
        self.eng.check(lib().mq_sigma(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(x), _ptr(y), prec),
                       "mq_sigma")
        self.eng.leave()
        return y

    def apply_batch(self, X, Y=None, prec=None):
        """Y[j] = H X[j] for the rows of a (nvec, n) device tensor, one launch per stage."""
        t = self.eng.torch
        X = X.contiguous()
        nvec = X.shape[0]
        if Y is None:
            Y = t.empty((nvec, self.n), dtype=t.float64, device=self.eng.device)
        prec = self.eng.prec if prec is None else prec
        self.eng.enter()
        self.eng.check(lib().mq_sigma_batch(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(X), _ptr(Y), nvec,
                                            prec), "mq_sigma_batch")
        self.eng.leave()
        return Y

    def apply_graph(self, x, y=None, prec=None):
        """y = H x through the plan's captured HIP graph (mq_sigma_graph)."""
        t = self.eng.torch
        if y is None:
            y = t.empty(self.n, dtype=t.float64, device=self.eng.device)
        prec = self.eng.prec if prec is None else prec
        self.eng.enter()
        self.eng.check(lib().mq_sigma_graph(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(x), _ptr(y), prec),
                       "mq_sigma_graph")
        self.eng.leave()
        return y

    def timed(self, x, y, prec=None):
        """One sigma with HIP events between its stages on the library stream.

        Returns (ms, work): ms = [boundary arrangement, psi reshape, T GEMM,
        W application, output GEMM] in milliseconds; work = [T GEMM flops,
        W-application flops, output GEMM flops, and the minimum HBM bytes of
        the same three stages]."""
        ms = (C.c_float * 5)()
        work = (C.c_double * 6)()
        prec = self.eng.prec if prec is None else prec
        self.eng.enter()
        self.eng.check(lib().mq_sigma_timed(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(x), _ptr(y), prec,
                                            ms, work), "mq_sigma_timed")
        self.eng.leave()
        return list(ms), list(work)

    def diagonal(self):
        """Diagonal of H_eff as a device vector (builds a diagonal plan once)."""
        t = self.eng.torch
        if getattr(self, "_dp", None) is None:
            self._dp = self.eng._plan("diag", self.site)
        d = t.empty(self.n, dtype=t.float64, device=self.eng.device)
        self.eng.enter()
        self.eng.check(lib().mq_diagonal(self._dp, _ptr(self.L.data), _ptr(self.R.data), _ptr(d)), "mq_diagonal")
        self.eng.leave()
        return d

    def davidson(self, x0, max_iter=10, tol=1e-8, ortho=(), prec=None, diag=None):
        t = self.eng.torch
        prec = self.eng.prec if prec is None else prec
        if diag is None:
            diag = self.diagonal()
        u = t.empty(self.n, dtype=t.float64, device=self.eng.device)
        th = C.c_double()
        its = C.c_int()
        od = [self.eng.to_dev(o) for o in ortho]
        arr = (C.c_void_p * max(1, len(od)))(*[o.data_ptr() for o in od])
        t.cuda.synchronize()
        self.eng.check(lib().mq_davidson(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(diag), _ptr(x0),
                                         _ptr(u), C.byref(th), max_iter, tol, prec, arr, len(od), C.byref(its)),
                       "mq_davidson")
        return th.value, u, its.value

    def block_davidson(self, X0, max_iter=50, tol=1e-8, ortho=(), prec=None, diag=None):
        """Lowest nroots eigenpairs from the rows of X0 (nroots, n): (thetas, U (nroots, n), iterations).
        diag: a device diagonal (SigmaPlan.diagonal()) to precondition the corrections."""
        t = self.eng.torch
        prec = self.eng.prec if prec is None else prec
        X0 = X0.contiguous()
        k = X0.shape[0]
        U = t.empty((k, self.n), dtype=t.float64, device=self.eng.device)
        th = (C.c_double * k)()
        its = C.c_int()
        od = [self.eng.to_dev(o) for o in ortho]
        arr = (C.c_void_p * max(1, len(od)))(*[o.data_ptr() for o in od])
        t.cuda.synchronize()
        self.eng.check(lib().mq_block_davidson(self.p, _ptr(self.L.data), _ptr(self.R.data),
                                               None if diag is None else _ptr(diag), _ptr(X0), _ptr(U), th,
                                               k, max_iter, tol, prec, arr, len(od), C.byref(its)),
                       "mq_block_davidson")
        return list(th), U, its.value

    def jd(self, x0, max_iter=10, tol=1e-8, ortho=(), prec=None):
        t = self.eng.torch
        prec = self.eng.prec if prec is None else prec
        u = t.empty(self.n, dtype=t.float64, device=self.eng.device)
        th = C.c_double()
        its = C.c_int()
        od = [self.eng.to_dev(o) for o in ortho]
        arr = (C.c_void_p * max(1, len(od)))(*[o.data_ptr() for o in od])
        t.cuda.synchronize()
        self.eng.check(lib().mq_jacobi_davidson(self.p, _ptr(self.L.data), _ptr(self.R.data), _ptr(x0), _ptr(u),
                                                C.byref(th), max_iter, tol, prec, arr, len(od),
                                                C.byref(its)), "mq_jacobi_davidson")
        return th.value, u, its.value
