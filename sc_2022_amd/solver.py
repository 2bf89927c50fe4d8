"""Python host side of the MI355X SBP-SAT Poisson solver.

``SbpPoisson`` owns one C-ABI handle (operator tables, multigrid hierarchy,
work vectors) and exposes the path on torch device tensors.  Grid functions
are torch tensors on a HIP device with N*N elements in the reference's layout
(element (i, j) at i + j*N, i fast); a row-major (N, N) tensor t is the
transpose view of the reference's column-major (Nx, Ny) array: t[j, i] = u(i, j).

Every call goes to libsbp.so; nothing here computes on the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _capi as capi

_DT = {torch.float64: capi.SBP_F64, torch.float32: capi.SBP_F32}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


@dataclass
class SolveResult:
    iters: int
    resnorm: float
    status: int
    history: list | None = None
    inner_iters: int | None = None
    level_iters: list | None = None      # nested iteration: CG iterations per grid, coarse to fine
    level_resnorms: list | None = None

    @property
    def converged(self):
        return self.status == capi.SBP_OK


class SbpPoisson:
    """Matrix-free SBP-SAT operator of order p on an N x N grid (N = 2^k+1 for
    multigrid).  Keyword arguments override sbp_params fields."""

    def __init__(self, N: int, p: int = 2, device=None, tuned: bool = True, coef=None, **params):
        """coef = (bx, by): the variable-coefficient operator d/dx (bx d/dx) +
        d/dy (by d/dy) with N-value profiles along the slow (Dirichlet) and
        fast (Neumann) index (sbp_create_var; p = 2, 4, 6; apply, CG, MG-PCG and the
        right-hand side only)."""
        self.lib = capi.load()
        self.N = int(N)
        self.p = int(p)
        self.n = self.N * self.N
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.params = capi.default_params(**params)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            if coef is None:
                capi.check(self.lib.sbp_create(self.N, self.p, C.byref(self.params), C.byref(h)), "sbp_create")
            else:
                import numpy as np
                bx, by = (np.ascontiguousarray(np.asarray(c, dtype=np.float64)) for c in coef)
                if bx.shape != (self.N,) or by.shape != (self.N,):
                    raise ValueError("coef: two arrays of N values")
                dp = C.POINTER(C.c_double)
                capi.check(self.lib.sbp_create_var(self.N, self.p, bx.ctypes.data_as(dp), by.ctypes.data_as(dp),
                                                   C.byref(self.params), C.byref(h)), "sbp_create_var")
                tuned = False
        self._h = h
        nl = C.c_int()
        self.lib.sbp_get_info(self._h, None, None, C.byref(nl))
        self.nlevels = nl.value
        if tuned and not (params.get("tile_i") or params.get("tile_j")):
            self.apply_tuned_tiles()

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.lib.sbp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device)
        self.lib.sbp_set_stream(self._h, C.c_void_p(s.cuda_stream))

    def _chk(self, t, n=None, dtype=None, name="array"):
        n = self.n if n is None else n
        if t.device != self.device or not t.is_contiguous() or t.numel() != n:
            raise ValueError(f"{name}: need a contiguous tensor of {n} elements on {self.device}")
        if dtype is not None and t.dtype != dtype:
            raise ValueError(f"{name}: need dtype {dtype}, got {t.dtype}")
        return t

    def level_size(self, level):
        N = self.N
        for _ in range(level):
            N = (N + 1) // 2
        return N

    def set_tiles(self, level, ti, tj, dtype=None, fused=False, aligned=False):
        """Tile shape for `level` (fused: the block shape of the level's
        two-sweep kernels; aligned: the line-aligned variant of a stencil
        shape); both precisions unless dtype is given."""
        dts = [capi.SBP_F64, capi.SBP_F32] if dtype is None else [_DT[dtype]]
        flags = (capi.SBP_FUSED if fused else 0) | (capi.SBP_ALIGNED if aligned else 0)
        for d in dts:
            capi.check(self.lib.sbp_set_tiles(self._h, level, d | flags, ti, tj), "sbp_set_tiles")

    def apply_tuned_tiles(self, cache=None):
        """Applies the autotuner's choices (tiles_gfx950.json) to every level."""
        from . import autotune
        c = autotune.load_cache() if cache is None else cache
        for level in range(max(1, self.nlevels)):
            Nl = self.level_size(level)
            for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
                t = autotune.lookup(self.p, tag, Nl, c)
                if t:
                    self.set_tiles(level, t[0], t[1], dt, aligned=len(t) > 2 and bool(t[2]))

    # -- operator ----------------------------------------------------------
    def apply(self, u, out=None, sign=1, level=0):
        """out = sign * A u (sign = -1 is the reference's matrix_free_A_full_GPU)."""
        n = self.level_size(level) ** 2
        self._chk(u, n, name="u")
        out = torch.empty_like(u) if out is None else self._chk(out, n, u.dtype, "out")
        self._sync_stream()
        capi.check(self.lib.sbp_apply(self._h, level, _ptr(u), _ptr(out), sign, _DT[u.dtype]), "sbp_apply")
        return out

    def apply_dot(self, u, out=None):
        self._chk(u, dtype=torch.float64, name="u")
        out = torch.empty_like(u) if out is None else self._chk(out, dtype=torch.float64, name="out")
        d = C.c_double()
        self._sync_stream()
        capi.check(self.lib.sbp_apply_dot(self._h, _ptr(u), _ptr(out), C.byref(d)), "sbp_apply_dot")
        return out, d.value

    def diag(self):
        d = torch.empty(self.n, dtype=torch.float64, device=self.device)
        self._sync_stream()
        capi.check(self.lib.sbp_diag(self._h, _ptr(d)), "sbp_diag")
        return d

    def assemble_rhs(self, f=None, gW=None, gE=None, gS=None, gN=None, out=None):
        """b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN), f = Laplace(u)."""
        for t, nm, n in ((f, "f", self.n), (gW, "gW", self.N), (gE, "gE", self.N), (gS, "gS", self.N),
                         (gN, "gN", self.N)):
            if t is not None:
                self._chk(t, n, torch.float64, nm)
        b = torch.empty(self.n, dtype=torch.float64, device=self.device) if out is None else out
        self._sync_stream()
        capi.check(self.lib.sbp_assemble_rhs(self._h, _ptr(f), _ptr(gW), _ptr(gE), _ptr(gS), _ptr(gN), _ptr(b)),
                   "sbp_assemble_rhs")
        return b

    def rhs_transpose(self, lam):
        self._chk(lam, dtype=torch.float64, name="lam")
        z = lambda n: torch.empty(n, dtype=torch.float64, device=self.device)
        df, gW, gE, gS, gN = z(self.n), z(self.N), z(self.N), z(self.N), z(self.N)
        self._sync_stream()
        capi.check(self.lib.sbp_rhs_transpose(self._h, _ptr(lam), _ptr(df), _ptr(gW), _ptr(gE), _ptr(gS), _ptr(gN)),
                   "sbp_rhs_transpose")
        return df, gW, gE, gS, gN

    def h_weights(self):
        """1-D norm weights hw (N values, host numpy): H~ = h^2 hw_i hw_j."""
        import numpy as np
        hw = np.empty(self.N)
        capi.check(self.lib.sbp_h_weights(self._h, hw.ctypes.data_as(C.POINTER(C.c_double))), "sbp_h_weights")
        return hw

    def h_tilde(self):
        """The diagonal of H~ (device, N*N values, element (i, j) at i + j N)."""
        hw = torch.as_tensor(self.h_weights(), dtype=torch.float64, device=self.device)
        h = 1.0 / (self.N - 1)
        return (h * h * hw[None, :] * hw[:, None]).reshape(-1).contiguous()

    def hnorm(self, v):
        self._chk(v, dtype=torch.float64, name="v")
        d = C.c_double()
        self._sync_stream()
        capi.check(self.lib.sbp_hnorm(self._h, _ptr(v), C.byref(d)), "sbp_hnorm")
        return d.value

    def dot(self, u, v):
        self._chk(u, dtype=torch.float64, name="u")
        self._chk(v, dtype=torch.float64, name="v")
        d = C.c_double()
        self._sync_stream()
        capi.check(self.lib.sbp_dot(self._h, _ptr(u), _ptr(v), C.byref(d)), "sbp_dot")
        return d.value

    # -- multigrid components ---------------------------------------------
    def restrict(self, rf, level=0, out=None):
        Nc = self.level_size(level + 1)
        self._chk(rf, self.level_size(level) ** 2, name="rf")
        out = torch.empty(Nc * Nc, dtype=rf.dtype, device=self.device) if out is None else out
        self._sync_stream()
        capi.check(self.lib.sbp_restrict(self._h, level, _ptr(rf), _ptr(out), _DT[rf.dtype]), "sbp_restrict")
        return out

    def prolong_add(self, ec, xf, level=0):
        self._chk(ec, self.level_size(level + 1) ** 2, name="ec")
        self._chk(xf, self.level_size(level) ** 2, ec.dtype, "xf")
        self._sync_stream()
        capi.check(self.lib.sbp_prolong_add(self._h, level, _ptr(ec), _ptr(xf), _DT[ec.dtype]), "sbp_prolong_add")
        return xf

    def smooth(self, b, x, nsweeps, level=0):
        n = self.level_size(level) ** 2
        self._chk(b, n, name="b")
        self._chk(x, n, b.dtype, "x")
        self._sync_stream()
        capi.check(self.lib.sbp_smooth(self._h, level, _ptr(b), _ptr(x), nsweeps, _DT[b.dtype]), "sbp_smooth")
        return x

    def coarse_solve(self, b):
        x = torch.empty_like(b)
        self._sync_stream()
        capi.check(self.lib.sbp_coarse_solve(self._h, _ptr(b), _ptr(x)), "sbp_coarse_solve")
        return x

    def vcycle(self, r, out=None):
        self._chk(r, name="r")
        out = torch.empty_like(r) if out is None else out
        self._sync_stream()
        capi.check(self.lib.sbp_vcycle(self._h, _ptr(r), _ptr(out), _DT[r.dtype]), "sbp_vcycle")
        return out

    # -- solvers -----------------------------------------------------------
    def _solve(self, fn, name, b, x, abstol, maxiter, history, raise_on_fail):
        self._chk(b, dtype=torch.float64, name="b")
        x = torch.zeros_like(b) if x is None else self._chk(x, dtype=torch.float64, name="x")
        it = C.c_int()
        rn = C.c_double()
        hist = torch.zeros(maxiter + 1, dtype=torch.float64, device=self.device) if history else None
        self._sync_stream()
        st = fn(self._h, _ptr(b), _ptr(x), float(abstol), int(maxiter), C.byref(it), C.byref(rn), _ptr(hist))
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, name)
        hl = hist[: it.value + 1].cpu().tolist() if history else None
        return x, SolveResult(it.value, rn.value, st, hl)

    def cg(self, b, x=None, abstol=1e-8, maxiter=None, history=False, raise_on_fail=True):
        maxiter = self.n if maxiter is None else maxiter
        return self._solve(self.lib.sbp_cg, "sbp_cg", b, x, abstol, maxiter, history, raise_on_fail)

    def mgpcg(self, b, x=None, abstol=1e-8, maxiter=500, history=False, raise_on_fail=True):
        return self._solve(self.lib.sbp_mgpcg, "sbp_mgpcg", b, x, abstol, maxiter, history, raise_on_fail)

    def cg_interp(self, b, b2h, b4h=None, x=None, abstol=1e-8, maxiter=None, history=False, raise_on_fail=True):
        """Initial-guess interpolation CG: exact solve on the coarsest grid
        (b2h on side (N+1)/2, or b4h on the next one down when given), bilinear
        prolongation, CG on every finer grid from the interpolated solution.
        Returns (x, SolveResult) with the fine grid's iterations and the
        per-grid ones in level_iters."""
        nlev = 2 if b4h is None else 3
        self._chk(b, dtype=torch.float64, name="b")
        self._chk(b2h, self.level_size(1) ** 2, torch.float64, "b2h")
        if b4h is not None:
            self._chk(b4h, self.level_size(2) ** 2, torch.float64, "b4h")
        x = torch.empty_like(b) if x is None else self._chk(x, dtype=torch.float64, name="x")
        maxiter = self.n if maxiter is None else maxiter
        its = (C.c_int * (nlev - 1))()
        rns = (C.c_double * (nlev - 1))()
        hist = torch.zeros(maxiter + 1, dtype=torch.float64, device=self.device) if history else None
        self._sync_stream()
        st = self.lib.sbp_cg_interp(self._h, nlev, _ptr(b), _ptr(b2h), _ptr(b4h), _ptr(x), float(abstol),
                                    int(maxiter), its, rns, _ptr(hist))
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_cg_interp")
        hl = hist[: its[nlev - 2] + 1].cpu().tolist() if history else None
        return x, SolveResult(its[nlev - 2], rns[nlev - 2], st, hl, None, list(its), list(rns))

    def mgpcg_mixed(self, b, x=None, abstol=1e-8, max_outer=None, raise_on_fail=True):
        self._chk(b, dtype=torch.float64, name="b")
        x = torch.zeros_like(b) if x is None else self._chk(x, dtype=torch.float64, name="x")
        max_outer = self.params.max_outer if max_outer is None else max_outer
        oi, ii, rn = C.c_int(), C.c_int(), C.c_double()
        self._sync_stream()
        st = self.lib.sbp_mgpcg_mixed(self._h, _ptr(b), _ptr(x), float(abstol), int(max_outer), C.byref(oi),
                                      C.byref(ii), C.byref(rn))
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_mgpcg_mixed")
        return x, SolveResult(oi.value, rn.value, st, None, ii.value)

    def adjoint_grad(self, u, u_obs, abstol=None, maxiter=500, lam=None):
        """Gradient of L = 1/2 (u - u_obs)' H~ (u - u_obs), u the forward solution,
        through the adjoint solve.  Returns (L, lam, df, dgW, dgE, dgS, dgN, iters)."""
        self._chk(u, dtype=torch.float64, name="u")
        self._chk(u_obs, dtype=torch.float64, name="u_obs")
        z = lambda n: torch.empty(n, dtype=torch.float64, device=self.device)
        lam = z(self.n) if lam is None else self._chk(lam, dtype=torch.float64, name="lam")
        df, gW, gE, gS, gN = z(self.n), z(self.N), z(self.N), z(self.N), z(self.N)
        if abstol is None:
            # 1e-13 relative to the adjoint source H~ (u - u_obs)
            d = u - u_obs
            abstol = 1e-13 * float(torch.linalg.vector_norm(d)) * (1.0 / (self.N - 1)) ** 2
        loss = C.c_double()
        it = C.c_int()
        self._sync_stream()
        capi.check(self.lib.sbp_adjoint_grad(self._h, _ptr(u), _ptr(u_obs), float(abstol), int(maxiter),
                                             C.byref(loss), _ptr(lam), _ptr(df), _ptr(gW), _ptr(gE), _ptr(gS),
                                             _ptr(gN), C.byref(it)), "sbp_adjoint_grad")
        return loss.value, lam, df, gW, gE, gS, gN, it.value

    # -- CSR comparison path -------------------------------------------------
    def csr_assemble(self):
        nnz = C.c_int64()
        self._sync_stream()
        capi.check(self.lib.sbp_csr_assemble(self._h, None, None, None, C.byref(nnz)), "sbp_csr_assemble")
        rp = torch.empty(self.n + 1, dtype=torch.int32, device=self.device)
        col = torch.empty(nnz.value, dtype=torch.int32, device=self.device)
        val = torch.empty(nnz.value, dtype=torch.float64, device=self.device)
        capi.check(self.lib.sbp_csr_assemble(self._h, _ptr(rp), _ptr(col), _ptr(val), C.byref(nnz)),
                   "sbp_csr_assemble")
        return rp, col, val

    def csr_spmv(self, csr, x, out=None):
        rp, col, val = csr
        out = torch.empty_like(x) if out is None else out
        self._sync_stream()
        capi.check(self.lib.sbp_csr_spmv(self._h, _ptr(rp), _ptr(col), _ptr(val), _ptr(x), _ptr(out)),
                   "sbp_csr_spmv")
        return out

    def csr_cg(self, csr, b, x=None, abstol=1e-8, maxiter=None, raise_on_fail=True):
        rp, col, val = csr
        maxiter = self.n if maxiter is None else maxiter
        x = torch.zeros_like(b) if x is None else x
        it, rn = C.c_int(), C.c_double()
        self._sync_stream()
        st = self.lib.sbp_csr_cg(self._h, _ptr(rp), _ptr(col), _ptr(val), _ptr(b), _ptr(x), float(abstol),
                                 int(maxiter), C.byref(it), C.byref(rn))
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_csr_cg")
        return x, SolveResult(it.value, rn.value, st)

    # -- the same CSR matrix through rocSPARSE (the vendor SpMV) ------------
    def csr_vendor_attach(self, csr, alg=0):
        """Hands the CSR arrays to rocSPARSE (kept by reference: the handle
        holds on to the tensors) and runs its analysis (alg 0: csr_adaptive,
        1: csr_rowsplit)."""
        rp, col, val = csr
        self._sync_stream()
        capi.check(self.lib.sbp_csr_vendor_attach(self._h, _ptr(rp), _ptr(col), _ptr(val), int(col.numel()), int(alg)),
                   "sbp_csr_vendor_attach")
        self._vendor_csr = csr

    def csr_vendor_spmv(self, x, out=None):
        out = torch.empty_like(x) if out is None else out
        self._sync_stream()
        capi.check(self.lib.sbp_csr_vendor_spmv(self._h, _ptr(x), _ptr(out)), "sbp_csr_vendor_spmv")
        return out

    def csr_vendor_cg(self, b, x=None, abstol=1e-8, maxiter=None, raise_on_fail=True):
        maxiter = self.n if maxiter is None else maxiter
        x = torch.zeros_like(b) if x is None else x
        it, rn = C.c_int(), C.c_double()
        self._sync_stream()
        st = self.lib.sbp_csr_vendor_cg(self._h, _ptr(b), _ptr(x), float(abstol), int(maxiter), C.byref(it),
                                        C.byref(rn))
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_csr_vendor_cg")
        return x, SolveResult(it.value, rn.value, st)

    def copy(self, src, dst):
        self._sync_stream()
        capi.check(self.lib.sbp_copy(self._h, _ptr(src), _ptr(dst), src.numel() * src.element_size()), "sbp_copy")
        return dst


class PoissonSolve(torch.autograd.Function):
    """u = A^-1 b(f, gW, gE, gS, gN) as a differentiable torch operation.

    forward: right-hand side assembly + FP64 MG-PCG; backward: the adjoint
    solve lambda = A^-T grad_u (A is symmetric, so the same MG-PCG) and the
    transposed right-hand-side map.  The Python counterpart of a
    ChainRulesCore.rrule for the solve."""

    @staticmethod
    def forward(ctx, op, f, gW, gE, gS, gN, rtol=1e-13):
        b = op.assemble_rhs(f, gW, gE, gS, gN)
        bn = float(torch.linalg.vector_norm(b))
        u, res = op.mgpcg(b, abstol=rtol * bn)
        if not res.converged:
            raise RuntimeError(f"PoissonSolve forward: MG-PCG did not converge ({res})")
        ctx.op = op
        ctx.rtol = rtol
        # data passed as None (zero) gets no gradient
        ctx.present = tuple(v is not None for v in (f, gW, gE, gS, gN))
        return u

    @staticmethod
    def backward(ctx, gu):
        op = ctx.op
        gu = gu.contiguous()
        gn = float(torch.linalg.vector_norm(gu))
        lam = torch.zeros_like(gu)
        if gn > 0:
            lam, res = op.mgpcg(gu, abstol=ctx.rtol * gn)
            if not res.converged:
                raise RuntimeError(f"PoissonSolve backward: adjoint MG-PCG did not converge ({res})")
        grads = op.rhs_transpose(lam)
        grads = [g if (have and ctx.needs_input_grad[k + 1]) else None
                 for k, (g, have) in enumerate(zip(grads, ctx.present))]
        return (None, *grads, None)


def poisson_solve(op, f, gW, gE, gS, gN, rtol=1e-13):
    return PoissonSolve.apply(op, f, gW, gE, gS, gN, rtol)
