"""Batched solves of many independent SBP-SAT Poisson problems.

``SbpBatch`` wraps one stacked handle (sbp_create_batch): nprob problems of
the same side N and order p whose grid functions sit back to back in one
device array, problem q at element offset q*N*N.  Every kernel launch covers
all of them; each problem has its own tolerance and device-side convergence
mask, so a converged problem's tiles fall through while the others iterate.

``RaggedBatch`` wraps one ragged handle (sbp_create_ragged): problems of one
order and uneven sides whose multigrid levels are shared, so each V-cycle
pass is one launch over every problem present on that level.
``UnevenBatch`` solves problems of mixed (N, p) with one ragged handle per
order, the orders running concurrently on their own streams.  Per-problem
times are device wall-clock stamps from the solve start to that problem's
convergence.

No counterpart in the reference, which solves one problem per call of
matrix_free_MGCG (initial_guess_cg.jl:303-346).
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field

import torch

from . import _capi as capi
from .solver import _DT, _ptr


@dataclass
class BatchResult:
    iters: list              # per problem: MG-PCG iterations, or refinement steps (mixed)
    resnorm: list            # per problem: last ||r||_2 evaluated
    seconds: list            # per problem: device time from the solve start to convergence (-1: none)
    status: int              # worst status over the problems
    statuses: list           # per problem status
    inner_iters: list | None = None


class SbpBatch:
    """nprob stacked problems of side N and order p.  Keyword arguments
    override sbp_params fields."""

    def __init__(self, N: int, p: int, nprob: int, device=None, tuned: bool = True, **params):
        self.lib = capi.load()
        self.N, self.p, self.nprob = int(N), int(p), int(nprob)
        self.n = self.N * self.N
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.params = capi.default_params(**params)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            capi.check(self.lib.sbp_create_batch(self.N, self.p, self.nprob, C.byref(self.params), C.byref(h)),
                       "sbp_create_batch")
        self._h = h
        nl = C.c_int()
        self.lib.sbp_get_info(self._h, None, None, C.byref(nl))
        self.nlevels = nl.value
        self.stream = None      # None: torch's current stream at each call
        if tuned and not (params.get("tile_i") or params.get("tile_j")):
            self.apply_tuned_tiles()

    def close(self):
        if getattr(self, "_h", None):
            self.lib.sbp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind_stream(self):
        cur = torch.cuda.current_stream(self.device)
        s = self.stream if self.stream is not None else cur
        if s != cur:
            s.wait_stream(cur)   # operands produced on the caller's stream are ready first
        self.lib.sbp_set_stream(self._h, C.c_void_p(s.cuda_stream))

    def _join(self):
        """Orders the caller's stream after this handle's launches."""
        if self.stream is not None:
            cur = torch.cuda.current_stream(self.device)
            if cur != self.stream:
                cur.wait_stream(self.stream)

    def _chk(self, t, dtype=torch.float64, name="array", n=None):
        n = self.n * self.nprob if n is None else n
        if t.device != self.device or not t.is_contiguous() or t.numel() != n or t.dtype != dtype:
            raise ValueError(f"{name}: need a contiguous {dtype} tensor of {n} elements on {self.device}")
        return t

    def level_size(self, level):
        N = self.N
        for _ in range(level):
            N = (N + 1) // 2
        return N

    def apply_tuned_tiles(self, cache=None):
        from . import autotune
        c = autotune.load_cache() if cache is None else cache
        for level in range(max(1, self.nlevels)):
            Nl = self.level_size(level)
            for dt, tag in ((capi.SBP_F64, "f64"), (capi.SBP_F32, "f32")):
                t = autotune.lookup(self.p, tag, Nl, c)
                if t:
                    fl = capi.SBP_ALIGNED if len(t) > 2 and t[2] else 0
                    capi.check(self.lib.sbp_set_tiles(self._h, level, dt | fl, t[0], t[1]), "sbp_set_tiles")

    def _abstol(self, abstol):
        if isinstance(abstol, (int, float)):
            abstol = [float(abstol)] * self.nprob
        if len(abstol) != self.nprob:
            raise ValueError(f"abstol: need {self.nprob} values")
        return (C.c_double * self.nprob)(*[float(a) for a in abstol])

    def apply(self, u, out=None, sign=1, level=0):
        """out = sign * A u for every problem (u: nprob * N_l^2 elements)."""
        n = self.level_size(level) ** 2 * self.nprob
        self._chk(u, u.dtype, "u", n)
        out = torch.empty_like(u) if out is None else self._chk(out, u.dtype, "out", n)
        self._bind_stream()
        capi.check(self.lib.sbp_apply(self._h, level, _ptr(u), _ptr(out), sign, _DT[u.dtype]), "sbp_apply")
        self._join()
        return out

    def apply_dot(self, u, out=None):
        """out = A u and the per-problem u_q' A u_q (FP64)."""
        self._chk(u, name="u")
        out = torch.empty_like(u) if out is None else self._chk(out, name="out")
        d = (C.c_double * self.nprob)()
        self._bind_stream()
        capi.check(self.lib.sbp_apply_dot(self._h, _ptr(u), _ptr(out), d), "sbp_apply_dot")
        return out, list(d)

    def slot(self, t, q, level=0):
        """Problem q's part of a stacked grid function (a view)."""
        n = self.level_size(level) ** 2
        return t[q * n:(q + 1) * n]

    def assemble_rhs(self, f, gW, gE, gS, gN, out):
        """Writes one problem's right-hand side into `out` (N*N FP64 elements,
        e.g. a slot view of a stacked array)."""
        for t in (f, gW, gE, gS, gN, out):
            if t is not None and (t.device != self.device or not t.is_contiguous() or t.dtype != torch.float64):
                raise ValueError("assemble_rhs: need contiguous FP64 tensors on the handle's device")
        if out.numel() != self.n or (f is not None and f.numel() != self.n):
            raise ValueError("assemble_rhs: f and out need N*N elements")
        self._bind_stream()
        capi.check(self.lib.sbp_assemble_rhs(self._h, _ptr(f), _ptr(gW), _ptr(gE), _ptr(gS), _ptr(gN), _ptr(out)),
                   "sbp_assemble_rhs")
        self._join()
        return out

    def hnorm(self, v):
        """sqrt(v' H~ v) of one problem's grid function."""
        if v.numel() != self.n or not v.is_contiguous():
            raise ValueError("hnorm: need N*N contiguous elements")
        o = C.c_double()
        self._bind_stream()
        capi.check(self.lib.sbp_hnorm(self._h, _ptr(v), C.byref(o)), "sbp_hnorm")
        return o.value

    def vcycle(self, r, out=None):
        self._chk(r, r.dtype, "r")
        out = torch.empty_like(r) if out is None else self._chk(out, r.dtype, "out")
        self._bind_stream()
        capi.check(self.lib.sbp_vcycle(self._h, _ptr(r), _ptr(out), _DT[r.dtype]), "sbp_vcycle")
        self._join()
        return out

    def _results(self):
        k = self.nprob
        return (C.c_int * k)(), (C.c_int * k)(), (C.c_double * k)(), (C.c_double * k)(), (C.c_int * k)()

    def mgpcg(self, b, x=None, abstol=1e-8, maxiter=500, raise_on_fail=True):
        """FP64 MG-PCG of every problem from x (zero when None)."""
        self._chk(b, name="b")
        x = torch.zeros_like(b) if x is None else self._chk(x, name="x")
        it, _, rn, sec, sts = self._results()
        tol = self._abstol(abstol)
        self._bind_stream()
        st = self.lib.sbp_mgpcg_batch(self._h, _ptr(b), _ptr(x), tol, int(maxiter), it, rn, sec, sts)
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_mgpcg_batch")
        return x, BatchResult(list(it), list(rn), list(sec), st, list(sts))

    def mgpcg_mixed(self, b, x=None, abstol=1e-8, max_outer=None, raise_on_fail=True):
        """FP32 inner MG-PCG inside FP64 refinement, every problem."""
        self._chk(b, name="b")
        x = torch.zeros_like(b) if x is None else self._chk(x, name="x")
        max_outer = self.params.max_outer if max_outer is None else max_outer
        oi, ii, rn, sec, sts = self._results()
        tol = self._abstol(abstol)
        self._bind_stream()
        st = self.lib.sbp_mgpcg_mixed_batch(self._h, _ptr(b), _ptr(x), tol, int(max_outer), oi, ii, rn, sec, sts)
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_mgpcg_mixed_batch")
        return x, BatchResult(list(oi), list(rn), list(sec), st, list(sts), list(ii))


class RaggedBatch:
    """Problems of one order p and uneven sides in one ragged handle
    (sbp_create_ragged): the multigrid levels are shared across sizes, so
    every V-cycle pass is one launch over all the problems present on its
    level.  Problems are held largest first (``rank``); grid functions are
    ragged tensors, problem of rank k at element offset ``offset[k]``."""

    def __init__(self, p: int, sides, device=None, tuned: bool = True, **params):
        self.lib = capi.load()
        self.p = int(p)
        sides = [int(N) for N in sides]
        # rank -> index in `sides` (stable, largest first)
        self.order = sorted(range(len(sides)), key=lambda q: -sides[q])
        self.rank = [0] * len(sides)
        for k, q in enumerate(self.order):
            self.rank[q] = k
        self.sides = [sides[q] for q in self.order]
        self.nprob = len(sides)
        self.offset = [0]
        for N in self.sides:
            self.offset.append(self.offset[-1] + N * N)
        self.total = self.offset[-1]
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.params = capi.default_params(**params)
        h = C.c_void_p()
        arr = (C.c_int * self.nprob)(*self.sides)
        with torch.cuda.device(self.device):
            capi.check(self.lib.sbp_create_ragged(self.p, self.nprob, arr, C.byref(self.params), C.byref(h)),
                       "sbp_create_ragged")
        self._h = h
        nl = C.c_int()
        self.lib.sbp_get_info(self._h, None, None, C.byref(nl))
        self.nlevels = nl.value
        self.stream = None
        if tuned and not (params.get("tile_i") or params.get("tile_j")):
            self.apply_tuned_tiles()

    close = SbpBatch.close
    __del__ = SbpBatch.__del__
    _bind_stream = SbpBatch._bind_stream
    _join = SbpBatch._join
    _results = SbpBatch._results

    def level_size(self, level):
        N = self.sides[0]
        for _ in range(level):
            N = (N + 1) // 2
        return N

    apply_tuned_tiles = SbpBatch.apply_tuned_tiles

    def view(self, t, k):
        """The slice of ragged tensor t that holds the problem of rank k."""
        return t[self.offset[k]:self.offset[k + 1]]

    def _chk(self, t, name):
        if t.device != self.device or not t.is_contiguous() or t.numel() != self.total or t.dtype != torch.float64:
            raise ValueError(f"{name}: need a contiguous float64 tensor of {self.total} elements on {self.device}")
        return t

    def _abstol(self, abstol):
        if isinstance(abstol, (int, float)):
            abstol = [float(abstol)] * self.nprob
        if len(abstol) != self.nprob:
            raise ValueError(f"abstol: need {self.nprob} values (by rank)")
        return (C.c_double * self.nprob)(*[float(a) for a in abstol])

    def mgpcg(self, b, x=None, abstol=1e-8, maxiter=500, raise_on_fail=True):
        """FP64 MG-PCG of every problem (b, x ragged, abstol by rank)."""
        self._chk(b, "b")
        x = torch.zeros_like(b) if x is None else self._chk(x, "x")
        it, _, rn, sec, sts = self._results()
        tol = self._abstol(abstol)
        self._bind_stream()
        st = self.lib.sbp_mgpcg_batch(self._h, _ptr(b), _ptr(x), tol, int(maxiter), it, rn, sec, sts)
        self._join()
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_mgpcg_batch")
        return x, BatchResult(list(it), list(rn), list(sec), st, list(sts))

    def mgpcg_mixed(self, b, x=None, abstol=1e-8, max_outer=None, raise_on_fail=True):
        """FP32 inner MG-PCG inside FP64 refinement, every problem."""
        self._chk(b, "b")
        x = torch.zeros_like(b) if x is None else self._chk(x, "x")
        max_outer = self.params.max_outer if max_outer is None else max_outer
        oi, ii, rn, sec, sts = self._results()
        tol = self._abstol(abstol)
        self._bind_stream()
        st = self.lib.sbp_mgpcg_mixed_batch(self._h, _ptr(b), _ptr(x), tol, int(max_outer), oi, ii, rn, sec, sts)
        self._join()
        if raise_on_fail and st not in (capi.SBP_OK, capi.SBP_ERR_NOT_CONVERGED):
            capi.check(st, "sbp_mgpcg_mixed_batch")
        return x, BatchResult(list(oi), list(rn), list(sec), st, list(sts), list(ii))


@dataclass
class _Group:
    p: int
    members: list                  # problem indices (submission order)
    op: RaggedBatch | None = None
    stream: torch.cuda.Stream | None = None
    b: torch.Tensor | None = None
    x: torch.Tensor | None = None
    result: BatchResult | None = None
    error: BaseException | None = None
    extra: dict = field(default_factory=dict)


class UnevenBatch:
    """Problems of mixed side and order, solved together.

    specs: sequence of (N, p).  One ragged handle per order holds all of that
    order's problems (shared multigrid levels: one launch per pass over every
    problem present on a level); the orders run concurrently, each from its
    own host thread on its own HIP stream (default priority), replaying its
    own captured iteration graph.  ``set_rhs`` fills problem q's right-hand
    side and ``solve`` runs every problem."""

    def __init__(self, specs, device=None, tuned=True, **params):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.specs = [(int(N), int(p)) for N, p in specs]
        self.params = params
        orders = {}
        for q, (N, p) in enumerate(self.specs):
            orders.setdefault(p, []).append(q)
        self.groups = []
        self.where = [None] * len(self.specs)    # problem -> (group, rank)
        # the order with the most points first: its launches are issued first
        for p, members in sorted(orders.items(), key=lambda kv: -sum(self.specs[q][0] ** 2 for q in kv[1])):
            g = _Group(p, members)
            g.op = RaggedBatch(p, [self.specs[q][0] for q in members], device=self.device, tuned=tuned, **params)
            g.stream = torch.cuda.Stream(device=self.device)
            g.op.stream = g.stream
            g.b = torch.zeros(g.op.total, dtype=torch.float64, device=self.device)
            g.x = torch.zeros(g.op.total, dtype=torch.float64, device=self.device)
            for i, q in enumerate(members):
                self.where[q] = (len(self.groups), g.op.rank[i])
            self.groups.append(g)
        self._single = {}

    def __len__(self):
        return len(self.specs)

    def _view(self, t_of_group, q):
        gi, k = self.where[q]
        g = self.groups[gi]
        return g.op.view(t_of_group(g), k)

    def rhs(self, q):
        """Problem q's right-hand-side slot (a view into its order's ragged tensor)."""
        return self._view(lambda g: g.b, q)

    def solution(self, q):
        return self._view(lambda g: g.x, q)

    def set_rhs(self, q, b):
        self.rhs(q).copy_(b)

    def single(self, q):
        """A single-problem handle of problem q's side and order (for its
        right-hand side and norms; cached per (N, p))."""
        from .solver import SbpPoisson
        key = self.specs[q]
        if key not in self._single:
            self._single[key] = SbpPoisson(key[0], key[1], device=self.device, tuned=False)
        return self._single[key]

    def hnorm(self, q, v):
        return self.single(q).hnorm(v)

    def set_fourier(self, q, seed):
        """Problem q's right-hand side from the seeded Fourier-mode data."""
        from .data import fourier_fields
        N, _ = self.specs[q]
        self.single(q).assemble_rhs(*fourier_fields(N, seed, self.device), out=self.rhs(q))

    def solve(self, abstol, mixed=True, maxiter=500, max_outer=None, zero_x=True):
        """Solves every problem; abstol[q] per problem.  Returns a list of
        per-problem dicts (iters, resnorm, seconds, status[, inner_iters])."""
        torch.cuda.synchronize(self.device)

        def run(g):
            try:
                with torch.cuda.stream(g.stream):
                    if zero_x:
                        g.x.zero_()
                    tol = [0.0] * len(g.members)
                    for i, q in enumerate(g.members):
                        tol[g.op.rank[i]] = abstol[q]
                    if mixed:
                        _, g.result = g.op.mgpcg_mixed(g.b, g.x, tol, max_outer, raise_on_fail=False)
                    else:
                        _, g.result = g.op.mgpcg(g.b, g.x, tol, maxiter, raise_on_fail=False)
            except BaseException as e:   # re-raised on the calling thread
                g.error = e

        threads = [threading.Thread(target=run, args=(g,)) for g in self.groups]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        torch.cuda.synchronize(self.device)
        for g in self.groups:
            if g.error is not None:
                e, g.error = g.error, None
                raise e
        out = [None] * len(self.specs)
        for g in self.groups:
            r = g.result
            for i, q in enumerate(g.members):
                k = g.op.rank[i]
                N, p = self.specs[q]
                d = {"N": N, "p": p, "iters": r.iters[k], "resnorm": r.resnorm[k],
                     "seconds": r.seconds[k], "status": r.statuses[k]}
                if r.inner_iters is not None:
                    d["inner_iters"] = r.inner_iters[k]
                out[q] = d
        return out

    def close(self):
        for g in self.groups:
            if g.op is not None:
                g.op.close()
        for h in self._single.values():
            h.close()
        self._single = {}
