"""Batched solves of many independent SBP-SAT Poisson problems.

``SbpBatch`` wraps one stacked handle (sbp_create_batch): nprob problems of
the same side N and order p whose grid functions sit back to back in one
device array, problem q at element offset q*N*N.  Every kernel launch covers
all of them; each problem has its own tolerance and device-side convergence
mask, so a converged problem's tiles fall through while the others iterate.

``UnevenBatch`` solves problems of mixed (N, p): it groups them by (N, p),
keeps one stacked handle per group, and drives the groups concurrently, each
from its own host thread on its own HIP stream (default priority), so the
small groups' launches overlap the large groups' on the GPU.  Per-problem
times are device wall-clock stamps from the group's solve start to that
problem's convergence.

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
                    capi.check(self.lib.sbp_set_tiles(self._h, level, dt, t[0], t[1]), "sbp_set_tiles")

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


@dataclass
class _Group:
    N: int
    p: int
    members: list                  # problem indices in submission order
    op: SbpBatch | None = None
    stream: torch.cuda.Stream | None = None
    b: torch.Tensor | None = None
    x: torch.Tensor | None = None
    result: BatchResult | None = None
    error: BaseException | None = None
    extra: dict = field(default_factory=dict)


class UnevenBatch:
    """Problems of mixed side and order, solved together.

    specs: sequence of (N, p).  The groups' handles, right-hand-side and
    solution stacks are allocated once here; ``set_rhs`` fills problem q's
    right-hand side and ``solve`` runs every group concurrently."""

    def __init__(self, specs, device=None, tuned=True, **params):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.specs = [(int(N), int(p)) for N, p in specs]
        groups = {}
        for q, key in enumerate(self.specs):
            groups.setdefault(key, []).append(q)
        self.groups = []
        self.where = [None] * len(self.specs)    # problem -> (group, slot)
        # largest groups first: their handles are launched first
        for (N, p), members in sorted(groups.items(), key=lambda kv: -kv[0][0] ** 2 * len(kv[1])):
            g = _Group(N, p, members)
            g.op = SbpBatch(N, p, len(members), device=self.device, tuned=tuned, **params)
            g.stream = torch.cuda.Stream(device=self.device)
            g.op.stream = g.stream
            n = N * N * len(members)
            g.b = torch.zeros(n, dtype=torch.float64, device=self.device)
            g.x = torch.zeros(n, dtype=torch.float64, device=self.device)
            for slot, q in enumerate(members):
                self.where[q] = (len(self.groups), slot)
            self.groups.append(g)

    def __len__(self):
        return len(self.specs)

    def _view(self, t, q):
        gi, slot = self.where[q]
        g = self.groups[gi]
        n = g.N * g.N
        return t[slot * n:(slot + 1) * n]

    def rhs(self, q):
        """Problem q's right-hand-side slot (a view into its group's stack)."""
        return self._view(self.groups[self.where[q][0]].b, q)

    def solution(self, q):
        return self._view(self.groups[self.where[q][0]].x, q)

    def set_rhs(self, q, b):
        self.rhs(q).copy_(b)

    def op(self, q):
        """The stacked handle that holds problem q."""
        return self.groups[self.where[q][0]].op

    def set_fourier(self, q, seed):
        """Problem q's right-hand side from the seeded Fourier-mode data."""
        from .data import fourier_fields
        N, _ = self.specs[q]
        op = self.op(q)
        saved, op.stream = op.stream, None   # torch's current stream: ordered after the field generation
        try:
            op.assemble_rhs(*fourier_fields(N, seed, self.device), out=self.rhs(q))
        finally:
            op.stream = saved

    def solve(self, abstol, mixed=True, maxiter=500, max_outer=None, zero_x=True):
        """Solves every problem; abstol[q] per problem.  Returns a list of
        per-problem dicts (iters, resnorm, seconds, status[, inner_iters])."""
        torch.cuda.synchronize(self.device)

        def run(g):
            try:
                with torch.cuda.stream(g.stream):
                    if zero_x:
                        g.x.zero_()
                    tol = [abstol[q] for q in g.members]
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
                raise g.error
        out = [None] * len(self.specs)
        for g in self.groups:
            r = g.result
            for slot, q in enumerate(g.members):
                d = {"N": g.N, "p": g.p, "iters": r.iters[slot], "resnorm": r.resnorm[slot],
                     "seconds": r.seconds[slot], "status": r.statuses[slot]}
                if r.inner_iters is not None:
                    d["inner_iters"] = r.inner_iters[slot]
                out[q] = d
        return out

    def close(self):
        for g in self.groups:
            if g.op is not None:
                g.op.close()
