"""Batched solves on the GPU: a stacked handle must give every problem the
same iterates as a single-problem handle (bit for bit: the kernels, tiles and
reduction orders are the same, only the launch covers more problems), the
device-side convergence masks must freeze finished problems, and the uneven
(mixed N, p) driver must reproduce the single-problem solves and the oracle."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import sbp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rhs_for(N, p, seed):
    from sc_2022_amd.data import fourier_rhs
    from sc_2022_amd.solver import SbpPoisson
    op = SbpPoisson(N, p, device=DEV)
    b = fourier_rhs(op, seed)
    op.close()
    return b


def stack(vs):
    return torch.cat([v.reshape(-1) for v in vs]).contiguous()


@pytest.mark.parametrize("p", [2, 4])
def test_batch_apply_matches_single(p):
    from sc_2022_amd.batch import SbpBatch
    from sc_2022_amd.solver import SbpPoisson
    N, k = 65, 3
    g = torch.Generator(device=DEV).manual_seed(5)
    us = [torch.randn(N * N, dtype=torch.float64, device=DEV, generator=g) for _ in range(k)]
    bt = SbpBatch(N, p, k, device=DEV)
    one = SbpPoisson(N, p, device=DEV)
    out = bt.apply(stack(us))
    out32 = bt.apply(stack([u.float() for u in us]))
    outd, dots = bt.apply_dot(stack(us))
    for q in range(k):
        ref = one.apply(us[q])
        assert torch.equal(out[q * N * N:(q + 1) * N * N], ref)
        assert torch.equal(out32[q * N * N:(q + 1) * N * N], one.apply(us[q].float()))
        assert torch.equal(outd[q * N * N:(q + 1) * N * N], ref)
        assert dots[q] == pytest.approx(float(torch.dot(us[q], ref)), rel=1e-12)
    # coarser levels of the hierarchy as well
    for level in (1, 2):
        Nl = bt.level_size(level)
        ul = [u[: Nl * Nl].clone() for u in us]
        o = bt.apply(stack(ul), level=level)
        for q in range(k):
            assert torch.equal(o[q * Nl * Nl:(q + 1) * Nl * Nl], one.apply(ul[q], level=level))


@pytest.mark.parametrize("p", [2, 4])
def test_batch_mgpcg_matches_single(p):
    """Per-problem tolerances; iterates identical to single-problem solves."""
    from sc_2022_amd.batch import SbpBatch
    from sc_2022_amd.solver import SbpPoisson
    N = 129
    bs = [rhs_for(N, p, s) for s in (1, 2, 3)]
    rt = [1e-10, 1e-6, 1e-12]
    tol = [r * float(torch.linalg.vector_norm(b)) for r, b in zip(rt, bs)]
    bt = SbpBatch(N, p, 3, device=DEV)
    xb, res = bt.mgpcg(stack(bs), abstol=tol)
    assert res.status == 0 and res.statuses == [0, 0, 0]
    one = SbpPoisson(N, p, device=DEV)
    for q in range(3):
        x1, r1 = one.mgpcg(bs[q], abstol=tol[q])
        assert res.iters[q] == r1.iters
        assert res.resnorm[q] == r1.resnorm
        assert torch.equal(xb[q * N * N:(q + 1) * N * N], x1)
        assert res.seconds[q] > 0
    # looser tolerance stops earlier
    assert res.iters[1] < res.iters[0] < res.iters[2]


@pytest.mark.parametrize("p", [2, 4])
def test_batch_mixed_matches_single(p):
    from sc_2022_amd.batch import SbpBatch
    from sc_2022_amd.solver import SbpPoisson
    N = 129
    bs = [rhs_for(N, p, s) for s in (4, 5)]
    tol = [1e-11 * float(torch.linalg.vector_norm(b)) for b in bs]
    bt = SbpBatch(N, p, 2, device=DEV)
    xb, res = bt.mgpcg_mixed(stack(bs), abstol=tol)
    assert res.status == 0
    one = SbpPoisson(N, p, device=DEV)
    for q in range(2):
        x1, r1 = one.mgpcg_mixed(bs[q], abstol=tol[q])
        assert (res.iters[q], res.inner_iters[q]) == (r1.iters, r1.inner_iters)
        assert res.resnorm[q] == r1.resnorm
        assert torch.equal(xb[q * N * N:(q + 1) * N * N], x1)


def test_batch_masks_freeze_finished_problems():
    """A zero right-hand side converges at iteration 0 and its iterate is never
    touched; its neighbours still match their single solves; an unreachable
    tolerance reports not-converged for that problem only."""
    from sc_2022_amd.batch import SbpBatch
    from sc_2022_amd.solver import SbpPoisson
    N, p = 65, 4
    b1 = rhs_for(N, p, 9)
    b0 = torch.zeros_like(b1)
    tol1 = 1e-10 * float(torch.linalg.vector_norm(b1))
    bt = SbpBatch(N, p, 3, device=DEV)
    x0 = torch.full((3 * N * N,), 0.0, dtype=torch.float64, device=DEV)
    xb, res = bt.mgpcg(stack([b1, b0, b1]), x=x0, abstol=[tol1, 1e-300, 0.0], maxiter=12, raise_on_fail=False)
    assert res.iters[1] == 0 and res.statuses[1] == 0
    assert torch.count_nonzero(xb[N * N:2 * N * N]) == 0
    assert res.statuses[0] == 0
    assert res.statuses[2] == 7 and res.iters[2] == 12      # abstol 0 is never met
    assert res.status == 7
    one = SbpPoisson(N, p, device=DEV)
    x1, r1 = one.mgpcg(b1, abstol=tol1)
    assert res.iters[0] == r1.iters and torch.equal(xb[:N * N], x1)
    # the capped problem ran exactly 12 iterations of the same recurrence
    x12, r12 = one.mgpcg(b1, abstol=0.0, maxiter=12, raise_on_fail=False)
    assert torch.equal(xb[2 * N * N:], x12)


def test_batch_rejects_single_problem_entry_points():
    from sc_2022_amd import _capi as capi
    from sc_2022_amd.batch import SbpBatch
    bt = SbpBatch(33, 2, 2, device=DEV)
    lib = capi.load()
    b = torch.zeros(2 * 33 * 33, dtype=torch.float64, device=DEV)
    it, rn = C.c_int(), C.c_double()
    st = lib.sbp_mgpcg(bt._h, C.c_void_p(b.data_ptr()), C.c_void_p(b.data_ptr()), 1.0, 5, C.byref(it),
                       C.byref(rn), None)
    assert st == capi.SBP_ERR_INVALID_ARG
    n = C.c_int()
    assert lib.sbp_get_batch(bt._h, C.byref(n)) == 0 and n.value == 2
    h = C.c_void_p()
    assert lib.sbp_create_batch(33, 2, 0, None, C.byref(h)) == capi.SBP_ERR_INVALID_ARG


def test_uneven_batch_matches_single_and_oracle():
    """Mixed sides and orders in one call: each problem equals its single
    solve bit for bit and the direct solve of the oracle's operator."""
    from sc_2022_amd.batch import UnevenBatch
    from sc_2022_amd.data import batch_specs, fourier_fields
    from sc_2022_amd.solver import SbpPoisson
    specs = batch_specs(31, count=10, kmin=4, kmax=7)
    assert len({(N, p) for N, p, _ in specs}) >= 3
    ub = UnevenBatch([(N, p) for N, p, _ in specs], device=DEV)
    tol = []
    for q, (N, p, s) in enumerate(specs):
        b = rhs_for(N, p, s)
        ub.set_rhs(q, b)
        tol.append(1e-11 * float(torch.linalg.vector_norm(b)))
    out = ub.solve(tol, mixed=True)
    for q, (N, p, s) in enumerate(specs):
        assert out[q]["status"] == 0 and out[q]["seconds"] > 0
        one = SbpPoisson(N, p, device=DEV)
        x1, r1 = one.mgpcg_mixed(ub.rhs(q).clone(), abstol=tol[q])
        assert out[q]["iters"] == r1.iters and out[q]["inner_iters"] == r1.inner_iters
        assert torch.equal(ub.solution(q), x1)
        level = int(round(np.log2(N - 1)))
        A = O.assemble_operator(level, p)
        xo = spla.spsolve(A.tocsc(), ub.rhs(q).cpu().numpy())
        hd = O.h_tilde_diag(level, p)
        xg = ub.solution(q).cpu().numpy()
        assert O.h_norm(xg - xo, hd) <= 1e-9 * O.h_norm(xo, hd), (N, p)
    # FP64 batched path through the same driver
    out64 = ub.solve(tol, mixed=False)
    assert all(o["status"] == 0 for o in out64)
    ub.close()


@pytest.mark.parametrize("p", [2, 4])
def test_ragged_matches_single_bitwise(p):
    """A ragged handle (sides 129, 129, 65, 33, 33 sharing their multigrid
    levels) gives every problem the iterates of its single-problem solve, bit
    for bit, FP64 and mixed, with the same kernels (fuse_n = 0 on both: the
    fused-kernel choice depends on the points per launch)."""
    from sc_2022_amd.batch import RaggedBatch
    from sc_2022_amd.solver import SbpPoisson
    sides = [65, 129, 33, 129, 33]
    seeds = [11, 12, 13, 14, 15]
    rb = RaggedBatch(p, sides, device=DEV, fuse_n=0)
    assert rb.sides == [129, 129, 65, 33, 33] and rb.nlevels == 4
    b = torch.zeros(rb.total, dtype=torch.float64, device=DEV)
    bs = {}
    for q, (N, s) in enumerate(zip(sides, seeds)):
        bs[q] = rhs_for(N, p, s)
        rb.view(b, rb.rank[q]).copy_(bs[q])
    tol = [0.0] * len(sides)
    for q in range(len(sides)):
        tol[rb.rank[q]] = 1e-11 * float(torch.linalg.vector_norm(bs[q]))
    x64, r64 = rb.mgpcg(b, abstol=tol)
    xmx, rmx = rb.mgpcg_mixed(b, abstol=tol)
    assert r64.status == 0 and rmx.status == 0
    for q, N in enumerate(sides):
        k = rb.rank[q]
        one = SbpPoisson(N, p, device=DEV, fuse_n=0)
        x1, res1 = one.mgpcg(bs[q], abstol=tol[k])
        assert r64.iters[k] == res1.iters and r64.resnorm[k] == res1.resnorm
        assert torch.equal(rb.view(x64, k), x1), (q, N)
        x2, res2 = one.mgpcg_mixed(bs[q], abstol=tol[k])
        assert (rmx.iters[k], rmx.inner_iters[k]) == (res2.iters, res2.inner_iters)
        assert torch.equal(rb.view(xmx, k), x2), (q, N)
        one.close()
    rb.close()


def test_ragged_256_problem_batch_against_single_and_oracle():
    """The 256-problem uneven workload (k in 6..10, p in {2, 4}, seeded
    Fourier data) through one ragged handle per order: every problem
    converges; a seeded sample of 12 problems matches its single-problem
    solve (iterations within one, solutions within 1e-10) and the C oracle's
    FP64 solution (within 1e-10)."""
    from oracle import c_oracle
    from sc_2022_amd.batch import UnevenBatch
    from sc_2022_amd.data import batch_specs
    from sc_2022_amd.solver import SbpPoisson
    specs = batch_specs(2022, count=256)
    ub = UnevenBatch([(N, p) for N, p, _ in specs], device=DEV)
    assert len(ub.groups) == 2
    for q, (N, p, s) in enumerate(specs):
        ub.set_fourier(q, s)
    tol = [1e-12 * float(torch.linalg.vector_norm(ub.rhs(q))) for q in range(len(specs))]
    out = ub.solve(tol, mixed=True)
    assert all(o["status"] == 0 and o["seconds"] > 0 for o in out)
    c_oracle.set_threads(16)
    rng = np.random.default_rng(7)
    for q in sorted(rng.choice(len(specs), size=12, replace=False)):
        N, p, s = specs[q]
        one = SbpPoisson(N, p, device=DEV)
        x1, r1 = one.mgpcg_mixed(ub.rhs(q).clone(), abstol=tol[q])
        assert abs(out[q]["inner_iters"] - r1.inner_iters) <= 1, (q, out[q], r1)
        xq = ub.solution(q)
        assert one.hnorm(xq - x1) <= 1e-10 * one.hnorm(x1), (q, N, p)
        bh = c_oracle.fourier_rhs(p, N, s)
        xo = torch.as_tensor(c_oracle.solve_reference(p, N, bh)[0], device=DEV)
        assert one.hnorm(xq - xo) <= 1e-10 * one.hnorm(xo), (q, N, p)
        one.close()
    ub.close()


def test_ragged_handle_validation_and_rejections():
    from sc_2022_amd import _capi as capi
    lib = capi.load()
    h = C.c_void_p()
    arr = lambda *v: (C.c_int * len(v))(*v)
    assert lib.sbp_create_ragged(2, 2, arr(33, 65), None, C.byref(h)) == capi.SBP_ERR_INVALID_ARG   # not sorted
    assert lib.sbp_create_ragged(2, 2, arr(65, 40), None, C.byref(h)) == capi.SBP_ERR_INVALID_SIZE  # off the hierarchy
    assert lib.sbp_create_ragged(2, 0, arr(65), None, C.byref(h)) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_create_ragged(4, 3, arr(65, 33, 33), None, C.byref(h)) == capi.SBP_OK
    n = C.c_int()
    assert lib.sbp_get_batch(h, C.byref(n)) == 0 and n.value == 3
    u = torch.zeros(65 * 65 + 2 * 33 * 33, dtype=torch.float64, device=DEV)
    pu = C.c_void_p(u.data_ptr())
    assert lib.sbp_apply(h, 0, pu, pu, 1, 0) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_vcycle(h, pu, pu, 0) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_destroy(h) == capi.SBP_OK
