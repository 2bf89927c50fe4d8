"""GPU parity of the variable-coefficient operator (sbp_create_var; the
reference's variable_diagonal_sbp_D2 on both axes) against the oracle's
restatement: apply (both signs, with the dot), right-hand side and its
transpose, CG, the V-cycle and MG-PCG on the rediscretised hierarchy, the
constant-coefficient limit, and the manufactured-solution convergence rates."""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import sbp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def T(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def relmax(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def profiles(level, seed):
    t = O.grid(level)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.2, 1.0, 3)
    bx = 1.0 + a[0] * np.sin(3 * t + a[1]) ** 2 + t
    by = 0.5 + a[2] * t ** 2 + 0.3 * np.cos(2 * t) ** 2
    return bx, by


@pytest.fixture(scope="module")
def SP():
    from sc_2022_amd.solver import SbpPoisson
    return SbpPoisson


@pytest.mark.parametrize("p", [2, 4, 6])
@pytest.mark.parametrize("level", [5, 6, 8])
def test_var_apply_matches_oracle(SP, level, p):
    N = 2 ** level + 1
    bx, by = profiles(level, level * 10 + p)
    A = O.assemble_operator_var(level, p, bx, by)
    h = SP(N, p, coef=(bx, by))
    u = np.random.default_rng(level + p).standard_normal(N * N)
    Au = A @ u
    assert relmax(h.apply(T(u)).cpu().numpy(), Au) < 1e-13
    assert relmax(h.apply(T(u), sign=-1).cpu().numpy(), -Au) < 1e-13
    out, d = h.apply_dot(T(u))
    assert relmax(out.cpu().numpy(), Au) < 1e-13
    assert abs(d - u @ Au) < 1e-12 * abs(u @ Au)


@pytest.mark.parametrize("p", [2, 4, 6])
def test_var_constant_limit_and_rhs(SP, p):
    """b = 1 gives the constant-coefficient handle's apply and right-hand
    side; with varying b the right-hand side and its transpose match the
    oracle (Dirichlet data scaled by the face coefficient, Neumann data as
    fluxes)."""
    level = 6
    N = 2 ** level + 1
    one = np.ones(N)
    hv, hc = SP(N, p, coef=(one, one)), SP(N, p)
    u = np.random.default_rng(1).standard_normal(N * N)
    assert relmax(hv.apply(T(u)).cpu().numpy(), hc.apply(T(u)).cpu().numpy()) < 1e-13
    bx, by, f, gW, gE, gS, gN, uu = O.var_data(level)
    h = SP(N, p, coef=(bx, by))
    b = h.assemble_rhs(T(f), T(gW), T(gE), T(gS), T(gN)).cpu().numpy()
    assert relmax(b, O.assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN)) < 1e-13
    lam = np.random.default_rng(2).standard_normal(N * N)
    df, dgW, dgE, dgS, dgN = (t.cpu().numpy() for t in h.rhs_transpose(T(lam)))
    # <b(f, g), lam> = <f, df> + <gW, dgW> + ... (the map is linear in (f, g))
    lhs = O.assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN) @ lam
    rhs = f @ df + gW @ dgW + gE @ dgE + gS @ dgS + gN @ dgN
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("p", [2, 4, 6])
def test_var_cg_against_oracle(SP, p):
    level = 6
    N = 2 ** level + 1
    bx, by, f, gW, gE, gS, gN, u = O.var_data(level)
    A = O.assemble_operator_var(level, p, bx, by)
    b = O.assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN)
    tol = 1e-10 * np.linalg.norm(b)
    xo, ito, _ = O.cg_cpu(A, b, abstol=tol)
    h = SP(N, p, coef=(bx, by))
    x, res = h.cg(T(b), abstol=tol, history=True)
    assert res.converged and abs(res.iters - ito) <= max(2, 0.01 * ito), (res.iters, ito)
    xg = x.cpu().numpy()
    assert np.linalg.norm(A @ xg - b) < 1.000001 * tol
    assert np.linalg.norm(xg - xo) < 1e-7 * np.linalg.norm(xo)


@pytest.mark.parametrize("p,rate,levels", [(2, 2.0, (6, 7, 8)), (4, 4.2, (5, 6, 7)), (6, 5.6, (5, 6, 7))])
def test_var_manufactured_rate_on_gpu(SP, p, rate, levels):
    errs = []
    for level in levels:
        N = 2 ** level + 1
        bx, by, f, gW, gE, gS, gN, u = O.var_data(level)
        h = SP(N, p, coef=(bx, by))
        b = h.assemble_rhs(T(f), T(gW), T(gE), T(gS), T(gN))
        x, res = h.cg(b, abstol=1e-13 * float(torch.linalg.vector_norm(b)))
        assert res.converged
        errs.append(O.h_norm(x.cpu().numpy() - u, O.h_tilde_diag(level, p)))
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert all(abs(r - rate) < 0.25 for r in rates), (errs, rates)


def coarse_level(p):
    return 5 if p == 6 else 4               # p = 6 closures need 24 points: coarsest 33^2


@pytest.mark.parametrize("p", [2, 4, 6])
@pytest.mark.parametrize("level", [6, 7])
def test_var_vcycle_matches_oracle(SP, level, p):
    """One V-cycle (2 + 2 damped-Jacobi sweeps, full weighting, bilinear
    prolongation, dense coarsest solve) on the hierarchy rediscretised with
    the injected profiles, against the oracle's VarHierarchy."""
    N = 2 ** level + 1
    bx, by = profiles(level, 3 * level + p)
    h = SP(N, p, coef=(bx, by))
    assert h.nlevels == level - coarse_level(p) + 1
    hier = O.VarHierarchy(level, p, bx, by, coarse_level=coarse_level(p))
    r = np.random.default_rng(level + p).standard_normal(N * N)
    z = h.vcycle(T(r)).cpu().numpy()
    assert relmax(z, O.vcycle(hier, r)) < 1e-12


@pytest.mark.parametrize("p", [2, 4, 6])
@pytest.mark.parametrize("level", [7, 10])
def test_var_vcycle_fused_matches_passes(SP, level, p):
    """The two-sweep kernels (levels of at most fuse_n^2 points: at 1025^2
    every level, the finest with 64 x 8 blocks, the others 64 x 4; interior,
    face and shifted last blocks) against the pass-per-sweep V-cycle
    (fuse_n = 0), in FP64 and inside the FP32-V-cycle MG-PCG."""
    N = 2 ** level + 1
    bx, by = profiles(level, 5 * level + p)
    hf = SP(N, p, coef=(bx, by))
    hs = SP(N, p, coef=(bx, by), fuse_n=0)
    r = np.random.default_rng(7 * level + p).standard_normal(N * N)
    zf = hf.vcycle(T(r)).cpu().numpy()
    zs = hs.vcycle(T(r)).cpu().numpy()
    assert relmax(zf, zs) < 1e-13
    b = T(r)
    tol = 1e-10 * float(torch.linalg.vector_norm(b))
    xf, rf = hf.mgpcg(b, abstol=tol)
    xs, rs = hs.mgpcg(b, abstol=tol)
    assert rf.converged and rs.converged and rf.iters == rs.iters
    assert relmax(xf.cpu().numpy(), xs.cpu().numpy()) < 1e-10
    xf32, rf32 = hf.mgpcg_mixed(b, abstol=tol)
    xs32, rs32 = hs.mgpcg_mixed(b, abstol=tol)
    assert rf32.converged and rs32.converged and abs(rf32.iters - rs32.iters) <= 1
    assert relmax(xf32.cpu().numpy(), xs32.cpu().numpy()) < 1e-8


@pytest.mark.parametrize("p", [2, 4, 6])
def test_var_mgpcg_against_oracle_and_direct(SP, p):
    level = 7
    N = 2 ** level + 1
    bx, by, f, gW, gE, gS, gN, u = O.var_data(level)
    hier = O.VarHierarchy(level, p, bx, by, coarse_level=coarse_level(p))
    b = O.assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN)
    xo, ito, _ = O.mgpcg(hier, b, rtol=1e-10)
    h = SP(N, p, coef=(bx, by))
    x, res = h.mgpcg(T(b), abstol=1e-10 * np.linalg.norm(b))
    assert res.converged and abs(res.iters - ito) <= 1 and res.iters <= 16, (res.iters, ito)
    xd = spla.spsolve(hier.A[0].tocsc(), b)
    assert np.linalg.norm(x.cpu().numpy() - xd) < 1e-8 * np.linalg.norm(xd)


@pytest.mark.parametrize("p,rate,levels", [(2, 2.0, (9, 10, 11)), (4, 4.2, (6, 7, 8))])
def test_var_mgpcg_manufactured_rate_large(SP, p, rate, levels):
    """MG-PCG to round-off (p = 2 up to 2049^2; p = 4 on coarser grids, its
    error reaching round-off near 513^2), the H-norm error rate of the
    variable-coefficient discretisation, and a level-independent iteration
    count."""
    errs, its = [], []
    for level in levels:
        N = 2 ** level + 1
        bx, by, f, gW, gE, gS, gN, u = O.var_data(level)
        h = SP(N, p, coef=(bx, by))
        b = h.assemble_rhs(T(f), T(gW), T(gE), T(gS), T(gN))
        x, res = h.mgpcg(b, abstol=1e-12 * float(torch.linalg.vector_norm(b)))
        assert res.converged
        its.append(res.iters)
        errs.append(O.h_norm(x.cpu().numpy() - u, O.h_tilde_diag(level, p)))
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert all(abs(r - rate) < 0.25 for r in rates), (errs, rates)
    assert max(its) - min(its) <= 2 and max(its) <= 20, its


@pytest.mark.parametrize("p,level", [(2, 7), (4, 7), (6, 7), (4, 10)])
def test_var_mgpcg_fp32_vcycle(SP, p, level):
    """FP64 MG-PCG with the V-cycle in FP32 (sbp_mgpcg_mixed on a
    variable-coefficient handle): the FP64 residual reaches the same tolerance
    and the solution agrees with the FP64 V-cycle's, in about as many
    iterations (the FP32 preconditioner is a perturbation of order 1e-7 of the
    FP64 one)."""
    N = 2 ** level + 1
    bx, by, f, gW, gE, gS, gN, u = O.var_data(level)
    h = SP(N, p, coef=(bx, by))
    b = h.assemble_rhs(T(f), T(gW), T(gE), T(gS), T(gN))
    tol = 1e-11 * float(torch.linalg.vector_norm(b))
    x64, r64 = h.mgpcg(b, abstol=tol)
    x32, r32 = h.mgpcg_mixed(b, abstol=tol)
    assert r64.converged and r32.converged
    assert r32.iters <= r64.iters + 2, (r32.iters, r64.iters)
    assert r32.inner_iters == r32.iters
    assert relmax(x32.cpu().numpy(), x64.cpu().numpy()) < 1e-9
    Ax = torch.empty_like(b)
    h.apply(x32, Ax)
    assert float(torch.linalg.vector_norm(b - Ax)) < 1.01 * tol
    # a second call on the same handle reuses the FP32 level buffers
    x2, r2 = h.mgpcg_mixed(b, abstol=tol)
    assert r2.iters == r32.iters and relmax(x2.cpu().numpy(), x32.cpu().numpy()) < 1e-13


def test_var_mgpcg_fp32_vcycle_iteration_cap(SP):
    """With the V-cycle in FP32 the solve runs at most
    inner_maxiter * max(max_outer, 1) iterations and reports not-converged
    when the tolerance is out of reach."""
    from sc_2022_amd import _capi as capi
    N = 129
    bx, by = profiles(7, 1)
    h = SP(N, 4, coef=(bx, by), inner_maxiter=2)
    b = T(np.random.default_rng(3).standard_normal(N * N))
    tol = 1e-15 * float(torch.linalg.vector_norm(b))
    x, res = h.mgpcg_mixed(b, abstol=tol, max_outer=3, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_NOT_CONVERGED
    assert res.iters == 6 and res.inner_iters == 6
    x, res = h.mgpcg_mixed(b, abstol=tol, max_outer=0, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_NOT_CONVERGED and res.iters == 2


def test_var_handle_rejects_what_it_does_not_serve(SP):
    from sc_2022_amd import _capi as capi
    N = 65
    bx, by = profiles(6, 0)
    h = SP(N, 4, coef=(bx, by))
    with pytest.raises(Exception):
        h.diag()
    with pytest.raises(Exception):
        h.apply(T(np.zeros(N * N), torch.float32))
    with pytest.raises(Exception):
        h.vcycle(T(np.zeros(N * N), torch.float32))
    hr = SP(60, 4, coef=(np.ones(60), np.ones(60)))   # not 2^k + 1: no hierarchy
    assert hr.nlevels == 0
    x, res = hr.mgpcg(T(np.ones(60 * 60)), abstol=1e-8, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_INVALID_SIZE
    x, res = hr.mgpcg_mixed(T(np.ones(60 * 60)), abstol=1e-8, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_INVALID_SIZE
    with pytest.raises(Exception):
        SP(N, 4, coef=(bx, -by))            # coefficients must be positive
    with pytest.raises(Exception):
        SP(N, 8, coef=(bx, by))             # orders 2, 4 and 6 only
    with pytest.raises(Exception):
        SP(23, 6, coef=(np.ones(23), np.ones(23)))   # p = 6 closures need N >= 24
