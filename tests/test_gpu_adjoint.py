"""Adjoint gradients of L = 1/2 (u - u_obs)' H~ (u - u_obs), u = A^-1 b(f, g):
against the CPU oracle's gradient (direct solves), against central finite
differences along seeded random directions (step 1e-6), and the autograd
Function against the fused C entry point."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from oracle import sbp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def data(level, seed):
    f, gW, gE, gS, gN = O.fourier_data(level, seed)
    N = 2 ** level + 1
    rng = np.random.default_rng(seed)
    uo_f, *uo_g = O.fourier_data(level, seed + 1)
    A = O.assemble_operator(level, 4)
    u_obs = spla.spsolve(A.tocsc(), O.assemble_rhs(level, 4, uo_f, *uo_g)) + 1e-3 * rng.standard_normal(N * N)
    return (f, gW, gE, gS, gN), u_obs


def oracle_grad(level, p, fg, u_obs):
    A = O.assemble_operator(level, p)
    b = O.assemble_rhs(level, p, *fg)
    u = spla.spsolve(A.tocsc(), b)
    hd = O.h_tilde_diag(level, p)
    L = 0.5 * float((u - u_obs) @ (hd * (u - u_obs)))
    lam = spla.spsolve(A.tocsc(), hd * (u - u_obs))
    return L, O.rhs_transpose(level, p, lam)


@pytest.mark.parametrize("p", [2, 4])
def test_adjoint_grad_matches_oracle(p):
    from sc_2022_amd.solver import SbpPoisson
    level = 6
    N = 2 ** level + 1
    fg, u_obs = data(level, 3)
    Lo, go = oracle_grad(level, p, fg, u_obs)
    h = SbpPoisson(N, p)
    b = h.assemble_rhs(*[T(v) for v in fg])
    u, _ = h.mgpcg(b, abstol=1e-14 * float(b.norm()))
    L, lam, *grads, it = h.adjoint_grad(u, T(u_obs))
    assert abs(L - Lo) < 1e-10 * abs(Lo)
    for g, r in zip(grads, go):
        g = g.cpu().numpy()
        assert np.abs(g - r).max() < 1e-9 * np.abs(r).max()


def observed(N, seed):
    """u_obs of the differentiation workload: a seeded random Fourier field
    (the source-type field of seed + 1) plus N(0, 1e-3) noise, made on the
    host so that the GPU and the C oracle see the same values."""
    from oracle import c_oracle as CO
    field = CO.fourier_fields(N, seed + 1)[0]
    return field + 1e-3 * np.random.default_rng(seed).standard_normal(N * N)


def test_adjoint_grad_2049_against_c_oracle():
    """The differentiation workload at its size (2049^2, p = 4): the GPU's
    gradient (GPU forward solve, adjoint MG-PCG, transposed data map) against
    the C oracle's FP64 gradient computed end to end on the host (its own
    right-hand side, forward MG-PCG and adjoint MG-PCG to 1e-14)."""
    from oracle import c_oracle as CO
    from sc_2022_amd.data import fourier_fields
    from sc_2022_amd.solver import SbpPoisson
    level, p, seed = 11, 4, 2022
    N = 2 ** level + 1
    CO.set_threads(16)
    u_obs = observed(N, seed)
    h = SbpPoisson(N, p)
    b = h.assemble_rhs(*fourier_fields(N, seed, DEV))
    u, _ = h.mgpcg(b, abstol=1e-14 * float(b.norm()), maxiter=100)
    L, lam, *grads, it = h.adjoint_grad(u, T(u_obs))
    bh = CO.fourier_rhs(p, N, seed)
    uh, _, _ = CO.solve_reference(p, N, bh)
    Lo, lam_o, *go, it_o = CO.adjoint_grad(p, N, uh, u_obs, rtol=1e-14)
    # both sides solve to ~1e-14 relative residual; L and the gradients are
    # then equal to ~1e-10 (the differences of two independent FP64 solves
    # at kappa ~ 1e6, amplified by L's cancellation u - u_obs)
    assert abs(L - Lo) <= 1e-9 * abs(Lo), (L, Lo)
    for g, r in zip(grads, go):
        g = g.cpu().numpy()
        assert np.abs(g - r).max() <= 1e-9 * np.abs(r).max()


@pytest.mark.parametrize("level,ndir", [(7, 8), (11, 32)])
def test_adjoint_grad_finite_differences(level, ndir):
    """Central differences (step 1e-6) along seeded random directions in (f, g),
    u_obs a seeded Fourier field plus N(0, 1e-3) noise; at 2049^2 (p = 4) along
    32 directions as the differentiation workload asks."""
    from sc_2022_amd.solver import SbpPoisson
    p = 4
    N = 2 ** level + 1
    h = SbpPoisson(N, p)
    f, gW, gE, gS, gN = [T(v) for v in O.fourier_data(level, 5)]
    u_obs = T(observed(N, 5))
    def loss(fv, gv):
        b = h.assemble_rhs(fv, *gv)
        u, _ = h.mgpcg(b, abstol=1e-14 * float(b.norm()))
        return u, 0.5 * h.hnorm(u - u_obs) ** 2

    u, L0 = loss(f, (gW, gE, gS, gN))
    L, lam, df, dW, dE, dS, dN, it = h.adjoint_grad(u, u_obs)
    assert abs(L - L0) < 1e-12 * abs(L0)
    gen = torch.Generator(DEV).manual_seed(123)
    eps = 1e-6
    worst = 0.0
    for k in range(ndir):
        d = [torch.randn(t.numel(), dtype=torch.float64, device=DEV, generator=gen) for t in (f, gW, gE, gS, gN)]
        # scale the perturbation to the data's size
        d = [di * float(t.norm()) / float(di.norm()) for di, t in zip(d, (f, gW, gE, gS, gN))]
        _, Lp = loss(f + eps * d[0], [g + eps * dg for g, dg in zip((gW, gE, gS, gN), d[1:])])
        _, Lm = loss(f - eps * d[0], [g - eps * dg for g, dg in zip((gW, gE, gS, gN), d[1:])])
        fd = (Lp - Lm) / (2 * eps)
        ad = sum(float(gr @ di) for gr, di in zip((df, dW, dE, dS, dN), d))
        # L is quadratic in (f, g): central differences are exact up to the
        # solves' 1e-14 residual and rounding, amplified by 1/eps
        assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-9 * abs(L0), (k, fd, ad)
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"FD level {level}: {ndir} directions, worst relative difference {worst:.2e}")


def test_autograd_function_matches_fused_entry():
    from sc_2022_amd.solver import SbpPoisson, poisson_solve
    level, p = 7, 4
    N = 2 ** level + 1
    h = SbpPoisson(N, p)
    fg = [T(v).requires_grad_(True) for v in O.fourier_data(level, 8)]
    u_obs = T(np.random.default_rng(2).standard_normal(N * N) * 1e-2)
    hd = T(O.h_tilde_diag(level, p))
    u = poisson_solve(h, *fg)
    L = 0.5 * ((u - u_obs) * hd * (u - u_obs)).sum()
    L.backward()
    Lc, lam, *grads, it = h.adjoint_grad(u.detach(), u_obs)
    assert abs(float(L.detach()) - Lc) < 1e-12 * abs(Lc)
    for t, g in zip(fg, grads):
        assert torch.allclose(t.grad, g, rtol=1e-9, atol=1e-9 * float(g.abs().max()))


def test_autograd_function_with_absent_data():
    """Data passed as None (zero) get no gradient and the present ones the
    same gradient as with explicit zeros."""
    from sc_2022_amd.solver import SbpPoisson, poisson_solve
    level, p = 6, 4
    N = 2 ** level + 1
    h = SbpPoisson(N, p)
    f, gW, gE, gS, gN = (T(v) for v in O.fourier_data(level, 4))
    hd = T(O.h_tilde_diag(level, p))
    f1, gS1, gN1 = (t.clone().requires_grad_(True) for t in (f, gS, gN))
    u = poisson_solve(h, f1, None, None, gS1, gN1)
    (0.5 * (u * hd * u).sum()).backward()
    full = [t.clone().requires_grad_(True) for t in (f, torch.zeros_like(gW), torch.zeros_like(gE), gS, gN)]
    u2 = poisson_solve(h, *full)
    (0.5 * (u2 * hd * u2).sum()).backward()
    assert torch.equal(u, u2)
    for a, b in ((f1, full[0]), (gS1, full[3]), (gN1, full[4])):
        assert torch.allclose(a.grad, b.grad, rtol=1e-12, atol=0)
