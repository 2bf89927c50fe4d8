"""Pins the CPU oracle (oracle/sbp_oracle.py) against facts that do not come
from the oracle itself: polynomial exactness and the SBP property of the 1-D
operators, the literal restatement of the reference GPU kernel, the closed
tensor form, transfer-operator identities, direct solves, manufactured-solution
convergence rates and the adjoint (dot-product) identity.  Each check is
chosen so that a plausible slip (a dropped closure term, a wrong sign or index,
a transposed kron) fails it."""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import sbp_oracle as O


@pytest.mark.parametrize("p", [2, 4, 6, 8, 10])
def test_d2_polynomial_exactness(p):
    N = 40
    D2, S0, SN, HI, H, r = O.diagonal_sbp_D2(p, N, xc=(0, 1))
    D2 = D2.toarray()
    bm = {2: 1, 4: 4, 6: 6, 8: 8, 10: 11}[p]
    kb = p // 2 + 1                  # boundary accuracy degree
    for k in range(0, p + 2):
        v = r ** k
        exact = k * (k - 1) * r ** (k - 2) if k >= 2 else np.zeros_like(r)
        err = np.abs(D2 @ v - exact)
        assert err[bm:-bm].max() < 1e-7 * max(1, k ** 2), (p, k)
        if k <= kb:
            assert err.max() < 1e-7 * max(1, k ** 2), (p, k)


@pytest.mark.parametrize("p", [2, 4, 6, 8, 10])
def test_d2_sbp_property(p):
    """H D2 = -M + BS with M symmetric positive semidefinite and M 1 = 0
    (diagonal_sbp.jl:159-160)."""
    N = 30
    D2, S0, SN, HI, H, r = O.diagonal_sbp_D2(p, N, xc=(0, 1))
    BS = (SN - S0).toarray()
    M = BS - (H @ D2).toarray()
    assert np.abs(M - M.T).max() < 1e-9 * np.abs(M).max()
    assert np.abs(M @ np.ones(N + 1)).max() < 1e-9 * np.abs(M).max()
    ev = np.linalg.eigvalsh(0.5 * (M + M.T))
    assert ev.min() > -1e-9 * np.abs(ev).max()
    assert abs(H.diagonal().sum() - 1.0) < 1e-14
    # S0/SN approximate -/+ first derivatives at the ends: exact on linears
    # (BS row = outward normal derivative)
    assert abs((S0 @ r)[0] - 1.0) < 1e-12 and abs((SN @ r)[-1] - 1.0) < 1e-12
    assert abs((BS @ r ** 2)[0] - 0.0) < 1e-9 and abs((BS @ r ** 2)[-1] - 2.0) < 1e-9


@pytest.mark.parametrize("p", [2, 4, 6, 8, 10])
def test_d1_sbp_property(p):
    N = 30
    D, HI, H, r = O.diagonal_sbp_D1(p, N, xc=(0, 1))
    Q = (H @ D).toarray()
    B = np.zeros((N + 1, N + 1)); B[0, 0] = -1; B[-1, -1] = 1
    assert np.abs(Q + Q.T - B).max() < 1e-12
    for k in range(0, p // 2 + 1):
        v = r ** k
        exact = k * r ** (k - 1) if k >= 1 else np.zeros_like(r)
        assert np.abs(D @ v - exact).max() < 1e-9
    # D1's norm equals D2's norm (the assembly mixes them)
    _, _, _, _, H2, _ = O.diagonal_sbp_D2(p, N, xc=(0, 1))
    assert np.abs(H.diagonal() - H2.diagonal()).max() < 1e-15


def test_invalid_order_and_size():
    with pytest.raises(ValueError):
        O.diagonal_sbp_D2(6 + 1, 20)
    with pytest.raises(ValueError):
        O.diagonal_sbp_D2(4, 5)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [3, 4])
def test_operator_spd_and_closed_form(level, p):
    A = O.assemble_operator(level, p)
    Ad = A.toarray()
    assert np.abs(Ad - Ad.T).max() < 1e-13 * np.abs(Ad).max()
    assert np.linalg.eigvalsh(Ad).min() > 0
    C = O.closed_form_operator(level, p)
    assert abs(A - C).max() < 1e-13 * abs(A).max()
    # A commutes with the 180-degree reversal of the grid
    R = np.eye(A.shape[0])[::-1]
    assert np.abs(R @ Ad @ R - Ad).max() < 1e-13 * np.abs(Ad).max()


@pytest.mark.parametrize("level,p", [(4, 6), (5, 6), (4, 8), (5, 8), (5, 10)])
def test_high_order_operator_closed_form_and_definiteness(level, p):
    """p = 6, 8, 10 (diagonal_sbp.jl:210-408): symmetric, equal to the closed
    form, 180-degree symmetric; positive definite for p = 6, but indefinite
    for p = 8, 10 with the reference's Dirichlet penalty tau = 13/h (the 1-D
    Dirichlet factor needs tau above ~800 for p = 8, ~3000 for p = 10)."""
    A = O.assemble_operator(level, p)
    Ad = A.toarray()
    assert np.abs(Ad - Ad.T).max() < 1e-13 * np.abs(Ad).max()
    assert abs(A - O.closed_form_operator(level, p)).max() < 1e-13 * abs(A).max()
    R = np.eye(A.shape[0])[::-1]
    assert np.abs(R @ Ad @ R - Ad).max() < 1e-13 * np.abs(Ad).max()
    ev = np.linalg.eigvalsh(Ad).min()
    if p == 6:
        assert ev > 0
    else:
        assert ev < -1.0
        hw, m, k = O.one_d_factors(p, 33)
        E = np.zeros((33, 33)); E[0, 0] = E[-1, -1] = 1
        tau = {8: 1000, 10: 3000}[p]
        assert np.linalg.eigvalsh(k + (tau - 13) * E).min() > 0
        assert np.linalg.eigvalsh(k + (tau / 10 - 13) * E).min() < 0


@pytest.mark.parametrize("p,levels,lo,hi", [(6, (5, 6, 7, 8), 5.2, 5.8), (10, (5, 6, 7), 6.8, 7.8)])
def test_high_order_manufactured_convergence(p, levels, lo, hi):
    """H-norm error of the direct solve per halving: about 5.5 for p = 6
    (boundary order 3 + 2) and about 7.3 for p = 10 (boundary order 5 + 2;
    the p = 10 operator is indefinite but nonsingular).  Finer p = 10 grids
    reach round-off."""
    errs = []
    for level in levels:
        A = O.assemble_operator(level, p)
        f, gW, gE, gS, gN, u = O.manufactured_data(level)
        b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
        x = spla.spsolve(A.tocsc(), b)
        errs.append(O.h_norm(x - u, O.h_tilde_diag(level, p)))
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(len(levels) - 1)]
    assert all(lo < r < hi for r in rates), rates


@pytest.mark.parametrize("level", [3, 4, 5])
def test_reference_gpu_kernel_is_minus_A(level):
    """The literal restatement of matrix_free_A_full_GPU (p=2) equals -A on
    every unit vector: two independent texts of the reference agree."""
    N = 2 ** level + 1
    A = O.assemble_operator(level, 2).toarray()
    I = np.eye(N * N)
    K = np.column_stack([O.reference_gpu_apply_p2(I[:, c], N) for c in range(N * N)])
    assert np.abs(K + A).max() < 1e-12


@pytest.mark.parametrize("p,rate_lo,rate_hi,e129", [(2, 1.95, 2.05, 1.77e-5), (4, 3.6, 4.3, 2.58e-9)])
def test_manufactured_convergence(p, rate_lo, rate_hi, e129):
    errs = []
    for level in (4, 5, 6, 7):
        A = O.assemble_operator(level, p)
        f, gW, gE, gS, gN, u = O.manufactured_data(level)
        b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
        x = spla.spsolve(A.tocsc(), b)
        errs.append(O.h_norm(x - u, O.h_tilde_diag(level, p)))
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(3)]
    assert rate_lo < rates[-1] < rate_hi, rates
    assert abs(errs[-1] / e129 - 1) < 0.02
    # the reference's own manufactured solution converges too
    A = O.assemble_operator(6, p)
    Ab, b, Ht, _, _, u = O.assembling_matrix(6, p)
    x = spla.spsolve(Ab.tocsc(), b)
    assert O.h_norm(x - u, Ht.diagonal()) < 1e-3


@pytest.mark.parametrize("p", [2, 4])
def test_rhs_matches_manufactured_limit(p):
    """b built from the data of u equals A u up to the truncation error: the
    SAT data maps carry the right signs and scalings."""
    for level, tol in ((6, 5e-2), (7, 2.5e-2)):
        A = O.assemble_operator(level, p)
        f, gW, gE, gS, gN, u = O.manufactured_data(level)
        b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
        hd = O.h_tilde_diag(level, p)
        rel = np.linalg.norm((A @ u - b) / np.sqrt(hd)) / np.linalg.norm(b / np.sqrt(hd))
        assert rel < tol


@pytest.mark.parametrize("p", [2, 4])
def test_rhs_transpose_identity(p):
    level = 4
    N = 2 ** level + 1
    rng = np.random.default_rng(3)
    f, gW, gE, gS, gN = (rng.standard_normal(N * N), *rng.standard_normal((4, N)))
    lam = rng.standard_normal(N * N)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    df, dW, dE, dS, dN = O.rhs_transpose(level, p, lam)
    lhs = b @ lam
    rhs = f @ df + gW @ dW + gE @ dE + gS @ dS + gN @ dN
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)


def test_transfers():
    for Nc in (5, 9, 17):
        Nf = 2 * Nc - 1
        P = O.prolongation_2d(Nc)
        R = O.restriction_2d(Nf)
        assert abs(R - P.T / 4).max() < 1e-15
        # P exact on bilinear functions
        tc = np.linspace(0, 1, Nc); tf = np.linspace(0, 1, Nf)
        g = lambda a, b: 1 + 2 * a + 3 * b + 4 * a * b
        Ec = g(tc[:, None], tc[None, :]).ravel(order="F")
        Ef = g(tf[:, None], tf[None, :]).ravel(order="F")
        assert np.abs(P @ Ec - Ef).max() < 1e-13
        # full weighting keeps constants at interior coarse points
        rc = (R @ np.ones(Nf * Nf)).reshape(Nc, Nc, order="F")
        assert np.abs(rc[1:-1, 1:-1] - 1).max() < 1e-15
        assert abs(rc[0, 0] - 9 / 16) < 1e-15


@pytest.mark.parametrize("p", [2, 4])
def test_cg_cpu_against_direct(p):
    level = 4
    A = O.assemble_operator(level, p)
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    xd = spla.spsolve(A.tocsc(), b)
    x, it, norms = O.cg_cpu(A, b, abstol=1e-12 * np.linalg.norm(b))
    assert norms[-1] < 1e-12 * np.linalg.norm(b)
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)
    assert 10 < it < A.shape[0]


@pytest.mark.parametrize("p", [2, 4])
def test_mgpcg_oracle_against_direct(p):
    level = 6
    hier = O.Hierarchy(level, p)
    A = hier.A[0]
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    xd = spla.spsolve(A.tocsc(), b)
    x, it, norms = O.mgpcg(hier, b, rtol=1e-12)
    assert np.linalg.norm(x - xd) < 1e-9 * np.linalg.norm(xd)
    assert it < 40
    # the V-cycle is a symmetric operator (needed by PCG)
    n = A.shape[0]
    rng = np.random.default_rng(5)
    a, c = rng.standard_normal(n), rng.standard_normal(n)
    assert abs(a @ O.vcycle(hier, c) - c @ O.vcycle(hier, a)) < 1e-10 * abs(a @ O.vcycle(hier, c))


def test_cg_iteration_count_129_p2():
    """Expected count for the 129^2 CG case (relative residual 1e-8)."""
    A = O.assemble_operator(7, 2)
    f, gW, gE, gS, gN, u = O.manufactured_data(7)
    b = O.assemble_rhs(7, 2, f, gW, gE, gS, gN)
    x, it, norms = O.cg_cpu(A, b, abstol=1e-8 * np.linalg.norm(b))
    assert 540 <= it <= 580


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("three", [False, True])
def test_initial_guess_interpolation_bilinear(p, three):
    """A bilinear u is the discrete solution on every grid (D2 and the
    boundary derivative are exact on linears), so the nested solve's
    interpolated guess must reproduce it to round-off; the same guess rotated
    by 180 degrees (the reference's matrix-free variant reverses its guess)
    must not."""
    level = 6
    A, A2, A4 = (O.assemble_operator(level - k, p) for k in range(3))
    bs = []
    for k in range(3):
        f, gW, gE, gS, gN, u_k = O.bilinear_data(level - k)
        bs.append(O.assemble_rhs(level - k, p, f, gW, gE, gS, gN))
        if k == 0:
            u = u_k
    assert np.linalg.norm(bs[0] - A @ u) < 1e-13 * np.linalg.norm(bs[0])
    kw = dict(A4h=A4, b4h=bs[2]) if three else {}
    x, iters, norms, x0 = O.initial_guess_interpolation_cg(A, bs[0], A2, bs[1], 1e-10 * np.linalg.norm(bs[0]), **kw)
    assert np.abs(x0 - u).max() < 1e-12
    assert np.abs(x0[::-1] - u).max() > 0.5
    assert len(iters) == (2 if three else 1)


@pytest.mark.parametrize("p", [2, 4])
def test_initial_guess_interpolation_cuts_cg_iterations(p):
    """On the manufactured problem the nested solve converges to the direct
    solution in fewer fine-grid CG iterations than CG from zero."""
    level = 6
    A, A2 = O.assemble_operator(level, p), O.assemble_operator(level - 1, p)
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    f2, gW2, gE2, gS2, gN2, _ = O.manufactured_data(level - 1)
    b2 = O.assemble_rhs(level - 1, p, f2, gW2, gE2, gS2, gN2)
    tol = 1e-10 * np.linalg.norm(b)
    x, iters, norms, x0 = O.initial_guess_interpolation_cg(A, b, A2, b2, tol)
    xd = spla.spsolve(A.tocsc(), b)
    _, it0, _ = O.cg_cpu(A, b, abstol=tol)
    assert norms[-1] < tol and np.linalg.norm(A @ x - b) < tol * 1.000001
    assert np.linalg.norm(x - xd) < 1e-8 * np.linalg.norm(xd)
    assert iters[0] < it0
    # the guess is O(h^2)-close to the fine solution already
    hd = O.h_tilde_diag(level, p)
    assert O.h_norm(x0 - xd, hd) < 1e-3 * O.h_norm(xd, hd)


@pytest.mark.parametrize("p", [2, 4])
def test_smoothers_literal_and_fixed_point(p):
    """jacobi_sweeps equals the reference's literal update
    x <- omega P^-1 (b - Q x) + (1 - omega) x with P = diag(A), Q = A - P
    (jacobi_brittany!, level_2_multigrid_new.jl:335-344), sweep by sweep;
    modified_richardson equals x <- x + omega (b - A x) (:346-350); both leave
    the exact solution A^-1 b unchanged, and a wrong sign of the update (a
    plausible slip) moves it."""
    level = 4
    N = 2 ** level + 1
    A = O.assemble_operator(level, p).tocsr()
    rng = np.random.default_rng(4)
    b, x0 = rng.standard_normal(N * N), rng.standard_normal(N * N)
    Pd = A.diagonal()
    Q = A - sp.diags(Pd)
    om = 2.0 / 3.0
    x = x0.copy()
    for nu in range(1, 4):
        x = om * (b - Q @ x) / Pd + (1 - om) * x
        got = O.jacobi_sweeps(A, x0, b, nu, om)
        assert np.abs(got - x).max() <= 1e-13 * np.abs(x).max()
    xr = x0.copy()
    for _ in range(3):
        xr = xr + 0.15 * (b - A @ xr)
    assert np.abs(O.modified_richardson(A, x0, b) - xr).max() <= 1e-13 * np.abs(xr).max()
    xs = spla.spsolve(A.tocsc(), b)
    for sm in (O.jacobi_sweeps(A, xs, b, 3), O.modified_richardson(A, xs, b, 3)):
        assert np.abs(sm - xs).max() <= 1e-10 * np.abs(xs).max()
    wrong = xs - om * (b - A @ (xs + 1e-3)) / Pd
    assert np.abs(O.jacobi_sweeps(A, xs + 1e-3, b, 1, om) - wrong).max() > 1e-6


def test_three_level_richardson_restatement():
    """The reference's three-level cycle as written (initial_guess_cg.jl:244-301):
    with nu = 0 it returns zero exactly (the middle grid never receives its
    right-hand side, so its correction -- built from -A_2h v = 0 -- is zero,
    where a standard cycle would return P (P A_4h^-1 R) R b); it equals its
    algebraic statement as one matrix; it is linear but not symmetric, and
    as the MGCG preconditioner (p = 2) it still reaches the direct solution,
    in several times the two-level cycle's iterations."""
    level = 5
    N = 2 ** level + 1
    A, A2, A4 = (O.assemble_operator(level - k, 2) for k in range(3))
    rng = np.random.default_rng(9)
    b, c = rng.standard_normal(N * N), rng.standard_normal(N * N)
    z0, _ = O.three_level_richardson(A, A2, A4, b, nu=0)
    assert not np.any(z0)
    for nu in (1, 3):
        M = O.three_level_richardson_matrix(A, A2, A4, nu=nu)
        z, _ = O.three_level_richardson(A, A2, A4, b, nu=nu)
        assert np.abs(z - M @ b).max() <= 1e-11 * np.abs(z).max()
    M3 = lambda r: O.three_level_richardson(A, A2, A4, r)[0]
    assert np.abs(M3(b + 2 * c) - M3(b) - 2 * M3(c)).max() <= 1e-11 * np.abs(M3(c)).max()
    assert abs(c @ M3(b) - b @ M3(c)) > 1e-6 * abs(c @ M3(b))
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    rhs = O.assemble_rhs(level, 2, f, gW, gE, gS, gN)
    tol = 1e-10 * np.linalg.norm(rhs)
    x, it, norms = O.matrix_free_mgcg(A, rhs, M3, abstol=tol)
    _, it2, _ = O.matrix_free_mgcg(A, rhs, lambda r: O.two_level_richardson(A, A2, r)[0], abstol=tol)
    assert norms[-1] < tol and it > 2 * it2
    xd = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - xd) < 1e-8 * np.linalg.norm(xd)


def test_two_level_richardson_restatement():
    """The reference's two-level cycle: with nu = 0 it is P A_2h^-1 FW b
    exactly; it is a symmetric operator (same smoother before and after,
    FW = P^T / 4); as the MGCG preconditioner (p = 2) it converges to the
    direct solution."""
    level = 5
    N = 2 ** level + 1
    A, A2 = O.assemble_operator(level, 2), O.assemble_operator(level - 1, 2)
    rng = np.random.default_rng(8)
    b, c = rng.standard_normal(N * N), rng.standard_normal(N * N)
    z0, _ = O.two_level_richardson(A, A2, b, nu=0)
    ref = O.prolongation_2d((N + 1) // 2) @ spla.spsolve(A2.tocsc(), O.restriction_2d(N) @ b)
    assert np.abs(z0 - ref).max() < 1e-12 * np.abs(ref).max()
    M = lambda r: O.two_level_richardson(A, A2, r)[0]
    assert abs(c @ M(b) - b @ M(c)) < 1e-10 * abs(c @ M(b))
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    rhs = O.assemble_rhs(level, 2, f, gW, gE, gS, gN)
    x, it, norms = O.matrix_free_mgcg(A, rhs, M, abstol=1e-10 * np.linalg.norm(rhs))
    assert norms[-1] < 1e-10 * np.linalg.norm(rhs) and it < 60
    xd = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - xd) < 1e-8 * np.linalg.norm(xd)


@pytest.mark.parametrize("p", [2, 4, 6])
def test_variable_d2_pins(p):
    """variable_diagonal_sbp_D2 (diagonal_sbp.jl:470-740): B = 1 gives the
    constant-coefficient D2 exactly; for a varying B, M is symmetric positive
    semidefinite with M 1 = 0 (every b_k's weights cancel along a row), and D
    approximates (B u')' to order p inside and p/2 at the faces.  For p = 6 a
    higher frequency and finer grids keep the interior error above round-off."""
    N = 40
    D2 = O.diagonal_sbp_D2(p, N, xc=(0, 1))[0]
    Dv = O.variable_diagonal_sbp_D2(p, N, np.ones(N + 1), xc=(0, 1))[0]
    assert abs(Dv - D2).max() < 1e-13 * abs(D2).max()
    r = np.linspace(0, 1, N + 1)
    B = 1 + 0.5 * np.sin(3 * r) + r ** 2
    M = O.variable_diagonal_sbp_D2(p, N, B, xc=(0, 1))[5].toarray()
    assert np.abs(M - M.T).max() == 0.0
    assert np.abs(M.sum(axis=1)).max() < 1e-13 * np.abs(M).max()
    assert np.linalg.eigvalsh(M).min() > -1e-12 * np.abs(M).max()
    ei, eb = [], []
    bm = {2: 1, 4: 6, 6: 9}[p]
    kf, ns = (2, (32, 64, 128)) if p < 6 else (8, (64, 128))
    for n in ns:
        x = np.linspace(0, 1, n + 1)
        Bb = 1 + x ** 2
        D = O.variable_diagonal_sbp_D2(p, n, Bb, xc=(0, 1))[0]
        err = np.abs(D @ np.sin(kf * x) - (2 * kf * x * np.cos(kf * x) - kf * kf * Bb * np.sin(kf * x)))
        ei.append(err[2 * bm:-2 * bm].max())
        eb.append(err.max())
    ri = math.log2(ei[-2] / ei[-1]); rb = math.log2(eb[-2] / eb[-1])
    assert abs(ri - p) < 0.25, ri
    assert abs(rb - p / 2) < 0.25, rb


@pytest.mark.parametrize("p", [2, 4, 6])
def test_variable_2d_operator(p):
    """The 2-D variable-coefficient operator: SPD, equal to the constant one
    at b = 1, equal to the tensor form k(bx) (x) hw + hw (x) m(by), and the
    manufactured solution converges at rate 2 (p = 2), about 4.2 (p = 4) and
    about 5.6 (p = 6)."""
    level = 4 if p < 6 else 5
    bx, by = O.var_coefficients(level)
    A = O.assemble_operator_var(level, p, bx, by)
    Ad = A.toarray()
    assert np.abs(Ad - Ad.T).max() < 1e-13 * np.abs(Ad).max()
    assert np.linalg.eigvalsh(Ad).min() > 0
    one = np.ones(2 ** level + 1)
    assert abs(O.assemble_operator_var(level, p, one, one) - O.assemble_operator(level, p)).max() < 1e-13 * abs(A).max()
    N = 2 ** level + 1
    hw, mx, kx = O.one_d_factors_var(p, N, bx)
    _, my, ky = O.one_d_factors_var(p, N, by)
    C = np.kron(kx, np.diag(hw)) + np.kron(np.diag(hw), my)
    assert np.abs(C - Ad).max() < 1e-12 * np.abs(Ad).max()
    errs = []
    for lv in (5, 6, 7):
        bx, by, f, gW, gE, gS, gN, u = O.var_data(lv)
        Av = O.assemble_operator_var(lv, p, bx, by)
        b = O.assemble_rhs_var(lv, p, bx, by, f, gW, gE, gS, gN)
        x = spla.spsolve(Av.tocsc(), b)
        errs.append(O.h_norm(x - u, O.h_tilde_diag(lv, p)))
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    want = {2: 2.0, 4: 4.2, 6: 5.6}[p]
    assert all(abs(r - want) < 0.2 for r in rates), rates


@pytest.mark.parametrize("p", [2, 4, 6])
def test_variable_hierarchy(p):
    """VarHierarchy: at b = 1 its V-cycle is the constant-coefficient one;
    with varying profiles MG-PCG reaches the sparse direct solution in a
    level-independent number of iterations (9 / 11 / 15 for p = 2 / 4 / 6)."""
    cl = 5 if p == 6 else 4
    level = 6
    N = 2 ** level + 1
    one = np.ones(N)
    r = np.random.default_rng(0).standard_normal(N * N)
    zc = O.vcycle(O.Hierarchy(level, p, coarse_level=cl), r)
    zv = O.vcycle(O.VarHierarchy(level, p, one, one, coarse_level=cl), r)
    assert np.abs(zv - zc).max() < 1e-12 * np.abs(zc).max()
    its = []
    for lv in (6, 7):
        bx, by, f, gW, gE, gS, gN, u = O.var_data(lv)
        H = O.VarHierarchy(lv, p, bx, by, coarse_level=cl)
        b = O.assemble_rhs_var(lv, p, bx, by, f, gW, gE, gS, gN)
        x, it, _ = O.mgpcg(H, b, rtol=1e-10)
        xd = spla.spsolve(H.A[0].tocsc(), b)
        assert np.linalg.norm(x - xd) < 1e-8 * np.linalg.norm(xd)
        its.append(it)
    assert its[0] == its[1] and its[0] <= 16, its


def test_fourier_modes_deterministic():
    s1, b1 = O.fourier_modes(123)
    s2, b2 = O.fourier_modes(123)
    assert s1 == s2 and b1 == b2
    assert len(s1) == 64 and all(len(m) == 8 for m in b1)
    for kx, ky, a, ph in s1:
        assert 1 <= kx <= 32 and 1 <= ky <= 32 and 0 <= ph < 2 * np.pi
        assert abs(a - 1 / (kx * kx + ky * ky + 1)) < 1e-15
    assert O.fourier_modes(124)[0] != s1


@pytest.mark.parametrize("level,p", [(4, 2), (4, 4), (5, 2), (5, 4)])
def test_golden_fixtures(golden_dir, level, p):
    d = np.load(os.path.join(golden_dir, f"sbp_L{level}_p{p}.npz"))
    A = O.assemble_operator(level, p)
    A0 = sp.csr_matrix((d["A_data"], d["A_indices"], d["A_indptr"]), shape=A.shape)
    assert abs(A - A0).max() < 1e-14
    assert np.abs(A @ d["u"] - d["Au"]).max() < 1e-12
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    assert np.abs(O.assemble_rhs(level, p, f, gW, gE, gS, gN) - d["b"]).max() < 1e-14
    assert np.abs(A @ d["x"] - d["b"]).max() < 1e-10
