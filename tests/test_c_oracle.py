"""Pins the C oracle (oracle/sbp_oracle.c) against the numpy/scipy oracle:
matrix-free apply, diagonal, CSR assembly, right-hand side, CG_CPU and the
MG-PCG (same V-cycle) on small grids."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import c_oracle as CO
from oracle import sbp_oracle as O


@pytest.fixture(scope="module", autouse=True)
def built():
    CO.build()
    CO.set_threads(2)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [4, 5, 6])
def test_apply_diag_csr(level, p):
    N = 2 ** level + 1
    A = O.assemble_operator(level, p)
    u = np.random.default_rng(level).standard_normal(N * N)
    assert np.abs(CO.apply(p, N, u) - A @ u).max() < 1e-12 * np.abs(A @ u).max()
    assert np.abs(CO.diag(p, N) - A.diagonal()).max() < 1e-14
    rp, col, val = CO.csr(p, N)
    Ac = sp.csr_matrix((val, col, rp), shape=A.shape)
    assert Ac.nnz == A.nnz and abs(Ac - A).max() < 1e-13


@pytest.mark.parametrize("p", [2, 4])
def test_ragged_apply(p):
    for N in (12, 19, 50):
        hw, m, k = O.one_d_factors(p, N)
        A = sp.kron(sp.csr_matrix(k), sp.diags(hw)) + sp.kron(sp.diags(hw), sp.csr_matrix(m))
        u = np.random.default_rng(N).standard_normal(N * N)
        assert np.abs(CO.apply(p, N, u) - A @ u).max() < 1e-12 * np.abs(A @ u).max()


@pytest.mark.parametrize("p", [2, 4])
def test_rhs(p):
    level = 5
    N = 2 ** level + 1
    f, gW, gE, gS, gN = O.fourier_data(level, 3)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    assert np.abs(CO.rhs(p, N, f, gW, gE, gS, gN) - b).max() < 1e-13 * np.abs(b).max()
    assert np.abs(CO.fourier_rhs(p, N, 3) - b).max() < 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("p", [2, 4])
def test_cg_and_mgpcg(p):
    level = 6
    N = 2 ** level + 1
    A = O.assemble_operator(level, p)
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    tol = 1e-8 * np.linalg.norm(b)
    xo, ito, _ = O.cg_cpu(A, b, abstol=tol)
    x, it, rn = CO.cg_csr(CO.csr(p, N), b, tol)
    assert abs(it - ito) <= 1 and rn < tol
    assert np.linalg.norm(x - xo) < 1e-6 * np.linalg.norm(xo)
    hier = O.Hierarchy(level, p)
    xm, itm, _ = O.mgpcg(hier, b, rtol=1e-12)
    x, it, rn = CO.mgpcg(p, N, b, 1e-12 * np.linalg.norm(b))
    assert abs(it - itm) <= 1
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) < 1e-10 * np.linalg.norm(xd)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [5, 6])
def test_rhs_transpose_and_hnorm(level, p):
    """The C transposed data map equals the kron-assembled one (SAT_X_r^T of
    level_2_multigrid_new.jl:164-167) and satisfies <b(f,g), lam> = <(f,g), b^T lam>."""
    N = 2 ** level + 1
    rng = np.random.default_rng(level + p)
    lam = rng.standard_normal(N * N)
    ref = O.rhs_transpose(level, p, lam)
    got = CO.rhs_transpose(p, N, lam)
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= 1e-13 * np.abs(r).max()
    f, gW, gE, gS, gN = (rng.standard_normal(N * N), *rng.standard_normal((4, N)))
    b = CO.rhs(p, N, f, gW, gE, gS, gN)
    lhs = b @ lam
    rhs = sum(a @ c for a, c in zip((f, gW, gE, gS, gN), got))
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)
    hd = O.h_tilde_diag(level, p)
    assert abs(CO.hnorm(p, N, lam) - O.h_norm(lam, hd)) < 1e-13 * O.h_norm(lam, hd)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [5, 6])
def test_adjoint_grad_against_direct(level, p):
    """The C FP64 gradient (adjoint MG-PCG + transposed data map) equals the
    gradient from sparse direct solves: lam = A^-1 H~ (u - u_obs),
    dL/d(f, g) = b^T lam."""
    N = 2 ** level + 1
    A = O.assemble_operator(level, p).tocsc()
    fg = O.fourier_data(level, 11)
    b = O.assemble_rhs(level, p, *fg)
    u = spla.spsolve(A, b)
    uo = spla.spsolve(A, O.assemble_rhs(level, p, *O.fourier_data(level, 12)))
    u_obs = uo + 1e-3 * np.random.default_rng(1).standard_normal(N * N)
    hd = O.h_tilde_diag(level, p)
    Lo = 0.5 * float((u - u_obs) @ (hd * (u - u_obs)))
    lam_o = spla.spsolve(A, hd * (u - u_obs))
    go = O.rhs_transpose(level, p, lam_o)
    w = hd * (u - u_obs)
    L, lam, *g, it = CO.adjoint_grad(p, N, u, u_obs, 1e-14 * np.linalg.norm(w))
    L2, lam2, *_ = CO.adjoint_grad(p, N, u, u_obs, rtol=1e-14)
    assert np.array_equal(lam, lam2) and L2 == L
    assert it <= 30
    assert abs(L - Lo) < 1e-12 * Lo
    assert np.linalg.norm(lam - lam_o) < 1e-11 * np.linalg.norm(lam_o)
    for a, r in zip(g, go):
        assert np.abs(a - r).max() <= 1e-10 * np.abs(r).max()


def test_fourier_fields_match_numpy():
    """The C evaluation of the seeded Fourier fields equals the numpy one."""
    for level in (4, 6):
        N = 2 ** level + 1
        for a, c in zip(O.fourier_data(level, 5), CO.fourier_fields(N, 5)):
            assert np.abs(a - c).max() <= 1e-14


def test_solve_reference_meets_tolerance():
    """The reference solution used for the headline errors: converged to
    1e-14 ||b|| and equal to the sparse direct solve."""
    level, p = 6, 4
    N = 2 ** level + 1
    b = CO.fourier_rhs(p, N, 2022)
    x, it, rr = CO.solve_reference(p, N, b)
    assert rr < 1e-14
    xd = spla.spsolve(O.assemble_operator(level, p).tocsc(), b)
    assert np.linalg.norm(x - xd) < 1e-12 * np.linalg.norm(xd)


@pytest.mark.parametrize("p", [2, 4])
def test_apply_keeps_exact_row_sums(p):
    """A 1 equals the rational row sums exactly: 0 on every row away from the
    Dirichlet faces (D2 1 = 0, BS 1 = 0, diagonal_sbp.jl:189-454) and
    hw_i (13 delta_j0 - bs_j) on the Dirichlet closure rows.  The plain sum of
    the rounded FP64 entries leaves ~1e-16 there for p = 4 (a diagonal shift
    that moves the 4097^2 solution by ~1e-10 relative), which the oracle's
    difference-form evaluation does not."""
    N = 65
    one = np.ones(N * N)
    hw = O.one_d_factors(p, N)[0]
    bs = {2: [1.5, -2.0, 0.5], 4: [11 / 6, -3.0, 1.5, -1 / 3]}[p]
    krs = np.zeros(N)
    krs[0] += 13.0
    krs[-1] += 13.0
    for r, v in enumerate(bs):
        krs[r] -= v
        krs[N - 1 - r] -= v
    want = (hw[:, None] * krs[None, :]).reshape(-1, order="F")   # i fast
    got = CO.apply(p, N, one)
    assert np.abs(got - want).max() <= 1e-15 * 13
    inner = np.zeros((N, N), dtype=bool)
    inner[8:-8, 8:-8] = True
    inner = inner.reshape(-1, order="F")
    assert not np.any(got[inner])
    plain = CO.apply_plain(p, N, one)
    if p == 4:
        assert np.abs(plain[inner]).max() > 1e-17
