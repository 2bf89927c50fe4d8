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
