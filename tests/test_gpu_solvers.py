"""GPU parity of the solvers against the CPU oracle: CG (configs: 129^2,
p = 2, relative residual 1e-8), MG-PCG FP64, mixed-precision iterative
refinement, the CSR CG, the reference-named mirror functions, and the
status codes of the failure paths (forced from the host)."""
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


@pytest.fixture(scope="module")
def SP():
    from sc_2022_amd.solver import SbpPoisson
    return SbpPoisson


def problem(level, p, data="manufactured"):
    A = O.assemble_operator(level, p)
    if data == "manufactured":
        f, gW, gE, gS, gN, u = O.manufactured_data(level)
    else:
        f, gW, gE, gS, gN = O.fourier_data(level, 7)
        u = None
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    return A, b, u, (f, gW, gE, gS, gN)


@pytest.mark.parametrize("p", [2, 4])
def test_cg_129(SP, p):
    """The reference's CPU-runnable case: FP64 matrix-free CG, 129^2, x0 = 0, ||r|| < 1e-8 ||b||."""
    level = 7
    A, b, u, _ = problem(level, p)
    tol = 1e-8 * np.linalg.norm(b)
    xo, ito, norms = O.cg_cpu(A, b, abstol=tol)
    h = SP(129, p)
    x, res = h.cg(T(b), abstol=tol, history=True)
    assert res.converged
    assert abs(res.iters - ito) <= 1, (res.iters, ito)
    xg = x.cpu().numpy()
    assert np.linalg.norm(A @ xg - b) < tol * 1.000001
    assert np.linalg.norm(xg - xo) < 1e-6 * np.linalg.norm(xo)
    # CG amplifies rounding differences over the iterations: tight early, looser later
    np.testing.assert_allclose(res.history[:10], norms[:10], rtol=1e-10)
    np.testing.assert_allclose(res.history[:50], norms[:50], rtol=1e-4)
    # discretisation error of the manufactured solution
    hd = O.h_tilde_diag(level, p)
    assert O.h_norm(xg - u, hd) < (2.0e-5 if p == 2 else 1e-6)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [5, 7, 9])
def test_mgpcg_f64_against_direct(SP, level, p):
    A, b, u, _ = problem(level, p)
    xd = spla.spsolve(A.tocsc(), b)
    h = SP(2 ** level + 1, p)
    x, res = h.mgpcg(T(b), abstol=1e-12 * np.linalg.norm(b), history=True)
    assert res.converged and res.iters < 40
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xd) < 1e-10 * np.linalg.norm(xd)
    if level <= 7:
        hier = O.Hierarchy(level, p)
        xo, ito, norms = O.mgpcg(hier, b, rtol=1e-12)
        assert abs(res.iters - ito) <= 1
        np.testing.assert_allclose(res.history[:5], norms[:5], rtol=1e-9)


@pytest.mark.parametrize("p", [2, 4])
def test_mgpcg_fourier_data(SP, p):
    level = 8
    A, b, _, _ = problem(level, p, "fourier")
    xd = spla.spsolve(A.tocsc(), b)
    h = SP(2 ** level + 1, p)
    x, res = h.mgpcg(T(b), abstol=1e-13 * np.linalg.norm(b))
    assert np.linalg.norm(x.cpu().numpy() - xd) < 1e-10 * np.linalg.norm(xd)


@pytest.mark.parametrize("p", [2, 4])
def test_mixed_precision_ir(SP, p):
    level = 9
    A, b, _, _ = problem(level, p, "fourier")
    xd = spla.spsolve(A.tocsc(), b)
    h = SP(2 ** level + 1, p)
    x, res = h.mgpcg_mixed(T(b), abstol=1e-12 * np.linalg.norm(b))
    assert res.converged, res
    xg = x.cpu().numpy()
    hd = O.h_tilde_diag(level, p)
    rel = O.h_norm(xg - xd, hd) / O.h_norm(xd, hd)
    assert rel <= 1e-10, rel
    assert 1 <= res.iters <= 6 and res.inner_iters >= res.iters
    # stopping on the correction-based error estimate
    he = SP(2 ** level + 1, p, ir_rtol=1e-11)
    x, res = he.mgpcg_mixed(T(b), abstol=1e-14 * np.linalg.norm(b))
    assert res.converged and res.iters <= 3, res
    rel = O.h_norm(x.cpu().numpy() - xd, hd) / O.h_norm(xd, hd)
    assert rel <= 1e-10, rel


@pytest.mark.parametrize("p", [2, 4])
def test_csr_cg(SP, p):
    level = 6
    A, b, _, _ = problem(level, p)
    tol = 1e-8 * np.linalg.norm(b)
    xo, ito, _ = O.cg_cpu(A, b, abstol=tol)
    h = SP(2 ** level + 1, p)
    csr = h.csr_assemble()
    x, res = h.csr_cg(csr, T(b), abstol=tol)
    assert res.converged and abs(res.iters - ito) <= 1
    assert np.linalg.norm(x.cpu().numpy() - xo) < 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("p", [2, 4])
def test_assembling_matrix_on_device(p):
    """Assembling_matrix (level_2_multigrid_new.jl:125-188) assembled on the
    device: CSR A, b for u = sin(pi (x + y)), H~ and u against the oracle's
    restatement."""
    import scipy.sparse as sps
    from sc_2022_amd import reference_api as R
    level = 5
    (rp, col, val), b, Ht, Nx, Ny, u = R.Assembling_matrix(level, p)
    A0, b0, H0, Nx0, Ny0, u0 = O.assembling_matrix(level, p)
    n = Nx * Ny
    A = sps.csr_matrix((val.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    assert (Nx, Ny) == (Nx0, Ny0)
    assert abs(A - A0).max() < 1e-13 * abs(A0).max()
    assert relmax(b.cpu().numpy(), b0) < 1e-13
    assert relmax(Ht.cpu().numpy(), H0.diagonal()) < 1e-15
    assert relmax(u.cpu().numpy(), u0) < 1e-15


@pytest.mark.parametrize("p", [2, 4, 6, 8, 10])
def test_h_weights_match_oracle(SP, p):
    N = 41
    hw = SP(N, p).h_weights()
    _, _, H, _ = O.diagonal_sbp_D1(p, N - 1, xc=(0, 1))
    assert np.abs(hw - H.diagonal() * (N - 1)).max() < 1e-14


def test_reference_api_mirror():
    from sc_2022_amd import reference_api as R
    level = 5
    N = 2 ** level + 1
    A, b, _, _ = problem(level, 2)
    u = np.random.default_rng(1).standard_normal(N * N)
    out = T(np.zeros(N * N))
    R.matrix_free_A_full_GPU(T(u), out)
    assert np.abs(out.cpu().numpy() - O.reference_gpu_apply_p2(u, N)).max() < 1e-12
    rc = T(np.zeros(17 * 17))
    R.matrix_free_restriction_2d_GPU(T(u), rc)
    assert np.abs(rc.cpu().numpy() - O.restriction_2d(N) @ u).max() < 1e-14
    ec = np.random.default_rng(2).standard_normal(17 * 17)
    fo = T(np.ones(N * N))
    R.matrix_free_prolongation_2d_GPU(T(ec), fo)
    assert np.abs(fo.cpu().numpy() - O.prolongation_2d(17) @ ec).max() < 1e-14
    xi, xo_ = T(np.zeros(N * N)), T(np.zeros(N * N))
    R.matrix_free_richardson(xi, xo_, T(b), maxiter=3, omega=0.15)
    assert np.abs(xo_.cpu().numpy() - O.modified_richardson(A, np.zeros(N * N), b)).max() < 1e-12
    x = T(np.zeros(N * N))
    iters, norms, x = R.matrix_free_MGCG(T(b), x)           # the reference's absolute sqrt(eps)
    assert np.linalg.norm(A @ x.cpu().numpy() - b) < 1.0000001 * math.sqrt(2.220446049250313e-16)
    x = T(np.zeros(N * N))
    iters, norms, x = R.matrix_free_MGCG(T(b), x, abstol=math.sqrt(2.2e-16) * np.linalg.norm(b))
    assert np.linalg.norm(A @ x.cpu().numpy() - b) < math.sqrt(2.2e-16) * np.linalg.norm(b)
    assert len(norms) == iters + 1
    x = T(np.zeros(N * N))
    it, tol, last = R.CG_full_GPU(T(b), x, abstol=1e-8 * np.linalg.norm(b))
    assert last < tol and np.linalg.norm(A @ x.cpu().numpy() - b) < 1.0001 * tol


def nested_problem(level, p, data):
    A, rhs = [], []
    for k in range(3):
        A.append(O.assemble_operator(level - k, p))
        d = data(level - k)
        rhs.append(O.assemble_rhs(level - k, p, *d[:5]))
    return A, rhs, data(level)[5]


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("three", [False, True])
def test_cg_interp_bilinear_guess(SP, p, three):
    """maxiter = 0 returns the interpolated guess; for a bilinear u (the
    discrete solution on every grid) it must be u to round-off."""
    level = 6
    A, rhs, u = nested_problem(level, p, O.bilinear_data)
    h = SP(2 ** level + 1, p)
    b4 = T(rhs[2]) if three else None
    x, res = h.cg_interp(T(rhs[0]), T(rhs[1]), b4, abstol=1e-10, maxiter=0, raise_on_fail=False)
    assert np.abs(x.cpu().numpy() - u).max() < 1e-11
    assert len(res.level_iters) == (2 if three else 1)


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("three", [False, True])
def test_cg_interp_against_oracle(SP, p, three):
    """Nested solve at 129^2 (and 65^2 / 33^2 below it) on the manufactured
    problem against the oracle's restatement: per-grid CG iteration counts,
    the oracle-recomputed residual, the solution."""
    level = 7
    A, rhs, u = nested_problem(level, p, O.manufactured_data)
    tol = 1e-9 * np.linalg.norm(rhs[0])
    kw = dict(A4h=A[2], b4h=rhs[2]) if three else {}
    xo, ito, nro, x0o = O.initial_guess_interpolation_cg(A[0], rhs[0], A[1], rhs[1], tol, **kw)
    h = SP(2 ** level + 1, p)
    b4 = T(rhs[2]) if three else None
    if not three:   # the guess itself (no CG iterations) against the oracle's
        x0, _ = h.cg_interp(T(rhs[0]), T(rhs[1]), None, abstol=tol, maxiter=0, raise_on_fail=False)
        assert np.linalg.norm(x0.cpu().numpy() - x0o) < 1e-10 * np.linalg.norm(x0o)
    x, res = h.cg_interp(T(rhs[0]), T(rhs[1]), b4, abstol=tol)
    assert res.converged
    assert all(abs(a - b) <= 2 for a, b in zip(res.level_iters, ito)), (res.level_iters, ito)
    xg = x.cpu().numpy()
    assert np.linalg.norm(A[0] @ xg - rhs[0]) < tol * 1.000001
    assert np.linalg.norm(xg - xo) < 1e-6 * np.linalg.norm(xo)
    # fewer fine-grid iterations than CG from zero
    _, r0 = h.cg(T(rhs[0]), abstol=tol)
    assert res.iters < r0.iters


def test_cg_interp_mirror_and_errors(SP):
    from sc_2022_amd import _capi as capi
    from sc_2022_amd import reference_api as R
    level, p = 6, 2
    A, rhs, u = nested_problem(level, p, O.manufactured_data)
    N = 2 ** level + 1
    tol = 1e-9 * np.linalg.norm(rhs[0])
    x = T(np.zeros(N * N))
    x, it, last = R.initial_guess_interpolation_CG_Matrix_Free_GPU(None, T(rhs[0]), T(rhs[1]), x, 33, abstol=tol)
    assert last < tol and np.linalg.norm(A[0] @ x.cpu().numpy() - rhs[0]) < 1.0001 * tol
    x = T(np.zeros(N * N))
    x, i2, ih, n2, nh = R.initial_guess_interpolation_three_level_CG_GPU(
        None, None, T(rhs[0]), T(rhs[1]), None, rhs[2], x, 33, 17, abstol=tol)
    assert n2 < tol and nh < tol and i2 > 0
    assert np.linalg.norm(A[0] @ x.cpu().numpy() - rhs[0]) < 1.0001 * tol
    # three levels need (N+3)/4 >= 17: not on 33^2
    h33 = SP(33, p)
    b33 = T(O.assemble_rhs(5, p, *O.manufactured_data(5)[:5]))
    b17 = T(O.assemble_rhs(4, p, *O.manufactured_data(4)[:5]))
    b9 = T(np.zeros(81))
    lib = h33.lib
    import ctypes as C
    its = (C.c_int * 2)()
    rns = (C.c_double * 2)()
    xo = torch.zeros_like(b33)
    st = lib.sbp_cg_interp(h33._h, 3, C.c_void_p(b33.data_ptr()), C.c_void_p(b17.data_ptr()),
                           C.c_void_p(b9.data_ptr()), C.c_void_p(xo.data_ptr()), 1e-8, 100, its, rns, None)
    assert st == capi.SBP_ERR_INVALID_SIZE
    st = lib.sbp_cg_interp(h33._h, 4, C.c_void_p(b33.data_ptr()), C.c_void_p(b17.data_ptr()),
                           C.c_void_p(b9.data_ptr()), C.c_void_p(xo.data_ptr()), 1e-8, 100, its, rns, None)
    assert st == capi.SBP_ERR_INVALID_ARG
    x, res = h33.cg_interp(b33, b17, abstol=1e-9 * float(torch.linalg.vector_norm(b33)))
    assert res.converged


@pytest.mark.parametrize("N", [65, 129])
def test_reference_two_level_mode(SP, N):
    """The reference's own preconditioner (two-level, 3 + 3 Richardson sweeps
    with weight 0.15, unscaled full weighting, exact coarse solve) against the
    oracle's restatement, alone and inside MGCG (p = 2).  At 65^2 the coarse
    33^2 grid takes the dense inverse; at 129^2 the coarse 65^2 grid takes the
    nested MG-PCG solve."""
    from sc_2022_amd import reference_api as R
    level = int(math.log2(N - 1))
    A, A2 = O.assemble_operator(level, 2), O.assemble_operator(level - 1, 2)
    h = SP(N, 2, **R.reference_mg_params(N, 2))
    assert h.nlevels == 2
    r = np.random.default_rng(6).standard_normal(N * N)
    zo, no = O.two_level_richardson(A, A2, r)
    z = h.vcycle(T(r)).cpu().numpy()
    assert np.abs(z - zo).max() < 1e-10 * np.abs(zo).max()
    zt, nt = R.matrix_free_Two_level_multigrid(T(r))
    assert abs(nt - no) < 1e-8 * no
    _, _, _, (f, gW, gE, gS, gN) = problem(level, 2)
    b = O.assemble_rhs(level, 2, f, gW, gE, gS, gN)
    tol = 1e-10 * np.linalg.norm(b)
    xo, ito, _ = O.matrix_free_mgcg(A, b, lambda q: O.two_level_richardson(A, A2, q)[0], abstol=tol)
    x, res = h.mgpcg(T(b), abstol=tol, maxiter=200)
    assert res.converged and abs(res.iters - ito) <= 1, (res.iters, ito)
    assert np.linalg.norm(A @ x.cpu().numpy() - b) < 1.000001 * tol
    it, norms, xr = R.matrix_free_MGCG(T(b), T(np.zeros(N * N)), abstol=tol)
    assert it == res.iters and len(norms) == it + 1


@pytest.mark.parametrize("N", [65, 129])
def test_reference_three_level_mode(SP, N):
    """The reference's three-level cycle as written (Richardson 3 + 3, weight
    0.15, unscaled full weighting, the middle grid's residual without its
    right-hand side, exact 4h solve; initial_guess_cg.jl:244-301) against the
    oracle's restatement, alone and inside matrix_free_MGCG_Three_level
    (:348-391), p = 2: same iteration count within one."""
    from sc_2022_amd import reference_api as R
    level = int(math.log2(N - 1))
    A, A2, A4 = (O.assemble_operator(level - k, 2) for k in range(3))
    h = SP(N, 2, **R.reference_mg_params(N, 3))
    assert h.nlevels == 3
    r = np.random.default_rng(7).standard_normal(N * N)
    zo, no = O.three_level_richardson(A, A2, A4, r)
    z = h.vcycle(T(r)).cpu().numpy()
    assert np.abs(z - zo).max() < 1e-10 * np.abs(zo).max()
    zt, nt = R.matrix_free_Three_level_multigrid(T(r))
    assert abs(nt - no) < 1e-8 * no
    _, _, _, (f, gW, gE, gS, gN) = problem(level, 2)
    b = O.assemble_rhs(level, 2, f, gW, gE, gS, gN)
    tol = 1e-10 * np.linalg.norm(b)
    xo, ito, _ = O.matrix_free_mgcg(A, b, lambda q: O.three_level_richardson(A, A2, A4, q)[0], abstol=tol)
    x = T(np.zeros(N * N))
    it, norms = R.matrix_free_MGCG_Three_level(T(b), x, abstol=tol)
    assert abs(it - ito) <= 1, (it, ito)
    assert len(norms) == it + 1 and norms[-1] < tol
    assert np.linalg.norm(A @ x.cpu().numpy() - b) < 1.000001 * tol


def test_reference_transfers_small_grids():
    """matrix_free_restriction_2d_GPU / matrix_free_prolongation_2d_GPU between
    17^2 and 9^2 (below the default 17^2 multigrid bottom)."""
    from sc_2022_amd import reference_api as R
    rng = np.random.default_rng(3)
    u = rng.standard_normal(17 * 17)
    rc = T(np.zeros(9 * 9))
    R.matrix_free_restriction_2d_GPU(T(u), rc)
    assert np.abs(rc.cpu().numpy() - O.restriction_2d(17) @ u).max() < 1e-14
    ec = rng.standard_normal(9 * 9)
    fo = T(np.ones(17 * 17))
    R.matrix_free_prolongation_2d_GPU(T(ec), fo)
    assert np.abs(fo.cpu().numpy() - O.prolongation_2d(9) @ ec).max() < 1e-14


def test_status_codes(SP):
    from sc_2022_amd import _capi as capi
    level, p = 5, 2
    A, b, _, _ = problem(level, p)
    h = SP(2 ** level + 1, p)
    # maxiter = 1 -> not converged, partial iterate returned
    x, res = h.cg(T(b), abstol=1e-14, maxiter=1, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_NOT_CONVERGED and res.iters == 1
    # NaN in b -> non-finite
    bn = b.copy(); bn[7] = np.nan
    x, res = h.cg(T(bn), abstol=1e-8, maxiter=10, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_NONFINITE
    x, res = h.mgpcg(T(bn), abstol=1e-8, maxiter=10, raise_on_fail=False)
    assert res.status == capi.SBP_ERR_NONFINITE
    # b = 0 -> converged at iteration 0
    x, res = h.mgpcg(T(np.zeros_like(b)), abstol=1e-8)
    assert res.converged and res.iters == 0
    # an indefinite preconditioner (negative smoothing weight is rejected at
    # create; a negative restriction scale flips the coarse correction sign)
    with pytest.raises(Exception):
        SP(33, 2, omega=-0.5)
    # multigrid needs 2^k+1
    hr = SP(40, 2)
    with pytest.raises(Exception):
        hr.mgpcg(T(np.ones(1600)), abstol=1e-8)


@pytest.mark.parametrize("p", [2, 4])
def test_graph_replay_matches_eager(SP, p):
    """Iterations replayed from a captured hipGraph give bit-identical results
    to eager launches, over several solves with eager work in between."""
    level = 7
    N = 2 ** level + 1
    A, b, _, _ = problem(level, p, "fourier")
    tol = 1e-12 * np.linalg.norm(b)
    hg = SP(N, p, use_graph=1)
    he = SP(N, p, use_graph=0)
    for rep in range(3):
        xg, rg = hg.mgpcg(T(b), abstol=tol, history=True)
        xe, re_ = he.mgpcg(T(b), abstol=tol, history=True)
        assert rg.iters == re_.iters and rg.history == re_.history
        assert torch.equal(xg, xe)
        cg_g, r1 = hg.cg(T(b), abstol=1e-6 * np.linalg.norm(b))
        cg_e, r2 = he.cg(T(b), abstol=1e-6 * np.linalg.norm(b))
        assert r1.iters == r2.iters and torch.equal(cg_g, cg_e)
        mg, m1 = hg.mgpcg_mixed(T(b), abstol=tol)
        me, m2 = he.mgpcg_mixed(T(b), abstol=tol)
        assert (m1.iters, m1.inner_iters) == (m2.iters, m2.inner_iters) and torch.equal(mg, me)
        hg.apply(xg)                      # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("p", [2, 4])
def test_adaptive_polls_match_polling_every_iteration(SP, p):
    """The adaptive poll schedule (the first poll after two iterations with a
    third already queued, later polls where the measured rate predicts
    convergence) stops at the same iteration as polling after every one, with
    the same history and a bit-identical iterate: for tolerances met first
    at iterations 1..4 (converged before, at and after the queued third
    iteration) and for maxiter 1..3 with an unreachable tolerance."""
    from sc_2022_amd import _capi as capi
    level = 7
    N = 2 ** level + 1
    A, b, _, _ = problem(level, p, "fourier")
    ha = SP(N, p)                  # poll_every = 0: adaptive
    h1 = SP(N, p, poll_every=1)
    _, ref = h1.mgpcg(T(b), abstol=1e-14 * np.linalg.norm(b), history=True)
    hist = ref.history
    for k in (1, 2, 3, 4):
        tol = 0.5 * (hist[k - 1] + hist[k])      # first met at iteration k
        xa, ra = ha.mgpcg(T(b), abstol=tol, history=True)
        x1, r1 = h1.mgpcg(T(b), abstol=tol, history=True)
        first = next(j for j in range(len(hist)) if hist[j] < tol)   # k unless the norms rise
        assert ra.converged and ra.iters == r1.iters == first and ra.history == r1.history
        assert torch.equal(xa, x1)
    for m in (1, 2, 3):
        xa, ra = ha.mgpcg(T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        x1, r1 = h1.mgpcg(T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        assert ra.status == r1.status == capi.SBP_ERR_NOT_CONVERGED and ra.iters == r1.iters == m
        assert torch.equal(xa, x1)
        xa, ra = ha.cg(T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        x1, r1 = h1.cg(T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        assert ra.iters == r1.iters == m and torch.equal(xa, x1)
    for rt in (1e-4, 1e-12):
        xa, ra = ha.mgpcg_mixed(T(b), abstol=rt * np.linalg.norm(b))
        x1, r1 = h1.mgpcg_mixed(T(b), abstol=rt * np.linalg.norm(b))
        assert ra.converged and (ra.iters, ra.inner_iters) == (r1.iters, r1.inner_iters)
        assert torch.equal(xa, x1)
    # the same schedule drives the CSR CG and the variable-coefficient MG-PCG
    ca, c1 = ha.csr_assemble(), h1.csr_assemble()
    for m in (2, 3):
        xa, ra = ha.csr_cg(ca, T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        x1, r1 = h1.csr_cg(c1, T(b), abstol=0.0, maxiter=m, raise_on_fail=False)
        assert ra.iters == r1.iters == m and torch.equal(xa, x1)
    ha.close(); h1.close()
    bx, by, f, gW, gE, gS, gN, _ = O.var_data(level)
    bv = T(O.assemble_rhs_var(level, p, bx, by, f, gW, gE, gS, gN))
    va, v1 = SP(N, p, coef=(bx, by)), SP(N, p, coef=(bx, by), poll_every=1)
    rn = [float(torch.linalg.vector_norm(bv))] + [
        v1.mgpcg(bv, abstol=0.0, maxiter=k, raise_on_fail=False)[1].resnorm for k in (1, 2, 3, 4)]
    for k in (1, 2, 3, 4):
        tol = 0.5 * (rn[k - 1] + rn[k])
        xa, ra = va.mgpcg(bv, abstol=tol)
        x1, r1 = v1.mgpcg(bv, abstol=tol)
        first = next(j for j in range(len(rn)) if rn[j] < tol)
        assert ra.converged and ra.iters == r1.iters == first and torch.equal(xa, x1)
    va.close(); v1.close()


@pytest.mark.parametrize("p", [2, 4])
def test_mgpcg_bottom_kernel_matches_tiled_levels(SP, p):
    """The single-workgroup bottom of the V-cycle and the all-tiled V-cycle
    give the same MG-PCG iteration count and solution to rounding."""
    level = 8
    N = 2 ** level + 1
    A, b, u, _ = problem(level, p, data="fourier")
    tol = 1e-11 * np.linalg.norm(b)
    h0 = SP(N, p, bottom_n=0)
    h1 = SP(N, p, bottom_n=65)
    x0, r0 = h0.mgpcg(T(b), abstol=tol)
    x1, r1 = h1.mgpcg(T(b), abstol=tol)
    assert r0.converged and r1.converged
    assert r0.iters == r1.iters
    assert torch.linalg.vector_norm(x0 - x1) <= 1e-11 * torch.linalg.vector_norm(x0)
    xm, rm = h1.mgpcg_mixed(T(b), abstol=tol)
    assert rm.converged
    assert torch.linalg.vector_norm(xm - x0) <= 1e-9 * torch.linalg.vector_norm(x0)


@pytest.mark.parametrize("p", [2, 4])
def test_mgpcg_fused_passes_match_separate_passes(SP, p):
    level = 9
    N = 2 ** level + 1
    A, b, u, _ = problem(level, p, data="fourier")
    tol = 1e-11 * np.linalg.norm(b)
    hf = SP(N, p, fuse_n=1 << 20)
    hs = SP(N, p, fuse_n=0)
    xf, rf = hf.mgpcg(T(b), abstol=tol)
    xs, rs = hs.mgpcg(T(b), abstol=tol)
    assert rf.converged and rs.converged and rf.iters == rs.iters
    assert torch.linalg.vector_norm(xf - xs) <= 1e-11 * torch.linalg.vector_norm(xs)
    xm, rm = hf.mgpcg_mixed(T(b), abstol=tol)
    assert rm.converged
    assert torch.linalg.vector_norm(xm - xs) <= 1e-9 * torch.linalg.vector_norm(xs)
