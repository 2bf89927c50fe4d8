"""GPU parity of the line-aligned stencil tiles (sbp_set_tiles with
SBP_ALIGNED: each block stores whole 128-byte lines of every grid row) against
the CPU oracle and against the unaligned tiles: the apply over sides whose rows
start at every offset within a line, destinations that are not line-aligned,
the fused inner product (each point counted once), and the MG-PCG / mixed
solves whose Krylov and V-cycle passes all run on aligned tiles."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from oracle import sbp_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
ALIGNED = [(256, 8), (256, 16), (512, 8)]


def T(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def relmax(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def closed_form(p, N):
    hw, m, k = O.one_d_factors(p, N)
    return sp.kron(sp.csr_matrix(k), sp.diags(hw)) + sp.kron(sp.diags(hw), sp.csr_matrix(m))


@pytest.fixture(scope="module")
def SP():
    from sc_2022_amd.solver import SbpPoisson
    return SbpPoisson


def aligned_handle(SP, N, p, shape, levels=(0,), **kw):
    h = SP(N, p, **kw)
    for lv in levels:
        h.set_tiles(lv, shape[0], shape[1], aligned=True)
    return h


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("shape", ALIGNED)
@pytest.mark.parametrize("N", [12, 33, 100, 129, 241, 256, 257, 300, 513, 1000])
def test_aligned_apply_closed_form(SP, N, p, shape):
    """Sides 12..1000 (N mod 16 and N mod 32 take many values, so the row
    starts fall at many offsets within a line), FP64 and FP32."""
    if p == 4 and N < 12:
        pytest.skip("below the p=4 minimum")
    A = closed_form(p, N)
    h = aligned_handle(SP, N, p, shape, tuned=False)
    rng = np.random.default_rng(N * 7 + p)
    u = rng.standard_normal(N * N)
    assert relmax(h.apply(T(u)).cpu().numpy(), A @ u) < 1e-13
    assert relmax(h.apply(T(u), sign=-1).cpu().numpy(), -(A @ u)) < 1e-13
    uf = u.astype(np.float32)
    out = h.apply(T(uf, torch.float32)).cpu().double().numpy()
    assert relmax(out, A @ uf.astype(np.float64)) < 2e-5
    h.close()


@pytest.mark.parametrize("off", [1, 3, 5, 17])
def test_aligned_apply_unaligned_destination(SP, off):
    """Destinations and operands that start mid-line (views at an element
    offset): the stored ranges follow the destination's own lines."""
    N, p = 257, 4
    A = closed_form(p, N)
    h = aligned_handle(SP, N, p, (256, 8), tuned=False)
    u = np.random.default_rng(off).standard_normal(N * N)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-5)):
        ub = torch.zeros(N * N + off, dtype=dt, device=DEV)
        ub[off:] = T(u, dt)
        ob = torch.full((N * N + 2 * off,), 7.0, dtype=dt, device=DEV)
        h.apply(ub[off:], ob[off:off + N * N])
        got = ob[off:off + N * N].cpu().double().numpy()
        ref = A @ u.astype(np.float32).astype(np.float64) if dt == torch.float32 else A @ u
        assert relmax(got, ref) < tol
        # nothing written outside the destination
        assert torch.all(ob[:off] == 7.0) and torch.all(ob[off + N * N:] == 7.0)
    h.close()


@pytest.mark.parametrize("shape", ALIGNED)
def test_aligned_apply_dot_counts_each_point_once(SP, shape):
    N, p = 1025, 4
    h = aligned_handle(SP, N, p, shape, tuned=False)
    u = torch.randn(N * N, dtype=torch.float64, device=DEV, generator=torch.Generator(device=DEV).manual_seed(3))
    Au, dot = h.apply_dot(u)
    ref = SP(N, p, tuned=False)
    Au0 = ref.apply(u)
    # the same stencil per point (up to the compiler's contraction of a tile shape)
    assert float((Au - Au0).abs().max()) <= 1e-14 * float(Au0.abs().max())
    d0 = float(torch.dot(u, Au0))
    assert abs(dot - d0) <= 1e-12 * abs(d0)
    h.close(); ref.close()


@pytest.mark.parametrize("p", [2, 4])
def test_aligned_mgpcg_matches_unaligned(SP, p):
    """Every pass of levels 0 and 1 (Krylov direction update, residual
    update inputs, smoothing, prolongation, last sweep with r'z) on aligned
    tiles: same iteration count and solution to rounding, FP64 and mixed."""
    level = 10
    N = 2 ** level + 1
    f, gW, gE, gS, gN = O.fourier_data(level, 7)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    tol = 1e-11 * np.linalg.norm(b)
    hu = SP(N, p, tuned=False, fuse_n=257)   # levels 0 and 1 as separate passes
    for shape in ALIGNED:
        ha = aligned_handle(SP, N, p, shape, levels=(0, 1), tuned=False, fuse_n=257)
        xu, ru = hu.mgpcg(T(b), abstol=tol)
        xa, ra = ha.mgpcg(T(b), abstol=tol)
        assert ru.converged and ra.converged and abs(ru.iters - ra.iters) <= 1
        assert torch.linalg.vector_norm(xa - xu) <= 1e-10 * torch.linalg.vector_norm(xu)
        xm, rm = ha.mgpcg_mixed(T(b), abstol=tol)
        assert rm.converged
        assert torch.linalg.vector_norm(xm - xu) <= 1e-9 * torch.linalg.vector_norm(xu)
        xc, rc = ha.cg(T(b), abstol=1e-6 * np.linalg.norm(b), maxiter=20000)
        xcu, rcu = hu.cg(T(b), abstol=1e-6 * np.linalg.norm(b), maxiter=20000)
        assert rc.converged and abs(rc.iters - rcu.iters) <= 2
        ha.close()
    hu.close()


def test_aligned_mgpcg_against_direct(SP):
    level, p = 8, 4
    N = 2 ** level + 1
    A = O.assemble_operator(level, p)
    f, gW, gE, gS, gN, u = O.manufactured_data(level)
    b = O.assemble_rhs(level, p, f, gW, gE, gS, gN)
    xd = spla.spsolve(A.tocsc(), b)
    h = aligned_handle(SP, N, p, (256, 8), levels=(0,), tuned=False)
    x, res = h.mgpcg(T(b), abstol=1e-12 * np.linalg.norm(b))
    assert res.converged
    assert np.linalg.norm(x.cpu().numpy() - xd) < 1e-10 * np.linalg.norm(xd)
    h.close()


def test_aligned_shapes_rejected_where_not_compiled(SP):
    from sc_2022_amd import _capi as capi
    h = SP(65, 4, tuned=False)
    lib = h.lib
    assert lib.sbp_set_tiles(h._h, 0, capi.SBP_F64 | capi.SBP_ALIGNED, 64, 4) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_set_tiles(h._h, 0, capi.SBP_F64 | capi.SBP_ALIGNED | capi.SBP_FUSED, 256, 8) == capi.SBP_ERR_INVALID_ARG
    assert lib.sbp_set_tiles(h._h, 0, capi.SBP_F64 | capi.SBP_ALIGNED, 256, 8) == capi.SBP_OK
    h.close()
    h6 = SP(65, 6, tuned=False)
    assert h6.lib.sbp_set_tiles(h6._h, 0, capi.SBP_F64 | capi.SBP_ALIGNED, 256, 8) == capi.SBP_ERR_INVALID_ARG
    h6.close()
