"""The reference's GPU call surface, on torch device tensors.

Each function keeps the name, argument order and in-place behaviour of the
Julia function it stands in for, and calls libsbp.so through ``SbpPoisson``.
Arrays are (N, N) or flat N*N device tensors in the reference's layout
(element (i, j) at i + j*N).  Differences from the reference are deliberate
and listed per function.
"""
from __future__ import annotations

import math

import torch

from .solver import SbpPoisson

_HANDLES: dict = {}


def handle(N, p=2, device=None, **params):
    """Cached solver handle for (device, N, p, params)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    key = (str(dev), int(N), int(p), tuple(sorted(params.items())))
    h = _HANDLES.get(key)
    if h is None:
        h = SbpPoisson(N, p, device=dev, **params)
        _HANDLES[key] = h
    return h


def _side(t):
    n = t.numel()
    N = int(round(math.sqrt(n)))
    if N * N != n:
        raise ValueError("grid arrays must hold N*N elements")
    return N


def Assembling_matrix(level, p=2, device=None):
    """(A, b, H_tilde, Nx, Ny, analy_sol) -- level_2_multigrid_new.jl:125-188,
    assembled on the device: A the SPD operator as device CSR (row pointers,
    columns, values; int32 / FP64), b its right-hand side for the reference's
    u = sin(pi (x + y)), H_tilde the diagonal of H~ (N*N values), and the
    analytic solution.  The reference returns Julia SparseMatrixCSC and dense
    vectors on the host."""
    from .data import reference_fields
    N = 2 ** level + 1
    h = handle(N, p, device)
    f, gW, gE, gS, gN, u = reference_fields(N, h.device)
    b = h.assemble_rhs(f, gW, gE, gS, gN)
    return h.csr_assemble(), b, h.h_tilde(), N, N, u


def matrix_free_A_full_GPU(idata, odata, p=2):
    """odata = -A idata (split_matrix_free.jl:444-486).  The reference exists for
    p = 2 only; p = 4 is the same operator family at fourth order."""
    h = handle(_side(idata), p, device=idata.device)
    h.apply(idata.reshape(-1), odata.reshape(-1), sign=-1)
    return None


def matrix_free_prolongation_2d_GPU(idata, odata):
    """odata = P idata, bilinear, coarse (n, n) -> fine (2n-1, 2n-1)
    (initial_guess_cg.jl:41-77)."""
    Nc = _side(idata)
    # a two-level handle whose coarse grid is Nc: any fine side whose coarse
    # operator table builds (17 -> 9 and up), not only the default 17^2 bottom
    h = handle(2 * Nc - 1, 2, device=idata.device, coarse_n=Nc)
    odata.zero_()
    h.prolong_add(idata.reshape(-1), odata.reshape(-1), level=0)
    return None


def matrix_free_restriction_2d_GPU(idata, odata):
    """odata = full weighting [1 2 1;2 4 2;1 2 1]/16 of idata with zero padding
    (initial_guess_cg.jl:98-131), without the padded host copy."""
    Nf = _side(idata)
    h = handle(Nf, 2, device=idata.device, restrict_scale=1.0, coarse_n=(Nf + 1) // 2)
    h.restrict(idata.reshape(-1), level=0, out=odata.reshape(-1))
    return None


def matrix_free_richardson(idata, odata, b, maxiter=3, omega=0.15, p=2):
    """maxiter Richardson sweeps x <- x + omega (b - A x) starting from idata;
    idata and odata both end holding the result (initial_guess_cg.jl:134-144)."""
    h = handle(_side(idata), p, device=idata.device, smoother=1, richardson_omega=float(omega))
    if maxiter == 0:
        odata.copy_(idata)
        return None
    h.smooth(b.reshape(-1), idata.reshape(-1), int(maxiter))
    odata.copy_(idata)
    return None


def reference_mg_params(N, levels=2, nu=3, omega=0.15):
    """sbp_params of the reference's own preconditioner: nu Richardson sweeps
    (weight omega) before and after, full weighting without scaling, and an
    exact solve on the grid `levels - 1` halvings down
    (initial_guess_cg.jl:134-180).  levels = 3 is the three-level cycle as
    the reference writes it (:244-301): the middle grid's residual is
    restricted without that grid's right-hand side (:281-283,
    drop_coarse_rhs); its aliased smoothing call at :296 is dead (the call
    at :297 overwrites its output) and has no counterpart."""
    Nc = N
    for _ in range(levels - 1):
        Nc = (Nc + 1) // 2
    return dict(smoother=1, richardson_omega=float(omega), nu_pre=int(nu), nu_post=int(nu),
                restrict_scale=1.0, coarse_n=Nc, drop_coarse_rhs=int(levels >= 3))


def matrix_free_Two_level_multigrid(b, A_2h=None, nu=3, NUM_V_CYCLES=1, SBPp=2):
    """One two-level cycle from zero (initial_guess_cg.jl:147-180) -> (z,
    ||A z - b||).  A_2h is unused: the coarse operator is rediscretised by the
    handle and solved exactly (dense inverse up to 33^2, nested MG-PCG to
    round-off above)."""
    if NUM_V_CYCLES != 1:
        raise ValueError("one cycle per call, as the reference's callers use it")
    N = _side(b)
    h = handle(N, SBPp, device=b.device, **reference_mg_params(N, 2, nu))
    bb = b.reshape(-1)
    z = h.vcycle(bb)
    r = h.apply(z) - bb
    return z.reshape(b.shape), float(torch.linalg.vector_norm(r))


def matrix_free_Three_level_multigrid(b, A_4h=None, nu=3, NUM_V_CYCLES=1, SBPp=2):
    """Three-level cycle as the reference writes it (initial_guess_cg.jl:244-301)
    -> (z, ||A z - b||): Richardson smoothing, the middle grid's residual
    restricted without its right-hand side (r_2h = -A_2h v, :282), exact
    solve on the 4h grid.  A_4h is unused: the handle rediscretises and
    inverts it.  The dead aliased smoothing call of :296 is not repeated."""
    if NUM_V_CYCLES != 1:
        raise ValueError("one cycle per call, as the reference's callers use it")
    N = _side(b)
    h = handle(N, SBPp, device=b.device, **reference_mg_params(N, 3, nu))
    bb = b.reshape(-1)
    z = h.vcycle(bb)
    r = h.apply(z) - bb
    return z.reshape(b.shape), float(torch.linalg.vector_norm(r))


def matrix_free_MGCG(b, x, maxiter=None, abstol=None, nu=3, p=2, levels=2):
    """MG-preconditioned CG (initial_guess_cg.jl:303-346) -> (iters, norms, x)
    with the reference's preconditioner (reference_mg_params: two-level
    Richardson, nu = 3, exact coarse solve); abstol defaults to the
    reference's absolute sqrt(eps) (:303; some drivers pass sqrt(eps) ||b||,
    matrix_free_mg_cg.jl:466-468).  The build's own multi-level V-cycle is
    SbpPoisson.mgpcg."""
    h = handle(_side(b), p, device=b.device, **reference_mg_params(_side(b), levels, nu))
    bb = b.reshape(-1)
    if abstol is None:
        abstol = math.sqrt(2.220446049250313e-16)
    maxiter = bb.numel() if maxiter is None else maxiter
    xs, res = h.mgpcg(bb, x.reshape(-1), abstol=abstol, maxiter=maxiter, history=True)
    return res.iters, res.history, x


def matrix_free_MGCG_Three_level(b, x, maxiter=None, abstol=None, nu=3, p=2):
    """matrix_free_MGCG_Three_level (initial_guess_cg.jl:348-391) -> (iters,
    norms): MG-preconditioned CG with the reference's three-level cycle
    (reference_mg_params(N, 3)); abstol defaults to the reference's absolute
    sqrt(eps).  x receives the solution (the reference returns only the
    iteration count and the residual norms)."""
    N = _side(b)
    h = handle(N, p, device=b.device, **reference_mg_params(N, 3, nu))
    bb = b.reshape(-1)
    abstol = math.sqrt(2.220446049250313e-16) if abstol is None else abstol
    maxiter = bb.numel() if maxiter is None else maxiter
    _, res = h.mgpcg(bb, x.reshape(-1), abstol=abstol, maxiter=maxiter, history=True, raise_on_fail=False)
    return res.iters, res.history


def CG_Matrix_Free_GPU_v2(x, Ap, b, Nx, Ny, abstol=None, p=2):
    """CG with the matrix-free apply (split_matrix_free.jl:688-714) ->
    (iters, abstol, last ||r||).  Solves the SPD system A x = b; the reference
    applies its -A kernel here and so converges to -A^-1 b."""
    h = handle(Nx, p, device=b.device)
    abstol = math.sqrt(2.220446049250313e-16) if abstol is None else abstol
    _, res = h.cg(b.reshape(-1), x.reshape(-1), abstol=abstol, maxiter=Nx * Ny)
    return res.iters, abstol, res.resnorm


def CG_full_GPU(b, x, abstol=None, p=2):
    """CG_full_GPU (split_matrix_free.jl:647-686) -> (iters, abstol, last ||r||)."""
    N = _side(b)
    return CG_Matrix_Free_GPU_v2(x, None, b, N, N, abstol=abstol, p=p)


def CG_GPU_sparse(x, A_csr, b, abstol=None, p=2):
    """CG with a CSR SpMV of the assembled operator (split_matrix_free.jl:744-769)
    -> (iters, abstol, last ||r||).  A_csr is the (row_ptr, col, val) triple of
    SbpPoisson.csr_assemble."""
    h = handle(_side(b), p, device=b.device)
    abstol = math.sqrt(2.220446049250313e-16) if abstol is None else abstol
    _, res = h.csr_cg(A_csr, b.reshape(-1), x.reshape(-1), abstol=abstol)
    return res.iters, abstol, res.resnorm


def initial_guess_interpolation_CG_Matrix_Free_GPU(A_GPU, b_GPU, b_2h, x, Nx_2h, abstol=None, p=2):
    """Coarse exact solve, bilinear interpolation, matrix-free CG
    (initial_guess_cg.jl:408-430) -> (x, iters, final ||r||).  A_GPU is unused:
    the operator is the handle's.  The reference's CG here stops at the
    absolute sqrt(eps); its reversal of the guess (:416) is a layout change
    that this layout does not need.  x receives the solution."""
    N = _side(b_GPU)
    if (N + 1) // 2 != Nx_2h:
        raise ValueError("b_2h must be on the grid of side (N+1)/2")
    h = handle(N, p, device=b_GPU.device)
    abstol = math.sqrt(2.220446049250313e-16) if abstol is None else abstol
    _, res = h.cg_interp(b_GPU.reshape(-1), b_2h.reshape(-1), x=x.reshape(-1), abstol=abstol)
    return x, res.iters, res.resnorm


def initial_guess_interpolation_CG_GPU(A_GPU, b_GPU, b_2h, x, Nx_2h, abstol=None, p=2):
    """initial_guess_interpolation_CG_GPU (initial_guess_cg.jl:401-406): the
    same nested solve (the reference runs its CG on the CSR matrix A_GPU; here
    the matrix-free apply of the same operator) -> (x, iters, final ||r||)."""
    return initial_guess_interpolation_CG_Matrix_Free_GPU(A_GPU, b_GPU, b_2h, x, Nx_2h, abstol=abstol, p=p)


def initial_guess_interpolation_three_level_CG_GPU(A_GPU, A_2h_GPU, b_GPU, b_2h_GPU, b_2h, b_4h, x, Nx_2h, Nx_4h,
                                                   abstol=None, p=2):
    """Three-level nested solve (initial_guess_cg.jl:447-456): exact solve on
    the 4h grid, CG on the 2h grid from its interpolation, CG on the fine grid
    from that -> (x, iters_2h, iters_h, ||r||_2h, ||r||_h).  b_2h duplicates
    b_2h_GPU in the reference (host copy for its LU) and is unused here."""
    N = _side(b_GPU)
    if (N + 1) // 2 != Nx_2h or (Nx_2h + 1) // 2 != Nx_4h:
        raise ValueError("coarse grids must have sides (N+1)/2 and (N+3)/4")
    h = handle(N, p, device=b_GPU.device)
    abstol = math.sqrt(2.220446049250313e-16) if abstol is None else abstol
    b4 = torch.as_tensor(b_4h, dtype=torch.float64, device=b_GPU.device).reshape(-1)
    _, res = h.cg_interp(b_GPU.reshape(-1), b_2h_GPU.reshape(-1), b4h=b4, x=x.reshape(-1), abstol=abstol)
    return x, res.level_iters[0], res.level_iters[1], res.level_resnorms[0], res.level_resnorms[1]
