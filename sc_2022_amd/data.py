"""Synthetic problem data, generated on the device.

Two families:

* the manufactured solution u = sin(pi x) sin(pi y) + x y of the plain CG /
  MG-PCG workloads (x on the slow, Dirichlet index j; y on the fast, Neumann
  index i), whose source and boundary data follow the reference's convention
  of level_2_multigrid_new.jl:36-38,173-177 (f = Laplace(u), g_W / g_E the
  Dirichlet values, g_S / g_N the outward normal derivatives);
* a seeded sum of Fourier modes: f = sum_m a_m cos(pi (kx_m x + ky_m y) + phi_m)
  with a_m = 1 / (kx^2 + ky^2 + 1), and each face datum
  g = sum_m a_m cos(pi k_m t + phi_m) with a_m = 1 / (k^2 + 1).  The integers
  kx, ky, k in 1..kmax and the phases phi in [0, 2 pi) come from a splitmix64
  stream of the seed in a fixed order (source modes, then W, E, S, N).

Grid functions are flat tensors of N*N elements, element (i, j) at i + j*N.
"""
from __future__ import annotations

import math

import torch

_M64 = (1 << 64) - 1


def _splitmix64(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


class _Stream:
    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> int:
        self.s, v = _splitmix64(self.s)
        return v

    def unit(self) -> float:               # [0, 1) with 53 bits
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def integer(self, lo: int, hi: int) -> int:   # lo..hi inclusive
        return lo + self.next() % (hi - lo + 1)


def fourier_modes(seed: int, n_src: int = 64, n_bnd: int = 8, kmax: int = 32):
    """(src, bnd): src = [(kx, ky, a, phi)] * n_src; bnd = 4 lists (W, E, S, N)
    of [(k, a, phi)] * n_bnd."""
    rs = _Stream(seed)
    src = []
    for _ in range(n_src):
        kx = rs.integer(1, kmax)
        ky = rs.integer(1, kmax)
        ph = 2.0 * math.pi * rs.unit()
        src.append((kx, ky, 1.0 / (kx * kx + ky * ky + 1.0), ph))
    bnd = []
    for _ in range(4):
        modes = []
        for _ in range(n_bnd):
            k = rs.integer(1, kmax)
            ph = 2.0 * math.pi * rs.unit()
            modes.append((k, 1.0 / (k * k + 1.0), ph))
        bnd.append(modes)
    return src, bnd


def grid(N: int, device=None, dtype=torch.float64):
    return torch.linspace(0.0, 1.0, N, dtype=dtype, device=device)


def fourier_fields(N: int, seed: int, device=None):
    """(f, gW, gE, gS, gN) of the seeded Fourier-mode problem on an N x N grid.

    The source is formed as two rank-64 products,
        F(i, j) = sum_m a_m [cos(pi ky y_i) cos(pi kx x_j + phi) - sin(pi ky y_i) sin(pi kx x_j + phi)],
    so generating a large grid costs two GEMMs instead of 64 full-grid passes."""
    src, bnd = fourier_modes(seed)
    t = grid(N, device)
    kx = torch.tensor([m[0] for m in src], dtype=torch.float64, device=device)
    ky = torch.tensor([m[1] for m in src], dtype=torch.float64, device=device)
    a = torch.tensor([m[2] for m in src], dtype=torch.float64, device=device)
    ph = torch.tensor([m[3] for m in src], dtype=torch.float64, device=device)
    ay = math.pi * t[:, None] * ky[None, :]                 # (N_i, M)
    ax = math.pi * t[:, None] * kx[None, :] + ph[None, :]   # (N_j, M)
    cy, sy = torch.cos(ay) * a[None, :], torch.sin(ay) * a[None, :]
    cx, sx = torch.cos(ax), torch.sin(ax)
    # F[j, i] (row-major over j) = element (i, j) at i + j N
    F = torch.matmul(cx, cy.t().contiguous()) - torch.matmul(sx, sy.t().contiguous())
    f = F.reshape(-1).contiguous()
    g = []
    for modes in bnd:
        k = torch.tensor([m[0] for m in modes], dtype=torch.float64, device=device)
        am = torch.tensor([m[1] for m in modes], dtype=torch.float64, device=device)
        pm = torch.tensor([m[2] for m in modes], dtype=torch.float64, device=device)
        g.append((am[None, :] * torch.cos(math.pi * t[:, None] * k[None, :] + pm[None, :])).sum(dim=1).contiguous())
    return (f, g[0], g[1], g[2], g[3])


def manufactured_fields(N: int, device=None):
    """(f, gW, gE, gS, gN, u) for u = sin(pi x) sin(pi y) + x y."""
    t = grid(N, device)
    y = t[None, :]    # fast index i (columns of the row-major [j, i] view)
    x = t[:, None]    # slow index j
    U = torch.sin(math.pi * x) * torch.sin(math.pi * y) + x * y
    Fm = -2.0 * math.pi ** 2 * torch.sin(math.pi * x) * torch.sin(math.pi * y)
    gW = torch.zeros_like(t)
    gE = t.clone()
    gS = -(math.pi * torch.sin(math.pi * t) + t)
    gN = -math.pi * torch.sin(math.pi * t) + t
    return (Fm.reshape(-1).contiguous(), gW, gE, gS, gN, U.reshape(-1).contiguous())


def reference_fields(N: int, device=None):
    """(f, gW, gE, gS, gN, u) of the reference's own manufactured solution
    u = sin(pi (x + y)) (level_2_multigrid_new.jl:36-38, 171-177): f = Laplace(u),
    Dirichlet data on the west / east faces, Neumann data on the south /
    north faces, in the flat layout (element (i, j) at i + j N)."""
    t = grid(N, device)
    U = torch.sin(math.pi * t[:, None] + math.pi * t[None, :])   # symmetric in (i, j)
    u = U.reshape(-1).contiguous()
    f = -2.0 * math.pi ** 2 * u
    gW = torch.sin(math.pi * t)
    gE = -torch.sin(math.pi * t)
    gS = -math.pi * torch.cos(math.pi * t)
    gN = math.pi * torch.cos(math.pi * t + math.pi)
    return (f, gW, gE, gS, gN, u)


def fourier_rhs(op, seed: int):
    """b of the seeded Fourier-mode problem for an SbpPoisson operator."""
    return op.assemble_rhs(*fourier_fields(op.N, seed, op.device))


def batch_specs(seed: int, count: int = 256, kmin: int = 6, kmax: int = 10, orders=(2, 4)):
    """Seeded list of (N, p, data_seed) for a batch of uneven problems:
    N = 2^k + 1 with k uniform in kmin..kmax, p uniform over `orders`."""
    rs = _Stream(seed ^ 0x5BD1E995)
    out = []
    for _ in range(count):
        k = rs.integer(kmin, kmax)
        p = orders[rs.integer(0, len(orders) - 1)]
        out.append((2 ** k + 1, p, rs.next() & 0x7FFFFFFF))
    return out
