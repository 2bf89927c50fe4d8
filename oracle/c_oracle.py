"""ctypes wrapper of the C oracle (oracle/sbp_oracle.c -> oracle/liboracle.so).

Test infrastructure and CPU baseline only.  ``build()`` runs the Makefile in
this directory (gcc, OpenMP)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_apply.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_apply_plain.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_diag.argtypes = [C.c_int, C.c_int, dp]
        L.orc_csr.restype = C.c_long
        L.orc_csr.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_rhs.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 5 + [dp]
        L.orc_cg_csr.argtypes = [C.c_long, ip, ip, dp, dp, dp, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.orc_mgpcg.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_double, C.POINTER(C.c_double)]
        L.orc_hnorm.restype = C.c_double
        L.orc_hnorm.argtypes = [C.c_int, C.c_int, dp]
        L.orc_rhs_transpose.argtypes = [C.c_int, C.c_int, dp] + [C.c_void_p] * 5
        L.orc_adjoint_grad.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, C.POINTER(C.c_double)] + \
            [C.c_void_p] * 6
        L.orc_fourier_field.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_fourier_edge.argtypes = [C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def set_threads(t):
    lib().orc_set_threads(int(t))


def apply(p, N, u):
    out = np.empty(N * N)
    assert lib().orc_apply(p, N, np.ascontiguousarray(u, dtype=np.float64), out) == 0
    return out


def apply_plain(p, N, u):
    """A u with each row a plain sum of the rounded FP64 entries (the
    evaluation the oracle's apply avoids; kept to show the difference)."""
    out = np.empty(N * N)
    assert lib().orc_apply_plain(p, N, np.ascontiguousarray(u, dtype=np.float64), out) == 0
    return out


def diag(p, N):
    d = np.empty(N * N)
    assert lib().orc_diag(p, N, d) == 0
    return d


def csr(p, N):
    L = lib()
    nnz = L.orc_csr(p, N, None, None, None)
    assert nnz > 0
    rp = np.empty(N * N + 1, dtype=np.int32)
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz)
    L.orc_csr(p, N, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return rp, col, val


def rhs(p, N, f=None, gW=None, gE=None, gS=None, gN=None):
    b = np.empty(N * N)
    arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (f, gW, gE, gS, gN)]
    args = [None if a is None else a.ctypes.data for a in arrs]
    assert lib().orc_rhs(p, N, *args, b) == 0
    return b


def cg_csr(A, b, abstol, maxiter=None, x=None):
    rp, col, val = A
    n = len(rp) - 1
    x = np.zeros(n) if x is None else np.array(x, dtype=np.float64)
    rn = C.c_double()
    it = lib().orc_cg_csr(n, rp, col, val, np.ascontiguousarray(b, dtype=np.float64), x, abstol,
                          n if maxiter is None else maxiter, C.byref(rn))
    return x, it, rn.value


def mgpcg(p, N, b, abstol, maxiter=200, nu_pre=2, nu_post=2, omega=2.0 / 3.0, rscale=4.0, x=None):
    x = np.zeros(N * N) if x is None else np.array(x, dtype=np.float64)
    rn = C.c_double()
    it = lib().orc_mgpcg(p, N, np.ascontiguousarray(b, dtype=np.float64), x, abstol, maxiter, nu_pre, nu_post,
                         omega, rscale, C.byref(rn))
    if it < 0:
        raise ValueError("orc_mgpcg: unsupported size")
    return x, it, rn.value


def hnorm(p, N, v):
    """sqrt(v' H~ v) on the N x N grid of order p."""
    r = lib().orc_hnorm(p, N, np.ascontiguousarray(v, dtype=np.float64))
    assert r >= 0.0
    return r


def rhs_transpose(p, N, lam):
    """Adjoint of ``rhs`` with respect to (f, gW, gE, gS, gN) at lam."""
    df = np.empty(N * N)
    g = [np.empty(N) for _ in range(4)]
    assert lib().orc_rhs_transpose(p, N, np.ascontiguousarray(lam, dtype=np.float64), df.ctypes.data,
                                   *[v.ctypes.data for v in g]) == 0
    return (df, *g)


def adjoint_grad(p, N, u, u_obs, abstol=None, maxiter=200, rtol=1e-14):
    """FP64 gradient of L = 1/2 (u - u_obs)' H~ (u - u_obs), u = A^-1 b(f, g):
    (L, lam, df, dgW, dgE, dgS, dgN, iterations of the adjoint MG-PCG).  The
    adjoint solve stops at ||r|| < abstol, or (abstol None) at
    rtol ||H~ (u - u_obs)||."""
    if abstol is None:
        abstol = -rtol
    lam = np.empty(N * N)
    df = np.empty(N * N)
    g = [np.empty(N) for _ in range(4)]
    loss = C.c_double()
    it = lib().orc_adjoint_grad(p, N, np.ascontiguousarray(u, dtype=np.float64),
                                np.ascontiguousarray(u_obs, dtype=np.float64), abstol, maxiter, C.byref(loss),
                                lam.ctypes.data, df.ctypes.data, *[v.ctypes.data for v in g])
    if it < 0:
        raise ValueError("orc_adjoint_grad: unsupported size")
    return (loss.value, lam, df, *g, it)


def fourier_fields(N, seed):
    """(f, gW, gE, gS, gN) of the seeded Fourier-mode problem (mode list of
    sbp_oracle.fourier_modes), evaluated mode by mode in C."""
    from .sbp_oracle import fourier_modes
    src, bnd = fourier_modes(seed)
    F = np.empty(N * N)
    lib().orc_fourier_field(N, len(src), np.ascontiguousarray(np.array(src, dtype=np.float64).ravel()), F)
    g = []
    for modes in bnd:
        v = np.empty(N)
        lib().orc_fourier_edge(N, len(modes), np.ascontiguousarray(np.array(modes, dtype=np.float64).ravel()), v)
        g.append(v)
    return (F, *g)


def fourier_rhs(p, N, seed):
    """Right-hand side of the seeded Fourier-mode problem (same mode list as
    sbp_oracle.fourier_modes), built on the host."""
    return rhs(p, N, *fourier_fields(N, seed))


def solve_reference(p, N, b, rtol=1e-14, maxiter=100):
    """FP64 reference solution x = A^-1 b by the C oracle's MG-PCG to
    ||r|| < rtol ||b|| (the headline errors are measured against it);
    raises if it does not get there."""
    bn = float(np.linalg.norm(b))
    x, it, rn = mgpcg(p, N, b, rtol * bn, maxiter=maxiter)
    if not rn < rtol * bn:
        raise RuntimeError(f"oracle MG-PCG stopped at ||r||/||b|| = {rn / bn:.3g} after {it} iterations")
    return x, it, rn / bn


def time_mgpcg_fourier(N, p, seed, rtol, threads, runs=1, warmup=0, b=None):
    """FP64 MG-PCG solves of the seeded Fourier problem on `threads` host
    threads (the bench's CPU-baseline sample): `warmup` untimed solves, then
    `runs` timed ones; reports the median."""
    set_threads(threads)
    b = fourier_rhs(p, N, seed) if b is None else b
    abstol = rtol * float(np.linalg.norm(b))
    for _ in range(warmup):
        mgpcg(p, N, b, abstol)
    ts = []
    for _ in range(max(1, runs)):
        t0 = time.perf_counter()
        x, it, rn = mgpcg(p, N, b, abstol)
        ts.append(time.perf_counter() - t0)
    return {"seconds": float(np.median(ts)), "all_seconds": ts, "iters": it, "resnorm": rn}


def time_cg_csr_sample(N, p, b, iters, threads=1):
    """The reference's CPU path (CSR assembly + CG_CPU) on `threads` host
    threads: the assembly time and the time of `iters` CG iterations (a
    bounded sample; the full solve needs thousands)."""
    set_threads(threads)
    t0 = time.perf_counter()
    A = csr(p, N)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, it, rn = cg_csr(A, b, 0.0, maxiter=iters)
    t_cg = time.perf_counter() - t0
    return {"assembly_s": t_asm, "iters": it, "s_per_iter": t_cg / max(1, it), "nnz": int(len(A[2]))}
