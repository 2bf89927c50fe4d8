#!/usr/bin/env python3
"""Benchmark of the SBP-SAT Poisson hot path on one MI355X.

Workload (one "step"): one mixed-precision MG-PCG solve (FP32 inner
MG-preconditioned CG inside FP64 iterative refinement) of the 4097 x 4097,
SBP order 4 problem with a seeded random Fourier-mode source (64 modes) and
boundary data (8 modes per face), from x0 = 0 to a relative residual that
gives a relative H-norm error <= 1e-10 against the FP64 oracle's solution
(the C restatement in oracle/, MG-PCG to 1e-14 on the host cores).  The same problem is also solved with the plain FP64
MG-PCG (reported beside it), and the matrix-free apply is timed alone with
rotating buffers larger than the 256 MB Infinity Cache for the roofline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4097] [--p 4]

Prints one JSON line.  --gpus above 1 is refused: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
REL_ERR_TARGET = 1e-10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4097)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2022)
    ap.add_argument("--rtol", type=float, default=1e-11,
                    help="FP64 leg: stop when ||r|| < rtol ||b||")
    ap.add_argument("--mixed-rtol", type=float, default=1e-10,
                    help="mixed leg: stop when the FP64 ||r|| < mixed_rtol ||b||")
    ap.add_argument("--inner-rtol", type=float, default=1e-5,
                    help="mixed: relative tolerance of each FP32 inner solve")
    ap.add_argument("--ir-rtol", type=float, default=0.0,
                    help="mixed: also stop when the correction-based error estimate <= ir_rtol ||x||")
    ap.add_argument("--no-extra", action="store_true", help="skip the 8193^2, CSR comparison, initial-guess CG, reference-algorithm, variable-coefficient and adjoint legs")
    ap.add_argument("--extra", default="", help="comma list of the extra legs to run (default: all of them)")
    ap.add_argument("--tol-sweep", action="store_true",
                    help="print iterations, error and time per stopping rule for both legs, then exit")
    ap.add_argument("--apply-reps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--only-apply", action="store_true", help="time the apply only (profiling runs)")
    ap.add_argument("--batch", type=int, default=256, help="problems in the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-rtol", type=float, default=1e-10,
                    help="batched leg: stop each problem when its FP64 ||r|| < batch_rtol ||b|| (the loosest rule "
                         "that meets rel err <= 1e-10 against the oracle for every problem: bench.py --batch-sweep)")
    ap.add_argument("--batch-inner-rtol", type=float, default=3e-6,
                    help="batched leg: relative tolerance of the FP32 inner solves")
    ap.add_argument("--batch-sweep", action="store_true",
                    help="print solves/s and the worst error against the oracle for a grid of batch stopping rules, then exit")
    return ap.parse_args()


def fourier_problem(h, seed):
    """b of the seeded Fourier-mode problem, generated on the device."""
    from sc_2022_amd.data import fourier_rhs
    return fourier_rhs(h, seed)


def oracle_reference(h, b, p, seed, torch, keep_b=None):
    """The FP64 oracle's solution of the seeded Fourier problem, against
    which every headline error is measured: the C oracle
    (oracle/sbp_oracle.c, an independent CPU restatement) builds the
    right-hand side on the host from the same mode list and solves it by
    MG-PCG to ||r|| < 1e-14 ||b|| on the host cores.  Returns the solution on
    the device, its H-norm, and what was checked (host vs device right-hand
    side, oracle iterations and residual, host seconds)."""
    from oracle import c_oracle
    N = h.N
    c_oracle.set_threads(min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    bh = c_oracle.fourier_rhs(p, N, seed)
    x, it, rr = c_oracle.solve_reference(p, N, bh)
    secs = time.perf_counter() - t0
    db = float(np.linalg.norm(b.cpu().numpy() - bh) / np.linalg.norm(bh))
    if db > 1e-13:
        raise RuntimeError(f"device and oracle right-hand sides differ: {db:.3g}")
    xd = torch.as_tensor(x, device=h.device)
    out = {"x": xd, "hnorm": h.hnorm(xd),
           "what": f"C oracle FP64 MG-PCG to ||r|| < 1e-14 ||b|| on the host ({N}^2, p={p})",
           "oracle_iters": it, "oracle_rel_resid": rr, "rhs_rel_diff": db, "host_s": secs}
    if keep_b is not None:
        keep_b.append(bh)
    return out


def vcycle_bytes(sides, es, fuse_n=1025):
    """Algorithmic bytes of one 2 + 2 V-cycle from the kernels it launches
    (distinct arrays read once plus arrays written once, coarse arrays at
    their own size, es bytes per value): on a level above fuse_n^2 points the
    separate passes -- first two sweeps (b in, x out: 2 n), residual and
    restriction (x, b in, coarse out: 2.25 n), prolongation and sweep (x,
    coarse, b in, x out: 3.25 n), last sweep (x, b in, x out: 3 n); on a fused
    level smooth2_restrict (2.25 n) and prolong_smooth2 (3.25 n); the coarsest
    level reads the FP64 inverse (n_c^2 x 8) and b and writes x."""
    tot = 0
    for l, N in enumerate(sides[:-1]):
        n = N * N
        tot += (5.5 if n <= fuse_n * fuse_n else 10.5) * n * es
    nc = sides[-1] ** 2
    return tot + nc * nc * 8 + 2 * nc * es


def mgpcg_bytes(sides, es, iters, residual=True):
    """Algorithmic bytes of an MG-PCG solve of `iters` iterations: per
    iteration the search-direction pass (z, p, x in; p, Ap, x out: 6 n), the
    residual update (r, Ap in; r out: 3 n) and a V-cycle; at the start the
    residual (x, b in, r out: 3 n; not for an inner solve from zero, whose
    residual is its right-hand side), a V-cycle and the copy p = z (2 n); at
    the end the pending x update (3 n)."""
    n = sides[0] ** 2
    vc = vcycle_bytes(sides, es)
    start = (3 * n * es if residual else 0) + vc + 2 * n * es
    return iters * (9 * n * es + vc) + start + 3 * n * es


def time_apply(h, torch, dtype, reps, nbuf):
    """Average duration of the stencil apply kernel over reps launches that
    rotate over nbuf input/output pairs (HIP events on the launch stream).
    The timed loop calls sbp_apply through ctypes with the pointers prepared
    beforehand, so the host issues launches faster than the kernel runs and
    the events see kernel time, not Python time between launches."""
    import ctypes as C
    from sc_2022_amd import _capi as capi
    n = h.n
    ins = [torch.randn(n, dtype=dtype, device=h.device) for _ in range(nbuf)]
    outs = [torch.empty(n, dtype=dtype, device=h.device) for _ in range(nbuf)]
    for k in range(nbuf):
        h.apply(ins[k], outs[k])          # checked launches; also sets the handle's stream
    torch.cuda.synchronize()
    fn, hh = h.lib.sbp_apply, h._h
    dt = capi.SBP_F64 if dtype == torch.float64 else capi.SBP_F32
    pin = [C.c_void_p(t.data_ptr()) for t in ins]
    pout = [C.c_void_p(t.data_ptr()) for t in outs]
    for r in range(max(50, reps // 2)):   # untimed: let the clocks settle under this load
        fn(hh, 0, pin[r % nbuf], pout[r % nbuf], 1, dt)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    bad = 0
    e0.record(s)
    for r in range(reps):
        k = r % nbuf
        bad |= fn(hh, 0, pin[k], pout[k], 1, dt)
    e1.record(s)
    torch.cuda.synchronize()
    if bad:
        raise RuntimeError(f"sbp_apply failed in the timed loop (status bits {bad})")
    ms = e0.elapsed_time(e1) / reps
    del ins, outs
    return ms


def time_copy(h, torch, reps, nbytes, nbuf):
    src = [torch.empty(nbytes // 8, dtype=torch.float64, device=h.device).fill_(1.0) for _ in range(nbuf)]
    dst = [torch.empty_like(src[0]) for _ in range(nbuf)]
    import ctypes as C
    for k in range(nbuf):
        h.copy(src[k], dst[k])
    torch.cuda.synchronize()
    fn, hh = h.lib.sbp_copy, h._h
    ps = [C.c_void_p(t.data_ptr()) for t in src]
    pd = [C.c_void_p(t.data_ptr()) for t in dst]
    nb = src[0].numel() * src[0].element_size()
    for r in range(max(50, reps // 2)):   # untimed warm-up, as for the apply
        fn(hh, ps[r % nbuf], pd[r % nbuf], nb)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    bad = 0
    e0.record(s)
    for r in range(reps):
        bad |= fn(hh, ps[r % nbuf], pd[r % nbuf], nb)
    e1.record(s)
    torch.cuda.synchronize()
    if bad:
        raise RuntimeError(f"sbp_copy failed in the timed loop (status bits {bad})")
    return e0.elapsed_time(e1) / reps


def batch_leg(a, dev, torch):
    """Uneven batch: a.batch problems, N = 2^k+1 with k uniform in 6..10 and p
    in {2, 4} drawn from the seed, each with its own Fourier-mode data; mixed
    MG-PCG with per-problem masks, one ragged handle per order (shared
    multigrid levels, one launch per pass).  Reports batch solves/s, p50/p99
    of the per-problem device times and the worst relative H-norm error
    against the C oracle's FP64 solution of every problem (host right-hand
    side from the same modes, MG-PCG to 1e-14): at the leg's rule (the
    loosest that meets 1e-10 for every problem, bench.py --batch-sweep) and,
    beside it, at the strict rule ||r|| < 1e-12 ||b|| with FP32 inner solves
    to 1e-5."""
    from oracle import c_oracle
    from sc_2022_amd.batch import UnevenBatch
    from sc_2022_amd.data import batch_specs
    specs = batch_specs(a.seed, count=a.batch)

    def make(inner):
        ub = UnevenBatch([(N, p) for N, p, _ in specs], device=dev, ir_rtol=a.ir_rtol, inner_rtol=inner)
        for q, (N, p, s) in enumerate(specs):
            ub.set_fourier(q, s)
        torch.cuda.synchronize()
        return ub

    ub = make(a.batch_inner_rtol)
    bnorm = [float(torch.linalg.vector_norm(ub.rhs(q))) for q in range(len(specs))]
    c_oracle.set_threads(min(16, os.cpu_count() or 1))
    t1 = time.perf_counter()
    xo, rhs_diff = [], 0.0
    for q, (N, p, s) in enumerate(specs):
        bh = c_oracle.fourier_rhs(p, N, s)
        rhs_diff = max(rhs_diff, float(np.linalg.norm(ub.rhs(q).cpu().numpy() - bh) / np.linalg.norm(bh)))
        xo.append(c_oracle.solve_reference(p, N, bh)[0])
    t_orc = time.perf_counter() - t1

    def run(ub, rtol, inner):
        tol = [rtol * bn for bn in bnorm]
        for _ in range(max(1, a.warmup)):
            out = ub.solve(tol, mixed=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.batch_steps):
            out = ub.solve(tol, mixed=True)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.batch_steps
        errs = []
        for q in range(len(specs)):
            x = torch.as_tensor(xo[q], device=dev)
            errs.append(ub.hnorm(q, ub.solution(q) - x) / ub.hnorm(q, x))
        sec = sorted(o["seconds"] for o in out)
        return {"rule": f"per problem FP64 ||r|| < {rtol:g} ||b||, FP32 inner solves to {inner:g}",
                "solves_per_s": len(specs) / dt, "ms_per_batch": dt * 1e3,
                "p50_ms": 1e3 * float(np.percentile(sec, 50)), "p99_ms": 1e3 * float(np.percentile(sec, 99)),
                "max_outer": max(o["iters"] for o in out), "max_inner": max(o["inner_iters"] for o in out),
                "max_rel_err_vs_oracle": max(errs), "all_converged": all(o["status"] == 0 for o in out)}

    if a.batch_sweep:
        for rt in (1e-10, 5e-11, 3e-11, 1e-11, 1e-12):
            print(json.dumps({"leg": "batched", "inner_rtol": a.batch_inner_rtol,
                              **run(ub, rt, a.batch_inner_rtol)}), flush=True)
        ub.close()
        return None
    res = run(ub, a.batch_rtol, a.batch_inner_rtol)
    ub.close()
    strict = make(1e-5)
    res["at_strict_rule"] = run(strict, 1e-12, 1e-5)
    strict.close()
    pts = sum(N * N for N, _, _ in specs)
    res.update({"problems": len(specs), "handles": 2, "points": pts,
                "design": "one ragged handle per order (shared multigrid levels, one launch per pass), "
                          "the two orders on two streams",
                "oracle_rhs_rel_diff": rhs_diff, "oracle_host_s": t_orc,
                "sizes": {f"{N}^2 p{p}": sum(1 for n, pp, _ in specs if (n, pp) == (N, p))
                          for N, p in sorted({(n, pp) for n, pp, _ in specs})}})
    return res


def csr_leg(h, torch, apply_us):
    """The assembled CSR operator (the comparison path of the reference's
    CG_GPU_sparse, which multiplies with CUSPARSE) against the matrix-free
    apply: time per product of the hand-written CSR SpMV and of rocSPARSE's
    (csr_adaptive and csr_rowsplit, analysis done once, outside the timing),
    with algorithmic bytes (values 8 B and column indices 4 B per nonzero,
    row pointers, x read once, y written)."""
    rp, col, val = h.csr_assemble()
    n, nnz = h.n, col.numel()
    x = torch.randn(n, dtype=torch.float64, device=h.device)
    y = torch.empty_like(x)
    nbytes = 12 * nnz + 4 * (n + 1) + 16 * n

    def timed(fn, reps=50):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(torch.cuda.current_stream())
        for _ in range(reps):
            fn()
        e1.record(torch.cuda.current_stream())
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    us = timed(lambda: h.csr_spmv((rp, col, val), x, y))
    ref = y.clone()
    res = {"nnz": nnz, "bytes": nbytes, "spmv_us": us, "spmv_GBs": nbytes / (us * 1e-6) / 1e9,
           "matrix_free_apply_us": apply_us, "matrix_free_speedup": us / apply_us}
    for alg, name in ((0, "rocsparse_adaptive"), (1, "rocsparse_rowsplit")):
        try:
            h.csr_vendor_attach((rp, col, val), alg)
            vus = timed(lambda: h.csr_vendor_spmv(x, y))
            diff = float((y - ref).abs().max() / ref.abs().max())
            res[name] = {"spmv_us": vus, "spmv_GBs": nbytes / (vus * 1e-6) / 1e9,
                         "matrix_free_speedup": vus / apply_us, "max_rel_diff_vs_hand_csr": diff}
        except Exception as e:  # reported, never silently replaced
            res[name] = {"error": f"{type(e).__name__}: {e}"}
    del rp, col, val
    return res


def leg_8193(a, dev, torch):
    """Mixed precision and FP64 MG-PCG at 8193^2 (p = 4, the largest single-GPU grid), each leg
    with the loosest rule that meets rel err <= 1e-10 at this size
    (bench.py --n 8193 --tol-sweep): FP64 ||r|| < rtol ||b||, mixed
    ||r|| < mixed_rtol ||b|| with inner FP32 solves to 3e-6; errors against
    the FP64 oracle's solution (oracle_reference)."""
    from sc_2022_amd.data import fourier_rhs
    from sc_2022_amd.solver import SbpPoisson
    N = 8193
    inner = 3e-6
    h = SbpPoisson(N, a.p, device=dev, ir_rtol=a.ir_rtol, inner_rtol=inner)
    b = fourier_rhs(h, a.seed)
    bn = float(torch.linalg.vector_norm(b))
    ref = oracle_reference(h, b, a.p, a.seed, torch)
    xref, href = ref.pop("x"), ref.pop("hnorm")
    x = torch.zeros_like(b)
    out = {}
    for leg in ("mixed", "f64"):
        def run():
            x.zero_()
            if leg == "mixed":
                return h.mgpcg_mixed(b, x, abstol=a.mixed_rtol * bn)[1]
            return h.mgpcg(b, x, abstol=a.rtol * bn, maxiter=200)[1]
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            r = run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 3
        out[leg] = {"ms": dt * 1e3, "solves_per_s": 1.0 / dt, "rel_err": h.hnorm(x - xref) / href,
                    "iters": r.iters, "inner": r.inner_iters}
    out["speedup_mixed_vs_f64"] = out["f64"]["ms"] / out["mixed"]["ms"]
    out["rules"] = {"mixed_rtol": a.mixed_rtol, "inner_rtol": inner, "f64_rtol": a.rtol}
    out["rel_err_reference"] = ref
    out["meets_rel_err_target"] = bool(max(out["mixed"]["rel_err"], out["f64"]["rel_err"]) <= REL_ERR_TARGET)
    h.close()
    return out


def nested_leg(a, dev, torch, N=1025, p=2, rtol=1e-10):
    """Initial-guess interpolation CG (the reference's
    initial_guess_interpolation_CG family) on the manufactured problem at
    N^2, p = 2, against matrix-free CG from zero and MG-PCG, all stopped at
    ||r|| < rtol ||b||: fine-grid iterations, time per solve (the nested
    solves include their coarse solves), H-norm error against u."""
    from sc_2022_amd.data import manufactured_fields
    from sc_2022_amd.solver import SbpPoisson
    hs = [SbpPoisson(N, p, device=dev), SbpPoisson((N + 1) // 2, p, device=dev),
          SbpPoisson((N + 3) // 4, p, device=dev)]
    bs, u = [], None
    for k, hk in enumerate(hs):
        f, gW, gE, gS, gN, uk = manufactured_fields(hk.N, dev)
        bs.append(hk.assemble_rhs(f, gW, gE, gS, gN))
        if k == 0:
            u = uk
    h = hs[0]
    tol = rtol * float(torch.linalg.vector_norm(bs[0]))
    x = torch.zeros_like(bs[0])
    runs = {
        "cg_from_zero": lambda: h.cg(bs[0], x.zero_(), abstol=tol)[1],
        "interp_2level": lambda: h.cg_interp(bs[0], bs[1], x=x, abstol=tol)[1],
        "interp_3level": lambda: h.cg_interp(bs[0], bs[1], bs[2], x=x, abstol=tol)[1],
        "mgpcg": lambda: h.mgpcg(bs[0], x.zero_(), abstol=tol)[1],
    }
    out = {"workload": f"N={N}, p={p}, u = sin(pi x) sin(pi y) + x y, stop ||r|| < {rtol:g} ||b||"}
    uh = h.hnorm(u)
    for name, fn in runs.items():
        r = fn()
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"ms": dt * 1e3, "iters": r.iters, "converged": r.converged,
                     "h_err_vs_u": h.hnorm(x - u) / uh}
        if r.level_iters:
            out[name]["level_iters"] = r.level_iters
    for hk in hs:
        hk.close()
    return out


def reference_modes_leg(a, dev, torch, N=1025, p=2):
    """The reference's own algorithms on this hardware, beside the build's:
    MGCG with the reference's preconditioner (two-level, 3 + 3 Richardson
    sweeps, exact coarse solve) against the multi-level MG-PCG, both to
    ||r|| < 1e-10 ||b||; and the Poisson_CG.jl-style efficiency report of
    plain CG (to 1e-8 ||b||), matrix-free and CSR: time in the operator
    products (iterations x one product) over the whole solve, and the rest as
    overhead."""
    from sc_2022_amd import reference_api as R
    from sc_2022_amd.data import manufactured_fields
    from sc_2022_amd.solver import SbpPoisson
    href = SbpPoisson(N, p, device=dev, **R.reference_mg_params(N, 2))
    h = SbpPoisson(N, p, device=dev)
    f, gW, gE, gS, gN, u = manufactured_fields(N, dev)
    b = h.assemble_rhs(f, gW, gE, gS, gN)
    bn = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)

    def timed(fn, reps=3):
        r = fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / reps * 1e3

    out = {"workload": f"N={N}, p={p}, u = sin(pi x) sin(pi y) + x y"}
    for name, hh in (("mgcg_reference_two_level", href), ("mgpcg_multilevel", h)):
        r, ms = timed(lambda: hh.mgpcg(b, x.zero_(), abstol=1e-10 * bn, maxiter=500)[1])
        out[name] = {"ms": ms, "iters": r.iters, "converged": r.converged}
    out["mgcg_speedup_multilevel_vs_reference"] = out["mgcg_reference_two_level"]["ms"] / out["mgpcg_multilevel"]["ms"]
    # efficiency report
    apply_ms = time_apply(h, torch, torch.float64, 200, 8)
    r, cg_ms = timed(lambda: h.cg(b, x.zero_(), abstol=1e-8 * bn)[1], reps=2)
    csr = h.csr_assemble()
    y = torch.empty_like(b)
    for _ in range(3):
        h.csr_spmv(csr, b, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        h.csr_spmv(csr, b, y)
    torch.cuda.synchronize()
    spmv_ms = (time.perf_counter() - t0) / 50 * 1e3
    rc, csr_ms = timed(lambda: h.csr_cg(csr, b, x.zero_(), abstol=1e-8 * bn)[1], reps=2)
    h.csr_vendor_attach(csr, 0)
    for _ in range(3):
        h.csr_vendor_spmv(b, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        h.csr_vendor_spmv(b, y)
    torch.cuda.synchronize()
    vspmv_ms = (time.perf_counter() - t0) / 50 * 1e3
    rv, vcsr_ms = timed(lambda: h.csr_vendor_cg(b, x.zero_(), abstol=1e-8 * bn)[1], reps=2)
    for name, it, tot, op in (("cg_matrix_free", r.iters, cg_ms, apply_ms), ("cg_csr", rc.iters, csr_ms, spmv_ms),
                              ("cg_csr_rocsparse", rv.iters, vcsr_ms, vspmv_ms)):
        t_op = it * op
        out[name] = {"iters": it, "ms": tot, "op_ms_each": op, "op_ms_in_cg": t_op,
                     "overhead_ms": tot - t_op, "efficiency": t_op / tot}
    href.close()
    h.close()
    return out


def variable_coefficient_leg(a, dev, torch, N=4097, p=4):
    """The variable-coefficient apply (sbp_create_var) at N^2 beside the
    constant-coefficient one: time per launch over rotating buffers and the
    algorithmic rate (2 n 8 bytes; the per-row coefficient bands are N (2H+1)
    doubles per axis and stay cache-resident)."""
    import numpy as np
    from sc_2022_amd.solver import SbpPoisson
    t = np.linspace(0.0, 1.0, N)
    hv = SbpPoisson(N, p, device=dev, coef=(1.0 + t / 2 + np.sin(3 * t) / 4, 1.0 + t ** 2))
    n = N * N
    nbuf = max(2, int((2 * 1024 ** 3) // (2 * n * 8)) + 1)
    ms = time_apply(hv, torch, torch.float64, 100, nbuf)
    # MG-PCG on the hierarchy rediscretised with the injected profiles, the
    # seeded Fourier-mode data, to ||r|| < 1e-10 ||b||
    b = fourier_problem(hv, a.seed)
    bn = float(torch.linalg.vector_norm(b))
    x = torch.zeros_like(b)

    def timed(solve, reps=3):
        solve()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = solve()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, res

    def fp64():
        x.zero_()
        return hv.mgpcg(b, x, abstol=1e-10 * bn, maxiter=200)[1]

    def fp32_vcycle():   # FP64 MG-PCG, V-cycle in FP32
        x.zero_()
        return hv.mgpcg_mixed(b, x, abstol=1e-10 * bn)[1]

    solve_ms, res = timed(fp64)
    mixed_ms, rm = timed(fp32_vcycle)
    hv.close()
    return {"workload": f"N={N}, p={p}, bx = 1 + x/2 + sin(3x)/4, by = 1 + y^2", "apply_us": ms * 1e3,
            "apply_GBs": 2 * n * 8 / (ms * 1e-3) / 1e9,
            "mgpcg": {"ms": solve_ms, "iters": res.iters, "converged": bool(res.converged), "rtol": 1e-10,
                      "levels": hv.nlevels},
            "mgpcg_fp32_vcycle": {"ms": mixed_ms, "iters": rm.iters, "converged": bool(rm.converged),
                                  "speedup": solve_ms / mixed_ms}}


def adjoint_leg(a, dev, torch, N=2049, p=4, ndir=32, eps=1e-6):
    """The differentiation workload: L = 1/2 (u - u_obs)' H~ (u - u_obs) at
    N^2, u = A^-1 b(f, g) with seeded Fourier-mode f and g, u_obs a seeded
    Fourier field (the source-type field of seed + 1) plus N(0, 1e-3) noise.
    Times the forward FP64 MG-PCG and the gradient (adjoint MG-PCG and the
    transposed right-hand-side map), both to 1e-13 relative residual; then
    checks the gradient against the C oracle's FP64 gradient (its own
    right-hand side, forward and adjoint MG-PCG on the host) and against
    central differences (step eps, solves to 1e-14) along ndir seeded
    directions in (f, g)."""
    from oracle import c_oracle
    from sc_2022_amd.data import fourier_fields
    from sc_2022_amd.solver import SbpPoisson
    h = SbpPoisson(N, p, device=dev)
    f, gW, gE, gS, gN = fourier_fields(N, a.seed, dev)
    b = h.assemble_rhs(f, gW, gE, gS, gN)
    c_oracle.set_threads(min(16, os.cpu_count() or 1))
    uo_h = c_oracle.fourier_fields(N, a.seed + 1)[0] + 1e-3 * np.random.default_rng(a.seed).standard_normal(N * N)
    u_obs = torch.as_tensor(uo_h, device=dev)
    bn = float(torch.linalg.vector_norm(b))
    u = torch.zeros_like(b)

    def forward():
        u.zero_()
        return h.mgpcg(b, u, abstol=1e-13 * bn, maxiter=200)[1]

    def gradient():
        return h.adjoint_grad(u, u_obs)

    reps = 3
    forward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rf = forward()
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / reps
    gradient()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        L, lam, df, dW, dE, dS, dN, it = gradient()
    torch.cuda.synchronize()
    t_g = (time.perf_counter() - t0) / reps
    grads = (df, dW, dE, dS, dN)
    # the C oracle's gradient, end to end on the host
    t0 = time.perf_counter()
    uh, _, _ = c_oracle.solve_reference(p, N, c_oracle.fourier_rhs(p, N, a.seed))
    Lo, _, *go, it_o = c_oracle.adjoint_grad(p, N, uh, uo_h, rtol=1e-14)
    t_o = time.perf_counter() - t0
    orc_err = max(float(np.abs(g.cpu().numpy() - r).max() / np.abs(r).max()) for g, r in zip(grads, go))
    # central differences along seeded directions
    def loss(fv, gv):
        bb = h.assemble_rhs(fv, *gv)
        uu, _ = h.mgpcg(bb, abstol=1e-14 * float(torch.linalg.vector_norm(bb)), maxiter=200)
        return 0.5 * h.hnorm(uu - u_obs) ** 2

    gen = torch.Generator(device=dev).manual_seed(123)
    fg = (f, gW, gE, gS, gN)
    fd_err = 0.0
    for _ in range(ndir):
        d = [torch.randn(t.numel(), dtype=torch.float64, device=dev, generator=gen) for t in fg]
        d = [di * float(torch.linalg.vector_norm(t)) / float(torch.linalg.vector_norm(di)) for di, t in zip(d, fg)]
        Lp = loss(f + eps * d[0], [g + eps * dg for g, dg in zip(fg[1:], d[1:])])
        Lm = loss(f - eps * d[0], [g - eps * dg for g, dg in zip(fg[1:], d[1:])])
        fd = (Lp - Lm) / (2 * eps)
        ad = sum(float(gr @ di) for gr, di in zip(grads, d))
        fd_err = max(fd_err, abs(fd - ad) / abs(ad))
    h.close()
    return {"workload": f"N={N}, p={p}, L = 1/2 |u - u_obs|_H^2, seeded Fourier f and g; u_obs a seeded Fourier "
                        f"field + N(0, 1e-3)",
            "forward_ms": t_f * 1e3, "forward_iters": rf.iters, "gradient_ms": t_g * 1e3, "adjoint_iters": it,
            "loss": L, "loss_oracle": Lo, "grad_f_norm": float(torch.linalg.vector_norm(df)),
            "grad_g_norm": float(sum(torch.linalg.vector_norm(v) ** 2 for v in (dW, dE, dS, dN)) ** 0.5),
            "oracle_rel_err": orc_err,
            "oracle": f"C oracle FP64 gradient (forward and adjoint MG-PCG to 1e-14, {it_o} adjoint iterations, "
                      f"{t_o:.1f} s on the host)",
            "fd_max_rel_err": fd_err, "fd_directions": ndir, "fd_step": eps}


def tol_sweep(a, h, b, x, bnorm, rel_err, torch):
    """Iterations, relative H-norm error and time of both legs over a grid of
    stopping rules (one JSON line each), to pick the loosest rule per leg
    that meets the error target."""
    from sc_2022_amd.solver import SbpPoisson

    def timeit(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    for rt in (1e-10, 3e-11, 1e-11, 5e-12, 3e-12, 1e-12):
        def f():
            x.zero_()
            return h.mgpcg(b, x, abstol=rt * bnorm, maxiter=200, raise_on_fail=False)[1]
        dt, r = timeit(f)
        print(json.dumps({"leg": "f64", "rtol": rt, "iters": r.iters, "rel_err": rel_err(x), "ms": dt * 1e3}),
              flush=True)
    for irt in (1e-5, 3e-6, 1e-6):
        for rt in (1e-10, 5e-11, 3e-11, 1e-11, 1e-12):
            hm = SbpPoisson(h.N, h.p, device=h.device, inner_rtol=irt, ir_rtol=a.ir_rtol)
            def f():
                x.zero_()
                return hm.mgpcg_mixed(b, x, abstol=rt * bnorm, raise_on_fail=False)[1]
            dt, r = timeit(f)
            print(json.dumps({"leg": "mixed", "inner_rtol": irt, "rtol": rt, "outer": r.iters,
                              "inner": r.inner_iters, "rel_err": rel_err(x), "ms": dt * 1e3}), flush=True)
            hm.close()


def host_cpu():
    """Model name and the CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "cpus_usable": usable, "cpus_online": os.cpu_count()}


def cpu_baseline(N, p, seed, threads, rtol, runs=5, cg_iters=10, b=None):
    """The C oracle's FP64 MG-PCG (same V-cycle, OpenMP) on the host cores:
    one warm-up and `runs` timed solves of the same problem (median), plus
    the reference's own CPU path -- CSR assembly and CG_CPU, one thread --
    over a bounded number of iterations."""
    from oracle import c_oracle
    cores = threads or min(16, os.cpu_count() or 1)
    b = c_oracle.fourier_rhs(p, N, seed) if b is None else b
    res = c_oracle.time_mgpcg_fourier(N, p, seed, rtol, cores, runs=runs, warmup=1, b=b)
    cg = c_oracle.time_cg_csr_sample(N, p, b, cg_iters, threads=1)
    cgm = c_oracle.time_cg_csr_sample(N, p, b, cg_iters, threads=cores)
    c_oracle.set_threads(cores)
    return {"value": 1.0 / res["seconds"], "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"FP64 MG-PCG solves of the same {N}^2 p={p} problem by oracle/sbp_oracle.c "
                      f"({res['iters']} iterations; median of {runs} after 1 warm-up, "
                      f"{res['seconds']:.2f} s each, OpenMP {cores} threads)",
            "all_seconds": [round(t, 4) for t in res["all_seconds"]],
            "host": host_cpu(),
            "csr_cg_one_thread": {
                "what": f"CSR assembly + {cg['iters']} CG iterations (CG_CPU's loop) at {N}^2 p={p}, 1 thread",
                "assembly_s": cg["assembly_s"], "ms_per_iter": cg["s_per_iter"] * 1e3, "nnz": cg["nnz"]},
            "csr_cg_all_cores": {"threads": cores, "assembly_s": cgm["assembly_s"],
                                 "ms_per_iter": cgm["s_per_iter"] * 1e3}}


def traffic_from_profiles(dtype_tag):
    path = os.path.join(ROOT, "profiles", "apply_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d.get(dtype_tag)
    except (OSError, ValueError):
        return None


def solve_rates(h, N, rmx, r64, tmx, t64):
    """Solve-level algorithmic GB/s: the FP64 solve's bytes (mgpcg_bytes at
    8 B), and the mixed solve's FP32 inner solves (mgpcg_bytes at 4 B, one
    start per correction) plus per correction the FP64 residual cast to FP32
    (x, b in, r out: 20 B a point) and the FP64 update (x, d in, x out: 20 B),
    and the final residual check (20 B)."""
    sides = [h.level_size(l) for l in range(h.nlevels)]
    n = N * N
    b64 = mgpcg_bytes(sides, 8, r64.iters)
    outer = rmx.iters
    inner = rmx.inner_iters or 0
    b32 = (mgpcg_bytes(sides, 4, 0, residual=False) * outer + inner * (9 * n * 4 + vcycle_bytes(sides, 4))
           + outer * 40 * n + 20 * n)
    return {"f64": b64 / t64 / 1e9, "mixed": b32 / tmx / 1e9, "f64_bytes": b64, "mixed_bytes": b32,
            "f64_bytes_per_iter": 9 * n * 8 + vcycle_bytes(sides, 8),
            "f32_bytes_per_iter": 9 * n * 4 + vcycle_bytes(sides, 4)}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}), flush=True)
        sys.exit(2)
    import torch
    from sc_2022_amd.solver import SbpPoisson

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if a.batch_sweep:
        batch_leg(a, dev, torch)
        return
    N, p = a.n, a.p
    n = N * N
    h = SbpPoisson(N, p, device=dev, ir_rtol=a.ir_rtol, inner_rtol=a.inner_rtol)

    # ---- apply roofline (FP64 and FP32), rotating > 1 GB of buffers ----
    nbuf64 = max(2, int((2 * 1024 ** 3) // (2 * n * 8)) + 1)
    nbuf32 = max(2, int((2 * 1024 ** 3) // (2 * n * 4)) + 1)
    apply_ms64 = time_apply(h, torch, torch.float64, a.apply_reps, nbuf64)
    apply_ms32 = time_apply(h, torch, torch.float32, a.apply_reps, nbuf32)
    bytes64 = 2 * n * 8
    bytes32 = 2 * n * 4
    gbs64 = bytes64 / (apply_ms64 * 1e-3) / 1e9
    gbs32 = bytes32 / (apply_ms32 * 1e-3) / 1e9
    copy_ms = time_copy(h, torch, 100, n * 8, nbuf64)
    copy_gbs = 2 * n * 8 / (copy_ms * 1e-3) / 1e9
    if a.only_apply:
        print(json.dumps({"apply_f64_us": apply_ms64 * 1e3, "apply_f64_GBs": gbs64,
                          "apply_f32_us": apply_ms32 * 1e3, "apply_f32_GBs": gbs32,
                          "copy_GBs": copy_gbs}), flush=True)
        return

    # ---- problem and the FP64 oracle's reference solution (untimed) ----
    b = fourier_problem(h, a.seed)
    bnorm = float(torch.linalg.vector_norm(b))
    ref_b_host = []
    ref = oracle_reference(h, b, p, a.seed, torch, keep_b=ref_b_host)
    xref, href = ref.pop("x"), ref.pop("hnorm")

    def rel_err(x):
        return h.hnorm(x - xref) / href

    # the GPU's own tightly converged FP64 solve, against the oracle's
    xg, rg = h.mgpcg(b, abstol=1e-14 * bnorm, maxiter=200, raise_on_fail=False)
    ref["gpu_f64_1e-14"] = {"status": rg.status, "rel_resid": rg.resnorm / bnorm, "rel_err_vs_oracle": rel_err(xg)}
    del xg

    x = torch.zeros_like(b)
    if a.tol_sweep:
        tol_sweep(a, h, b, x, bnorm, rel_err, torch)
        return

    def solve_mixed():
        x.zero_()
        return h.mgpcg_mixed(b, x, abstol=a.mixed_rtol * bnorm)[1]

    def solve_f64():
        x.zero_()
        return h.mgpcg(b, x, abstol=a.rtol * bnorm, maxiter=200)[1]

    # ---- FP64 MG-PCG (reported beside the headline) ----
    for _ in range(max(1, a.warmup)):
        r64 = solve_f64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        r64 = solve_f64()
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / a.steps
    err64 = rel_err(x)

    # ---- headline: mixed-precision solves ----
    for _ in range(a.warmup):
        rmx = solve_mixed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rmx = solve_mixed()
    torch.cuda.synchronize()
    tmx = (time.perf_counter() - t0) / a.steps
    errmx = rel_err(x)

    extra = {}
    if not a.no_extra:
        for name, fn in (("csr_spmv", lambda: csr_leg(h, torch, apply_ms64 * 1e3)),
                         ("n8193", lambda: leg_8193(a, dev, torch)),
                         ("initial_guess_cg", lambda: nested_leg(a, dev, torch)),
                         ("reference_algorithms", lambda: reference_modes_leg(a, dev, torch)),
                         ("variable_coefficient", lambda: variable_coefficient_leg(a, dev, torch)),
                         ("adjoint", lambda: adjoint_leg(a, dev, torch))):
            if a.extra and name not in a.extra.split(","):
                continue
            t_leg = time.perf_counter()
            try:
                extra[name] = fn()
            except Exception as e:  # reported, never silently replaced
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"# {name}: {time.perf_counter() - t_leg:.1f} s", file=sys.stderr, flush=True)

    batched = None
    if a.batch > 0:
        try:
            batched = batch_leg(a, dev, torch)
        except Exception as e:  # reported, never silently replaced
            batched = {"error": f"{type(e).__name__}: {e}"}

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(N, p, a.seed, a.cpu_threads, a.rtol, b=ref_b_host[0])
        except Exception as e:  # reported, never silently replaced
            cpu = {"error": f"{type(e).__name__}: {e}"}

    out = {
        "metric": "mixed-precision MG-PCG solves/s at 4097^2 SBP4 (rel err <= 1e-10 vs FP64), + apply HBM GB/s",
        "value": 1.0 / tmx,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": tmx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 outer",
        "data": "synthetic: seeded Fourier-mode source (64 modes) and boundary data (8 modes/face)",
        "config": {"workload": f"MG-PCG solve, N={N} (n={n}), SBP p={p}, V-cycle 2+2 damped Jacobi to 17^2; "
                               f"each leg stops at the loosest rule meeting rel err <= {REL_ERR_TARGET:g} "
                               f"(bench.py --tol-sweep): FP64 ||r|| < {a.rtol:g} ||b||, mixed FP64 ||r|| < "
                               f"{a.mixed_rtol:g} ||b|| with FP32 inner solves to {a.inner_rtol:g}"
                               + (f" and error estimate < {a.ir_rtol:g} ||x||" if a.ir_rtol > 0 else ""),
                   "seed": a.seed},
        "rel_err_mixed": errmx,
        "rel_err_f64": err64,
        "rel_err_reference": ref,
        "meets_rel_err_target": bool(errmx <= REL_ERR_TARGET and err64 <= REL_ERR_TARGET),
        "mixed": {"solves_per_s": 1.0 / tmx, "outer": rmx.iters, "inner": rmx.inner_iters,
                  "ms": tmx * 1e3},
        "f64": {"solves_per_s": 1.0 / t64, "iters": r64.iters, "ms": t64 * 1e3},
        "speedup_mixed_vs_f64": t64 / tmx,
        "solve_GBs": solve_rates(h, N, rmx, r64, tmx, t64),
        "apply": {"f64_us": apply_ms64 * 1e3, "f64_GBs": gbs64, "f32_us": apply_ms32 * 1e3, "f32_GBs": gbs32,
                  "copy_GBs": copy_gbs, "rotating_buffers": nbuf64},
        "roofline": {"bound": "hbm", "achieved": gbs64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs64 / HBM_PEAK_GBS, "traffic": traffic_from_profiles("f64"),
                     "traffic_source": "profiles/apply_traffic.json: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this "
                                       "kernel at this size (a profiled run, not this run), scaled by a plain copy; "
                                       "also profiles/solve_counters_4097_p4.md",
                     "kernel": "stencil_pass FP64 apply (OutStore), algorithmic bytes 2*n*8 per launch"},
        "batched": batched,
        **extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
