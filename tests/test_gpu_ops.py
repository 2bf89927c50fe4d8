"""GPU parity of the operator-level entry points against the CPU oracle:
apply (both signs, FP64 and FP32), diag, right-hand side and its transpose,
H-norm, restriction, prolongation, smoothing, coarse solve, V-cycle, CSR
assembly and SpMV.  FP64 bar: 1e-13 relative (max-norm) unless stated."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from oracle import sbp_oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def T(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def A_of(level, p, _cache={}):
    key = (level, p)
    if key not in _cache:
        _cache[key] = O.assemble_operator(level, p)
    return _cache[key]


def relmax(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def SP():
    from sc_2022_amd.solver import SbpPoisson
    return SbpPoisson


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("level", [4, 5, 6, 7])
def test_apply_f64_matches_assembled(SP, level, p):
    N = 2 ** level + 1
    h = SP(N, p)
    A = A_of(level, p)
    rng = np.random.default_rng(level * 10 + p)
    u = rng.standard_normal(N * N)
    out = h.apply(T(u)).cpu().numpy()
    assert relmax(out, A @ u) < 1e-13
    neg = h.apply(T(u), sign=-1).cpu().numpy()
    assert relmax(neg, -(A @ u)) < 1e-13


@pytest.mark.parametrize("level", [4, 5])
def test_apply_sign_minus_is_reference_kernel(SP, level):
    """sign = -1 reproduces the literal restatement of matrix_free_A_full_GPU."""
    N = 2 ** level + 1
    h = SP(N, 2)
    u = np.random.default_rng(0).standard_normal(N * N)
    ref = O.reference_gpu_apply_p2(u, N)
    assert relmax(h.apply(T(u), sign=-1).cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize("p", [2, 4])
def test_apply_golden(SP, golden_dir, p):
    d = np.load(os.path.join(golden_dir, f"sbp_L5_p{p}.npz"))
    h = SP(int(d["N"]), p)
    assert relmax(h.apply(T(d["u"])).cpu().numpy(), d["Au"]) < 1e-13
    assert relmax(h.apply(T(d["u"]), sign=-1).cpu().numpy(), d["ref_minus_Au"]) < 1e-13
    assert relmax(h.diag().cpu().numpy(), d["diag"]) < 1e-15


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("tiles", [(64, 4, 0), (64, 8, 0), (128, 8, 0), (256, 4, 0), (256, 8, 0), (256, 16, 0),
                                   (256, 8, 1), (256, 16, 1), (512, 8, 1)])
def test_apply_every_tile_shape(SP, p, tiles):
    level = 7
    N = 2 ** level + 1
    h = SP(N, p)
    h.set_tiles(0, tiles[0], tiles[1], aligned=bool(tiles[2]))
    u = np.random.default_rng(11).standard_normal(N * N)
    assert relmax(h.apply(T(u)).cpu().numpy(), A_of(level, p) @ u) < 1e-13
    uf = u.astype(np.float32)
    e32 = relmax(h.apply(T(uf, torch.float32)).cpu().double().numpy(), A_of(level, p) @ uf.astype(np.float64))
    assert e32 < 2e-5


@pytest.mark.parametrize("p", [2, 4])
@pytest.mark.parametrize("N", [12, 13, 20, 37, 100, 300])
def test_apply_ragged_sizes_against_closed_form(SP, N, p):
    """Sides that are not 2^k+1 (the apply accepts any N >= 2C)."""
    if p == 4 and N < 12:
        pytest.skip("below the p=4 minimum")
    hw, m, k = O.one_d_factors(p, N)
    A = sp.kron(sp.csr_matrix(k), sp.diags(hw)) + sp.kron(sp.diags(hw), sp.csr_matrix(m))
    h = SP(N, p)
    u = np.random.default_rng(N).standard_normal(N * N)
    assert relmax(h.apply(T(u)).cpu().numpy(), A @ u) < 1e-13


def test_apply_seeded_random_cases(SP):
    """30 seeded random cases -- order 2..10, side 2C..400 (mostly not
    2^k + 1), a compiled tile shape for the order, sign +-1, FP64 and FP32 --
    against the closed form kron(k, hw) + kron(hw, m)."""
    from sc_2022_amd.autotune import compiled_shapes
    rng = np.random.default_rng(20261203)
    cmin = {2: 8, 4: 12, 6: 18, 8: 24, 10: 32}
    for case in range(30):
        p = int(rng.choice([2, 4, 6, 8, 10]))
        N = int(rng.integers(cmin[p], 401))
        shapes = compiled_shapes(p)
        ti, tj, ta = shapes[int(rng.integers(len(shapes)))]
        sign = int(rng.choice([1, -1]))
        hw, m, k = O.one_d_factors(p, N)
        A = sp.kron(sp.csr_matrix(k), sp.diags(hw)) + sp.kron(sp.diags(hw), sp.csr_matrix(m))
        h = SP(N, p, tile_i=ti, tile_j=tj) if not ta else SP(N, p)
        if ta:
            h.set_tiles(0, ti, tj, aligned=True)
        u = rng.standard_normal(N * N)
        ref = sign * (A @ u)
        e64 = relmax(h.apply(T(u), sign=sign).cpu().numpy(), ref)
        uf = u.astype(np.float32)
        e32 = relmax(h.apply(T(uf, torch.float32), sign=sign).cpu().double().numpy(),
                     sign * (A @ uf.astype(np.float64)))
        assert e64 < 1e-13 and e32 < 2e-5, (case, p, N, ti, tj, ta, sign, e64, e32)
        h.close()


@pytest.mark.parametrize("p", [2, 4])
def test_apply_f32(SP, p):
    level = 6
    N = 2 ** level + 1
    h = SP(N, p)
    u = np.random.default_rng(2).standard_normal(N * N)
    out = h.apply(T(u, torch.float32)).cpu().double().numpy()
    assert relmax(out, A_of(level, p) @ u.astype(np.float32).astype(np.float64)) < 5e-6


@pytest.mark.parametrize("p", [2, 4])
def test_apply_large_properties(SP, p):
    """4097^2 (the headline size): symmetry <Au, v> = <u, Av>, the closed
    form of A 1, and agreement with a 1-D-factor reconstruction on a slab."""
    N = 4097
    h = SP(N, p)
    g = torch.Generator(device=DEV).manual_seed(5)
    u = torch.randn(N * N, dtype=torch.float64, device=DEV, generator=g)
    v = torch.randn(N * N, dtype=torch.float64, device=DEV, generator=g)
    Au, Av = h.apply(u), h.apply(v)
    a, b = float(Au @ v), float(u @ Av)
    assert abs(a - b) < 1e-13 * float(Au.norm()) * float(v.norm())
    one = torch.ones(N * N, dtype=torch.float64, device=DEV)
    A1 = h.apply(one).reshape(N, N).cpu().numpy()      # A1[j, i]
    hw, m, k = O.one_d_factors(p, 64)
    k1 = k.sum(axis=1)                                   # k 1 near the face
    hwN = np.ones(N); nb = 1 if p == 2 else 4
    hwN[:nb] = hw[:nb]; hwN[-nb:] = hw[-nb:]
    kN = np.zeros(N); kN[:8] = k1[:8]; kN[-8:] = k1[-8:]
    expect = kN[:, None] * hwN[None, :]                 # (A 1)(i, j) = hw_i (k 1)_j
    assert np.abs(A1 - expect).max() < 1e-12
    # slab check: the first 16 columns j against a dense 1-D reconstruction
    U = u.reshape(N, N).cpu().numpy()
    rows = slice(0, 16)
    # k-part: rows j 0..15 need columns j' 0..21
    Ah = Au.reshape(N, N).cpu().numpy()
    Mi = sp.diags([np.full(N - abs(s), c) for s, c in _interior(p)], [s for s, _ in _interior(p)]).tolil()
    _patch_faces(Mi, m, N)
    Ki = sp.diags([np.full(N - abs(s), c) for s, c in _interior(p)], [s for s, _ in _interior(p)]).tolil()
    _patch_faces(Ki, k, N)
    Mi, Ki = Mi.tocsr(), Ki.tocsr()
    ref = (Ki[rows] @ U) * hwN[None, :] + hwN[rows][:, None] * (U[rows] @ Mi.T)
    assert relmax(Ah[rows], ref) < 1e-13


def _interior(p):
    if p == 2:
        return [(-1, -1.0), (0, 2.0), (1, -1.0)]
    return [(-2, 1 / 12), (-1, -4 / 3), (0, 2.5), (1, -4 / 3), (2, 1 / 12)]


def _patch_faces(M, small, N):
    """Copy the closure rows of a small dense 1-D factor into the big one."""
    n0 = small.shape[0]
    R = 4
    for r in range(R + 2):
        for c in range(n0):
            M[r, c] = small[r, c]
            M[N - 1 - r, N - 1 - c] = small[n0 - 1 - r, n0 - 1 - c]


@pytest.mark.parametrize("p", [2, 4])
def test_rhs_and_transpose(SP, golden_dir, p):
    d = np.load(os.path.join(golden_dir, f"sbp_L5_p{p}.npz"))
    h = SP(int(d["N"]), p)
    b = h.assemble_rhs(T(d["f"]), T(d["gW"]), T(d["gE"]), T(d["gS"]), T(d["gN"])).cpu().numpy()
    assert relmax(b, d["b"]) < 1e-13
    bf = h.assemble_rhs(T(d["ff"]), T(d["fW"]), T(d["fE"]), T(d["fS"]), T(d["fN"])).cpu().numpy()
    assert relmax(bf, d["bf"]) < 1e-13
    outs = h.rhs_transpose(T(d["lam"]))
    for got, key in zip(outs, ("df", "dgW", "dgE", "dgS", "dgN")):
        assert relmax(got.cpu().numpy(), d[key]) < 1e-13, key
    hn = h.hnorm(T(d["u"]))
    assert abs(hn - O.h_norm(d["u"], d["hdiag"])) < 1e-13 * hn


@pytest.mark.parametrize("p", [2, 4])
def test_restrict_prolong(SP, golden_dir, p):
    d = np.load(os.path.join(golden_dir, f"sbp_L5_p{p}.npz"))
    N = int(d["N"])
    h1 = SP(N, p, restrict_scale=1.0)
    rc = h1.restrict(T(d["r"])).cpu().numpy()
    assert relmax(rc, d["R_fw_r"]) < 1e-15
    h4 = SP(N, p)
    assert relmax(h4.restrict(T(d["r"])).cpu().numpy(), 4 * d["R_fw_r"]) < 1e-15
    xf = T(np.zeros(N * N))
    h4.prolong_add(T(d["ec"]), xf)
    assert relmax(xf.cpu().numpy(), d["P_e"]) < 1e-15
    # FP32
    rc32 = h1.restrict(T(d["r"], torch.float32)).cpu().double().numpy()
    assert relmax(rc32, d["R_fw_r"]) < 1e-6


@pytest.mark.parametrize("p", [2, 4])
def test_smoother_matches_oracle(SP, p):
    level = 6
    N = 2 ** level + 1
    A = A_of(level, p)
    rng = np.random.default_rng(4)
    b, x0 = rng.standard_normal(N * N), rng.standard_normal(N * N)
    h = SP(N, p)
    x = T(x0)
    h.smooth(T(b), x, 3)
    ref = O.jacobi_sweeps(A, x0, b, 3)
    assert relmax(x.cpu().numpy(), ref) < 1e-13
    hr = SP(N, p, smoother=1, richardson_omega=0.15)
    x = T(x0)
    hr.smooth(T(b), x, 3)
    assert relmax(x.cpu().numpy(), O.modified_richardson(A, x0, b, 3, 0.15)) < 1e-13


@pytest.mark.parametrize("bottom", [0, 65])
@pytest.mark.parametrize("p", [2, 4])
def test_coarse_solve_and_vcycle(SP, p, bottom):
    """65^2: with bottom_n = 65 the levels 65 and 33 run in the
    single-workgroup bottom kernels around the 17^2 coarse solve; with 0 they
    run the fused two-sweep kernels."""
    level = 6
    N = 2 ** level + 1
    h = SP(N, p, bottom_n=bottom)
    assert h.nlevels == 3
    Ac = A_of(4, p)
    bc = np.random.default_rng(8).standard_normal(17 * 17)
    xc = h.coarse_solve(T(bc)).cpu().numpy()
    assert relmax(xc, spla.spsolve(Ac.tocsc(), bc)) < 1e-12
    hier = O.Hierarchy(level, p)
    r = np.random.default_rng(9).standard_normal(N * N)
    z = h.vcycle(T(r)).cpu().numpy()
    assert relmax(z, O.vcycle(hier, r)) < 1e-12
    z32 = h.vcycle(T(r, torch.float32)).cpu().double().numpy()
    assert relmax(z32, O.vcycle(hier, r)) < 1e-4


@pytest.mark.parametrize("bottom", [0, 65])
@pytest.mark.parametrize("nu", [(1, 1), (3, 0), (0, 3), (2, 2)])
def test_vcycle_sweep_counts(SP, nu, bottom):
    level, p = 5, 4
    N = 2 ** level + 1
    h = SP(N, p, nu_pre=nu[0], nu_post=nu[1], bottom_n=bottom)
    hier = O.Hierarchy(level, p)
    r = np.random.default_rng(10).standard_normal(N * N)
    z = h.vcycle(T(r)).cpu().numpy()
    assert relmax(z, O.vcycle(hier, r, nu[0], nu[1])) < 1e-12


@pytest.mark.parametrize("level", [6, 9])
@pytest.mark.parametrize("p", [2, 4])
def test_bottom_kernels_match_fused_levels(SP, p, level):
    """The bottom kernels (the levels of side <= 65 in one workgroup per
    problem, down and up leg around the coarse solve) evaluate every point as
    the fused two-sweep kernels do: V-cycles with bottom_n = 65 and 0 agree to
    rounding, FP64 and FP32, also over a stacked batch of three problems."""
    N = 2 ** level + 1
    hf = SP(N, p, bottom_n=0)
    hb = SP(N, p, bottom_n=65)
    g = torch.Generator(device=DEV).manual_seed(level * 10 + p)
    r = torch.randn(N * N, dtype=torch.float64, device=DEV, generator=g)
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        zf = hf.vcycle(r.to(dt))
        zb = hb.vcycle(r.to(dt))
        assert float((zb - zf).abs().max()) <= tol * float(zf.abs().max())
    hf.close(); hb.close()
    from sc_2022_amd.batch import SbpBatch
    bb = SbpBatch(N, p, 3, device=DEV, bottom_n=65)
    rs = torch.randn(3 * N * N, dtype=torch.float64, device=DEV, generator=g)
    hf = SP(N, p, bottom_n=0)
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        zs = bb.vcycle(rs.to(dt))
        for q in range(3):
            zf = hf.vcycle(rs[q * N * N:(q + 1) * N * N].to(dt))
            assert float((zs[q * N * N:(q + 1) * N * N] - zf).abs().max()) <= tol * float(zf.abs().max())
    bb.close(); hf.close()


@pytest.mark.parametrize("p", [2, 4])
def test_csr_assembly_and_spmv(SP, p):
    level = 5
    N = 2 ** level + 1
    h = SP(N, p)
    rp, col, val = h.csr_assemble()
    A = sp.csr_matrix((val.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(N * N, N * N))
    Ao = A_of(level, p)
    assert abs(A - Ao).max() < 1e-13 * abs(Ao).max()
    assert A.nnz == Ao.nnz
    u = np.random.default_rng(12).standard_normal(N * N)
    y = h.csr_spmv((rp, col, val), T(u)).cpu().numpy()
    assert relmax(y, Ao @ u) < 1e-13


def test_copy_kernel(SP):
    h = SP(17, 2)
    a = torch.randn(1 << 20, dtype=torch.float64, device=DEV)
    b = torch.empty_like(a)
    h.copy(a, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("bottom", [0, 33, 65])
@pytest.mark.parametrize("p", [2, 4])
def test_vcycle_large_against_oracle(SP, p, bottom):
    """Level 9 (513^2): every V-cycle kernel runs its interior-tile path
    (including the fused residual + restriction); the levels at or below
    bottom_n run in the single-workgroup bottom kernels."""
    level = 9
    N = 2 ** level + 1
    h = SP(N, p, bottom_n=bottom)
    hier = O.Hierarchy(level, p)
    r = np.random.default_rng(21).standard_normal(N * N)
    z = h.vcycle(T(r)).cpu().numpy()
    assert relmax(z, O.vcycle(hier, r)) < 1e-12
    z32 = h.vcycle(T(r, torch.float32)).cpu().double().numpy()
    assert relmax(z32, O.vcycle(hier, r)) < 1e-4


@pytest.mark.parametrize("p,level,shape", [(p, lv, sh) for p in (2, 4) for lv in (5, 7, 10)
                                            for sh in ((64, 4), (64, 8), (128, 8), (256, 8))]
                         + [(6, lv, sh) for lv in (7, 10) for sh in ((64, 4), (64, 8), (128, 8))])
def test_vcycle_fused_passes_match_separate_passes(SP, p, level, shape):
    """The temporally blocked two-sweep kernels (pre-smoothing + residual +
    restriction; prolongation + post-smoothing) against the pass-per-sweep
    V-cycle and the oracle, for every compiled block shape: level 10 (1025^2)
    has interior blocks, shifted last blocks and face blocks on every level
    down to 33^2; with 4-row blocks the second block row holds closure rows
    whose reads leave the block's tiles (recomputed from global memory)."""
    N = 2 ** level + 1
    r = np.random.default_rng(30 + level).standard_normal(N * N)
    hf = SP(N, p, fuse_n=1 << 20, bottom_n=0)
    for lv in range(hf.nlevels - 1):
        hf.set_tiles(lv, shape[0], shape[1], fused=True)
    hs = SP(N, p, fuse_n=0, bottom_n=0)
    zf = hf.vcycle(T(r)).cpu().numpy()
    zs = hs.vcycle(T(r)).cpu().numpy()
    assert relmax(zf, zs) < 1e-13
    if level <= 7:
        assert relmax(zf, O.vcycle(O.Hierarchy(level, p, coarse_level=5 if p == 6 else 4), r)) < 1e-12
    zf32 = hf.vcycle(T(r, torch.float32)).cpu().double().numpy()
    zs32 = hs.vcycle(T(r, torch.float32)).cpu().double().numpy()
    assert relmax(zf32, zs32) < 1e-5
    # Richardson smoother through the same kernels
    hr = SP(N, p, fuse_n=1 << 20, smoother=1, richardson_omega=0.15, bottom_n=0)
    for lv in range(hr.nlevels - 1):
        hr.set_tiles(lv, shape[0], shape[1], fused=True)
    hr0 = SP(N, p, fuse_n=0, smoother=1, richardson_omega=0.15, bottom_n=0)
    assert relmax(hr.vcycle(T(r)).cpu().numpy(), hr0.vcycle(T(r)).cpu().numpy()) < 1e-13
