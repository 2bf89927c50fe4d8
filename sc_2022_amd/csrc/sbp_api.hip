// sbp_api.hip -- C ABI (include/sbp.h) over the gfx950 kernels in sbp_kernels.h.
//
// The handle owns everything the reference re-allocated per call (the MG
// hierarchy buffers of initial_guess_cg.jl:152-161, the apply temporaries of
// split_matrix_free.jl:459-462, the coarse LU of initial_guess_cg.jl:521),
// allocated once at creation.  A handle holds nprob >= 1 independent problems
// of one size and order stacked in its buffers (problem q at offset q * n of
// every grid function); every launch covers all of them.  Krylov scalars live
// on the device, one block per problem; kernels test that problem's "done"
// flag and fall through once it has converged, so the host only polls now
// and then (every poll_every iterations, or where the measured convergence
// rate predicts convergence) while each problem's iteration count keeps the
// reference's semantics (the first k with ||r_k||_2 < abstol).
#include <hip/hip_runtime.h>

#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/sbp.h"
#include "sbp_bottom.h"
#include "sbp_var.h"
#include "sbp_kernels.h"
#include "sbp_launch.h"
#include "sbp_tables.h"

using namespace sbp;

namespace {

constexpr int kMaxLevels = 16;
constexpr int kEwThreads = 256;

struct Level {
    int N = 0;
    size_t n = 0;                         // points of one problem
    OpTab tb{};
    OpTab* dtab = nullptr;                // device copy of tb (tb.dev points here)
    int ti[2] = {0, 0}, tj[2] = {0, 0};   // tile shape per precision (0 = f64, 1 = f32)
    int ta[2] = {0, 0};                   // 1: the shape's line-aligned variant
    int fti[2] = {128, 128}, ftj[2] = {8, 8};   // fused two-sweep block shape per precision
    // per precision (0 = f64, 1 = f32), nprob problems each: rhs, two sweep
    // buffers, output, residual
    void* b[2] = {nullptr, nullptr};
    void* xa[2] = {nullptr, nullptr};
    void* xb[2] = {nullptr, nullptr};
    void* xo[2] = {nullptr, nullptr};
    void* r[2] = {nullptr, nullptr};
};

}  // namespace

struct sbp_handle_s {
    int N = 0, p = 0, nlev = 0, nprob = 1;
    bool mg = false;
    // Problems present on each level: all nprob for a stacked handle; for a
    // ragged handle (sbp_create_ragged: problems of several sides, sorted by
    // side, largest first) the problems whose side is at least the level's.
    // Problem rank q sits in slot q of every level's stacks.  j0[l] is the
    // first rank whose finest level is l (np[l] - j0[l] of them join there);
    // koff[l] is the element offset of rank j0[l] in the ragged layout of the
    // caller's arrays (problem q after all problems of rank < q, N_q^2 each).
    bool ragged = false;
    int np[kMaxLevels] = {};
    int j0[kMaxLevels] = {};
    size_t koff[kMaxLevels] = {};
    size_t rtot = 0;             // ragged: total points over the problems
    size_t roff[kMaxLevels] = {};   // ragged: partial-slot offset of level l (np[l] * npl[l] slots)
    size_t npl[kMaxLevels] = {};    // ragged: partial slots per problem on level l
    int bottom_l0 = -1;          // first level run by the single-workgroup bottom kernel (-1: none)
    sbp_params prm{};
    hipStream_t st = nullptr;
    RhsTab rt{};
    Level L[kMaxLevels];
    double* Ainv = nullptr;
    int nc = 0;
    // fine Krylov vectors (nprob problems each)
    double *r = nullptr, *pv = nullptr, *pv2 = nullptr, *Ap = nullptr, *z = nullptr, *x = nullptr;
    float *r32 = nullptr, *p32 = nullptr, *p32b = nullptr, *Ap32 = nullptr, *z32 = nullptr, *d32 = nullptr;
    double* partial = nullptr;   // block partials, nprob * npp slots
    size_t npp = 0;
    Scalars* sc = nullptr;       // outer / FP64 solver, nprob blocks
    Scalars* sc_in = nullptr;    // inner FP32 solver, nprob blocks
    Scalars* sc_host = nullptr;  // pinned mirror, nprob blocks
    double* dabs = nullptr;      // device per-problem tolerances
    double* habs = nullptr;      // pinned staging for them
    double* dtmp = nullptr;      // small device scalars (nprob)
    double* w = nullptr;         // adjoint source (allocated on first use)
    double* lam = nullptr;       // adjoint state when the caller passes none
    bool have32 = false;
    double clock_khz = 0.0;      // wall_clock64 rate
    // Captured Krylov iterations: [precision][0 = CG, 1 = MG-PCG]
    hipStream_t cap = nullptr;   // capture stream (default priority)
    hipEvent_t pev = nullptr;    // end of a poll's scalar copy (poll_start)
    // [precision][0 = CG, 1 = MG-PCG][parity of the iteration]
    hipGraphExec_t gexec[2][2][2] = {};
    hipGraph_t graph[2][2][2] = {};
    // Initial-guess interpolation: the handle of the next coarser grid
    // ((N+1)/2, same order and parameters) and the coarse solution on it,
    // both made on first use.
    sbp_handle child = nullptr;
    double* xc = nullptr;
    // Coarsest grids above 33 x 33 (coarse_n > 33, e.g. the reference's
    // two-level cycle): solved "exactly" by a nested FP64 MG-PCG on a handle
    // of that size (default parameters), made on first use; cbuf stages
    // FP32 right-hand sides and solutions as FP64.
    bool coarse_nested = false;
    sbp_handle cexact = nullptr;
    double* cbuf = nullptr;
    // Variable-coefficient operator (sbp_create_var): device table and bands.
    bool var = false;
    VarTab* vtab = nullptr;
    double* vband = nullptr;
    // Its multigrid hierarchy (N = 2^k + 1): per level the device table and
    // bands (level 0: vtab / vband) and the level's right-hand side b,
    // iterate x and sweep buffer t (level 0: b is the caller's, x and t are
    // z / pv2).  The coarsest level is solved with Ainv (nc x nc).
    struct VarLevel {
        int N = 0;
        size_t n = 0;
        VarTab* dtab = nullptr;
        double* band = nullptr;
        double *b = nullptr, *x = nullptr, *t = nullptr;
        float *b32 = nullptr, *x32 = nullptr, *t32 = nullptr;   // FP32 V-cycle (level 0: r32 / z32 / Ap32)
    };
    std::vector<VarLevel> vl;
    bool vmg = false;
    int dev = 0;                 // device current at creation
    // The CSR operator through rocSPARSE (sbp_csr_vendor_*): library context,
    // matrix and vector descriptors, the SpMV work buffer from its analysis.
    rocsparse_handle vs = nullptr;
    rocsparse_spmat_descr vs_mat = nullptr;
    rocsparse_dnvec_descr vs_x = nullptr, vs_y = nullptr;
    rocsparse_spmv_alg vs_alg = rocsparse_spmv_alg_csr_adaptive;
    void* vs_buf = nullptr;
    size_t vs_bufsz = 0;
};

namespace {

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return SBP_ERR_HIP;   \
    } while (0)

#define STCHK(x)                                    \
    do {                                            \
        int s_ = (x);                               \
        if (s_ != SBP_OK) return s_;                \
    } while (0)

inline int ew_blocks(size_t n) {
    size_t b = (n + kEwThreads - 1) / kEwThreads;
    return (int)std::max<size_t>(1, std::min<size_t>(b, 4096));
}

// Tile shapes that are compiled in.
inline bool tile_ok(int p, int ti, int tj, int ta = 0) {
#define SBP_TILE_OK(TI_, J_, A_) if (ti == TI_ && tj == J_ && ta == A_) return true;
    if (p > 4) {
        SBP_HI_TILE_SHAPES(SBP_TILE_OK)
        return false;
    }
    SBP_TILE_SHAPES(SBP_TILE_OK)
#undef SBP_TILE_OK
    return false;
}

// Default stencil tile of a level of side N in precision q (0 = FP64, 1 = FP32)
// of an order-p handle; the autotuner's table (tiles_gfx950.json) refines it.
inline void default_tile(int N, int p, int q, int* ti, int* tj, int* ta) {
    *ta = 0;
    if (p <= 4 && N >= 4097 && (q == 0 || N < 8193)) { *ti = 256; *tj = 8; *ta = 1; }   // line-aligned apply
    else if (N >= 1025) { *ti = 256; *tj = 8; }
    else if (N >= 129) { *ti = 128; *tj = 8; }
    else { *ti = 64; *tj = 4; }
}

inline size_t pass_blocks(int N, int ti, int tj, int aln = 0) { return (size_t)nparts_of(N, ti, tj, aln); }

// The tile shape a pass with output policy OUT runs on level L in precision
// T: the level's shape, except that policies that do not take line-aligned
// stores (OUT::kAlign false) run an aligned level on the unaligned shape of
// the same size (256 x 8 where that size has none).
struct Shape { int ti, tj, ta; };
template <typename T, class OUT> inline Shape pass_shape(int p, const Level& L);

// Blocks (and block partials per problem) of that pass.
template <typename T, class OUT> inline size_t pass_nblocks(int p, const Level& L);

template <typename T> inline int prec_index();
template <> inline int prec_index<double>() { return 0; }
template <> inline int prec_index<float>() { return 1; }

template <typename T, class OUT> inline Shape pass_shape(int p, const Level& L) {
    const int q = prec_index<T>();
    Shape s{L.ti[q], L.tj[q], L.ta[q]};
    if (s.ta && !OUT::kAlign) {
        s.ta = 0;
        if (!tile_ok(p, s.ti, s.tj, 0)) { s.ti = 256; s.tj = 8; }
    }
    return s;
}

template <typename T, class OUT> inline size_t pass_nblocks(int p, const Level& L) {
    const Shape s = pass_shape<T, OUT>(p, L);
    return pass_blocks(L.N, s.ti, s.tj, aln_of(s.ta, (int)sizeof(T)));
}



inline Batch batch_of(const sbp_handle h, int l) {
    Batch bt;
    bt.nprob = h->np[l];
    bt.sf = (long long)h->L[l].n;
    bt.sc = l + 1 < h->nlev ? (long long)h->L[l + 1].n : 0;
    bt.ss = kScalarDoubles;
    return bt;
}

inline dim3 ew_grid(const sbp_handle h, size_t n) { return dim3(ew_blocks(n), h->nprob); }
// The same over the problems present on level l.
inline dim3 ew_grid_l(const sbp_handle h, int l) { return dim3(ew_blocks(h->L[l].n), h->np[l]); }

// One stencil pass on level l over every problem; *nparts receives the
// number of block partials per problem.
template <typename T, class IN, class OUT>
int launch_pass(sbp_handle h, int l, const IN& in, const OUT& out, const Scalars* scb, double* partial,
                int* nparts, int nprob = -1) {
    const Level& L = h->L[l];
    const Shape sh = pass_shape<T, OUT>(h->p, L);
    const int ti = sh.ti, tj = sh.tj, ta = sh.ta;
    if (nparts) *nparts = (int)pass_nblocks<T, OUT>(h->p, L);
    PassCtl c;
    c.done = scb ? &scb->done : nullptr;
    c.ds = kScalarInts;
    c.partial = partial;
    c.nbx = c.nby = 0;
    c.bt = batch_of(h, l);
    if (nprob >= 0) c.bt.nprob = nprob;   // a subset of the level's problems (pointers pre-offset)
#define SBP_TILE_CASE(TI_, J_, A_)                                                              \
    if (ti == TI_ && tj == J_ && ta == A_) {                                                    \
        if constexpr (A_ == 0 || OUT::kAlign)                                                   \
            return launch_tile<TI_, J_, A_, T, IN, OUT>(h->p, L.tb, in, out, c, h->st);        \
        else                                                                                    \
            return SBP_ERR_INVALID_ARG;   /* pass_shape never selects this */                   \
    }
    SBP_TILE_SHAPES(SBP_TILE_CASE)
#undef SBP_TILE_CASE
    return SBP_ERR_INVALID_ARG;
}

// rc = scale * FW(b - A u) from level l to l+1, every problem.
template <typename T>
int launch_rr(sbp_handle h, int l, const T* u, const T* b, T* rc, const Scalars* scb) {
    const Level& L = h->L[l];
    // the residual + restriction kernel is compiled for the unaligned shapes
    const bool al = L.ta[prec_index<T>()] != 0;
    const int ti = al ? 256 : L.ti[prec_index<T>()], tj = al ? 8 : L.tj[prec_index<T>()];
    const Batch bt = batch_of(h, l);
    const int Nc = h->L[l + 1].N;
    const T scale = (T)h->prm.restrict_scale;
    const int* done = scb ? &scb->done : nullptr;
#define SBP_RR_CASE(TI_, J_, A_) \
    if (ti == TI_ && tj == J_ && A_ == 0)   \
        return launch_resid_restrict<TI_, J_, T>(h->p, L.tb, u, b, rc, Nc, scale, done, kScalarInts, bt, h->st);
    SBP_TILE_SHAPES(SBP_RR_CASE)
#undef SBP_RR_CASE
    return SBP_ERR_INVALID_ARG;
}

int alloc_level_buffers(sbp_handle h, int prec) {
    const size_t es = prec == 0 ? sizeof(double) : sizeof(float);
    for (int l = 0; l < h->nlev; ++l) {
        Level& L = h->L[l];
        const size_t bytes = L.n * es * h->np[l];
        if (l > 0 || h->ragged) HIPCHK(hipMalloc(&L.b[prec], bytes));
        HIPCHK(hipMalloc(&L.xo[prec], bytes));
        if (l < h->nlev - 1) {
            HIPCHK(hipMalloc(&L.xa[prec], bytes));
            HIPCHK(hipMalloc(&L.xb[prec], bytes));
        }
    }
    return SBP_OK;
}

// ---------------------------------------------------------------------------
// Smoothing / V-cycle
// ---------------------------------------------------------------------------
// One smoothing pass: dst = S(operand) with the operand formed by IN.
template <typename T, class IN>
int smooth_pass(sbp_handle h, int l, const IN& in, const T* b, T* dst, const T* dotv, const Scalars* scb,
                double* partial) {
    const bool rich = h->prm.smoother == 1;
    const T om = (T)(rich ? h->prm.richardson_omega : h->prm.omega);
    if (dotv) {
        OutSmoothDot<T> o{b, dst, om, dotv, rich ? 0 : 1};
        return launch_pass<T>(h, l, in, o, scb, partial, nullptr);
    }
    OutSmooth<T> o{b, dst, om, rich ? 0 : 1};
    return launch_pass<T>(h, l, in, o, scb, nullptr, nullptr);
}

// Whether level l runs the fused two-sweep kernels: when the level's points
// over all stacked problems are at most fuse_n^2 (launch-latency-bound
// levels; above that the redundant halo work of the fused blocks costs more
// than the passes it saves).
inline bool fuse_level(const sbp_handle h, int l) {
    if (h->p > 6 || h->prm.drop_coarse_rhs) return false;   // the two-sweep kernels are compiled for p = 2, 4, 6 (fused_shape_ok)
    if (!fused_shape_ok(h->L[l].fti[0], h->L[l].ftj[0], h->p) || !fused_shape_ok(h->L[l].fti[1], h->L[l].ftj[1], h->p))
        return false;
    const long long pts = (long long)h->L[l].n * h->np[l];
    return pts <= (long long)h->prm.fuse_n * h->prm.fuse_n;
}

template <typename T>
int vcycle_level(sbp_handle h, int l, const T* b, T* out, const T* dotv, const Scalars* scb, double* partial);

// Levels bottom_l0 .. nlev-2 of the V-cycle in two single-workgroup launches
// per problem, the down and the up leg, around the coarsest level's solve.
template <typename T>
int launch_bottom(sbp_handle h, const T* b, T* out, const T* dotv, const Scalars* scb, double* partial) {
    const int l0 = h->bottom_l0, lc = h->nlev - 1;
    const int pi = prec_index<T>();
    BottomCtl c{};
    c.nlev = lc - l0;
    for (int k = 0; k < c.nlev; ++k) {
        const Level& L = h->L[l0 + k];
        c.tab[k] = L.dtab;
        c.side[k] = L.N;
        c.x2[k] = L.xa[pi];
        c.b[k] = k == 0 ? (void*)b : L.b[pi];
        c.sz[k] = (long long)L.n;
    }
    const Level& Lc = h->L[lc];
    c.side[c.nlev] = Lc.N;
    c.b[c.nlev] = Lc.b[pi];
    c.sz[c.nlev] = (long long)Lc.n;
    c.ec = Lc.xo[pi];
    c.nu1 = h->prm.nu_pre;
    c.nu2 = h->prm.nu_post;
    c.use_diag = h->prm.smoother == 1 ? 0 : 1;
    c.omega = h->prm.smoother == 1 ? h->prm.richardson_omega : h->prm.omega;
    c.scale = h->prm.restrict_scale;
    c.done = scb ? &scb->done : nullptr;
    c.ds = kScalarInts;
    c.ss = 1;
    c.partial = partial;
    if (h->p == 2)
        hipLaunchKernelGGL((vcycle_bottom_down<2, T>), dim3(h->nprob), dim3(kBottomThreads), 0, h->st, c);
    else if (h->p == 4)
        hipLaunchKernelGGL((vcycle_bottom_down<4, T>), dim3(h->nprob), dim3(kBottomThreads), 0, h->st, c);
    else
        return SBP_ERR_INVALID_ORDER;   // not compiled for p > 4 (never selected for them)
    HIPCHK(hipGetLastError());
    STCHK(vcycle_level<T>(h, lc, (const T*)Lc.b[pi], (T*)Lc.xo[pi], nullptr, scb, nullptr));
    if (h->p == 2)
        hipLaunchKernelGGL((vcycle_bottom_up<2, T>), dim3(h->nprob), dim3(kBottomThreads), 0, h->st, out, dotv, c);
    else
        hipLaunchKernelGGL((vcycle_bottom_up<4, T>), dim3(h->nprob), dim3(kBottomThreads), 0, h->st, out, dotv, c);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

constexpr double kExactRtol = 1e-13;   // "exact" solves: ||r|| < 1e-13 ||b||
constexpr int kExactMaxiter = 100;

// x = A^-1 b on handle c, for the reference's sparse LU solve A_2h \ b_2h:
// MG-PCG from zero to ||r|| < 1e-13 ||b|| (round-off level for FP64).  An
// iterate short of that after kExactMaxiter iterations is still returned as
// the guess; HIP errors, NaN/Inf and breakdown are reported.
int exact_solve(sbp_handle c, const double* b, double* x) {
    HIPCHK(hipMemsetAsync(x, 0, c->L[0].n * sizeof(double), c->st));
    double bb = 0.0;
    STCHK(sbp_dot(c, b, b, &bb));
    if (!std::isfinite(bb)) return SBP_ERR_NONFINITE;
    if (bb == 0.0) return SBP_OK;
    int it = 0;
    double rn = 0.0;
    const int rc = sbp_mgpcg(c, b, x, kExactRtol * std::sqrt(bb), kExactMaxiter, &it, &rn, nullptr);
    return rc == SBP_ERR_NOT_CONVERGED ? SBP_OK : rc;
}

// Coarsest level above 33 x 33: out = A_c^-1 b by exact_solve on the nested
// handle (one problem; FP32 data is staged through FP64), and with dotv the
// single partial dotv' out.
template <typename T>
int coarse_exact(sbp_handle h, const T* b, T* out, const T* dotv, double* partial) {
    const Level& L = h->L[h->nlev - 1];
    if (h->nprob != 1) return SBP_ERR_INVALID_ARG;
    if (!h->cexact) {
        sbp_params cp;
        sbp_default_params(&cp);
        STCHK(sbp_create(L.N, h->p, &cp, &h->cexact));
        if (hipMalloc(&h->cbuf, 2 * L.n * sizeof(double)) != hipSuccess) return SBP_ERR_HIP;
    }
    h->cexact->st = h->st;
    if constexpr (std::is_same<T, double>::value) {
        STCHK(exact_solve(h->cexact, b, out));
    } else {
        double* b64 = h->cbuf;
        double* x64 = h->cbuf + L.n;
        hipLaunchKernelGGL((cast_vec<T, double, kEwThreads>), dim3(ew_blocks(L.n)), dim3(kEwThreads), 0, h->st, b,
                           b64, L.n);
        HIPCHK(hipGetLastError());
        STCHK(exact_solve(h->cexact, b64, x64));
        hipLaunchKernelGGL((cast_vec<double, T, kEwThreads>), dim3(ew_blocks(L.n)), dim3(kEwThreads), 0, h->st,
                           (const double*)x64, out, L.n);
        HIPCHK(hipGetLastError());
    }
    if (dotv) {
        hipLaunchKernelGGL((dot_partial<T, kEwThreads>), dim3(1, 1), dim3(kEwThreads), 0, h->st, dotv, (const T*)out,
                           L.n, partial);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

// V-cycle on level l: x_out ~ A_l^-1 b (zero initial guess), every problem.
// When dotv is non-null the last pass also writes block partials of dotv' x_out.
template <typename T>
int vcycle_level(sbp_handle h, int l, const T* b, T* out, const T* dotv, const Scalars* scb, double* partial) {
    const int pi = prec_index<T>();
    Level& L = h->L[l];
    hipStream_t st = h->st;
    if (l == h->bottom_l0) return launch_bottom<T>(h, b, out, dotv, scb, partial);
    if (l == h->nlev - 1) {
        if (h->coarse_nested) return coarse_exact<T>(h, b, out, dotv, partial);
        hipLaunchKernelGGL((coarse_solve<T>), dim3(((int)L.n + kCoarseRows - 1) / kCoarseRows, h->np[l]), dim3(256), 0,
                           st, h->Ainv, b, out, (int)L.n, scb);
        HIPCHK(hipGetLastError());
        if (dotv) {
            hipLaunchKernelGGL((dot_partial<T, kEwThreads>), dim3(1, h->np[l]), dim3(kEwThreads), 0, st, dotv,
                               (const T*)out, L.n, partial);
            HIPCHK(hipGetLastError());
        }
        return SBP_OK;
    }
    T* xa = (T*)L.xa[pi];
    T* xb = (T*)L.xb[pi];
    const int nu1 = h->prm.nu_pre, nu2 = h->prm.nu_post;
    const T om = (T)(h->prm.smoother == 1 ? h->prm.richardson_omega : h->prm.omega);
    // Pre-smoothing from zero.  The first sweep x1 = w D^-1 b is formed on
    // the fly inside the second sweep's pass.
    T* cur = xa;
    const int use_diag = h->prm.smoother == 1 ? 0 : 1;
    const int* done = scb ? &scb->done : nullptr;
    const bool fused_pre = nu1 == 2 && fuse_level(h, l);
    const bool fused_post = nu2 == 2 && fuse_level(h, l);
    if (fused_pre) {
        // x2 = S(S(0)) and the restricted residual in one pass
        const int rc = launch_smooth2_restrict<T>(h->p, L.fti[pi], L.ftj[pi], L.tb, b, xa, (T*)h->L[l + 1].b[pi], om, use_diag,
                                                  (T)h->prm.restrict_scale, done, kScalarInts, batch_of(h, l), st);
        if (rc != SBP_OK) return rc;
    } else if (nu1 == 0) {
        hipLaunchKernelGGL((fill<T, kEwThreads>), ew_grid_l(h, l), dim3(kEwThreads), 0, st, xa, T(0), L.n, scb);
        HIPCHK(hipGetLastError());
    } else {
        const InSmooth0<T> in0{b, om, h->prm.smoother == 1 ? 0 : 1};
        if (nu1 == 1) {
            STCHK(launch_pass<T>(h, l, in0, OutCopy<T>{xa}, scb, nullptr, nullptr));
        } else {
            STCHK(smooth_pass<T>(h, l, in0, b, xa, nullptr, scb, nullptr));
            for (int s = 2; s < nu1; ++s) {
                T* nxt = (cur == xa) ? xb : xa;
                STCHK(smooth_pass<T>(h, l, InPlain<T>{cur}, b, nxt, nullptr, scb, nullptr));
                cur = nxt;
            }
        }
    }
    // Residual and full-weighting restriction, fused (the fine residual stays on chip).
    Level& Lc = h->L[l + 1];
    // (drop_coarse_rhs: below the finest level the restricted residual is -A x,
    // the reference's three-level cycle)
    if (!fused_pre)
        STCHK(launch_rr<T>(h, l, (const T*)cur, (l > 0 && h->prm.drop_coarse_rhs) ? nullptr : b, (T*)Lc.b[pi], scb));
    T* ec = (T*)Lc.xo[pi];
    // ragged: problems whose finest level is l + 1 take r'z from that level's last pass
    const bool rz_c = h->ragged && h->np[l + 1] > h->j0[l + 1];
    STCHK(vcycle_level<T>(h, l + 1, (const T*)Lc.b[pi], ec, rz_c ? (const T*)Lc.b[pi] : nullptr, scb,
                          rz_c ? h->partial + h->roff[l + 1] : nullptr));
    // Prolongation folded into the first post-smoothing pass.
    if (fused_post)
        return launch_prolong_smooth2<T>(h->p, L.fti[pi], L.ftj[pi], L.tb, (const T*)cur, (const T*)ec, b, out, dotv, om, use_diag,
                                         done, kScalarInts, partial, batch_of(h, l), st);
    if (nu2 == 0) {
        hipLaunchKernelGGL((prolong_add<T>), dim3((L.N + 255) / 256, L.N, h->np[l]), dim3(256), 0, st, (const T*)ec,
                           cur, L.N, Lc.N, scb);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((copy_vec<T, kEwThreads>), ew_grid_l(h, l), dim3(kEwThreads), 0, st, (const T*)cur, out,
                           L.n, scb);
        HIPCHK(hipGetLastError());
        if (dotv) {
            hipLaunchKernelGGL((dot_partial<T, kEwThreads>), ew_grid_l(h, l), dim3(kEwThreads), 0, st, dotv,
                               (const T*)out, L.n, partial);
            HIPCHK(hipGetLastError());
        }
        return SBP_OK;
    }
    T* other = (cur == xa) ? xb : xa;
    {
        T* dst = (nu2 == 1) ? out : other;
        STCHK(smooth_pass<T>(h, l, InProlong<T>{cur, ec, Lc.N}, b, dst, nu2 == 1 ? dotv : nullptr, scb,
                             nu2 == 1 ? partial : nullptr));
        cur = dst;
    }
    for (int s = 1; s < nu2; ++s) {
        const bool last = (s == nu2 - 1);
        T* dst = last ? out : ((cur == xa) ? xb : xa);
        STCHK(smooth_pass<T>(h, l, InPlain<T>{cur}, b, dst, last ? dotv : nullptr, scb, last ? partial : nullptr));
        cur = dst;
    }
    return SBP_OK;
}

// Number of partials per problem the V-cycle's final pass writes on level l.
template <typename T>
int vcycle_nparts(sbp_handle h, int l = 0) {
    const Level& L = h->L[l];
    if (l == h->nlev - 1 || h->bottom_l0 == l) return 1;
    if (h->prm.nu_post == 2 && fuse_level(h, l))
        return fused_post_blocks(L.N, L.fti[prec_index<T>()], L.ftj[prec_index<T>()]);
    if (h->prm.nu_post == 0) return ew_blocks(L.n);
    return (int)pass_nblocks<T, OutSmoothDot<T>>(h->p, L);
}

void drop_graphs(sbp_handle h) {
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            for (int c = 0; c < 2; ++c) {
                if (h->gexec[a][b][c]) (void)hipGraphExecDestroy(h->gexec[a][b][c]);
                if (h->graph[a][b][c]) (void)hipGraphDestroy(h->graph[a][b][c]);
                h->gexec[a][b][c] = nullptr;
                h->graph[a][b][c] = nullptr;
            }
        }
}

// Copies every problem's scalar block to the pinned mirror; *all_done (when
// given) tells whether every problem is done.
int poll(sbp_handle h, const Scalars* sc, bool* all_done = nullptr) {
    HIPCHK(hipMemcpyAsync(h->sc_host, sc, sizeof(Scalars) * h->nprob, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (all_done) {
        bool d = true;
        for (int q = 0; q < h->nprob; ++q) d = d && h->sc_host[q].done;
        *all_done = d;
    }
    return SBP_OK;
}

// A poll in two halves: poll_start queues the scalar blocks' copy to the
// pinned mirror and an event after it; poll_wait waits for that event only,
// so work queued between the two runs on while the host waits.
int poll_start(sbp_handle h, const Scalars* sc) {
    if (!h->pev) HIPCHK(hipEventCreateWithFlags(&h->pev, hipEventDisableTiming));
    HIPCHK(hipMemcpyAsync(h->sc_host, sc, sizeof(Scalars) * h->nprob, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipEventRecord(h->pev, h->st));
    return SBP_OK;
}

int poll_wait(sbp_handle h, bool* all_done) {
    HIPCHK(hipEventSynchronize(h->pev));
    bool d = true;
    for (int q = 0; q < h->nprob; ++q) d = d && h->sc_host[q].done;
    *all_done = d;
    return SBP_OK;
}

int status_of(const Scalars& s) {
    if (s.status == 5) return SBP_ERR_NONFINITE;
    if (s.status == 6) return SBP_ERR_BREAKDOWN;
    if (s.status == 7 || !s.done) return SBP_ERR_NOT_CONVERGED;
    return SBP_OK;
}

// Worst status over the problems (errors first, then not-converged).
int batch_status(sbp_handle h) {
    int worst = SBP_OK;
    for (int q = 0; q < h->nprob; ++q) {
        const int s = status_of(h->sc_host[q]);
        if (s == SBP_ERR_NONFINITE || s == SBP_ERR_BREAKDOWN) return s;
        if (s != SBP_OK) worst = s;
    }
    return worst;
}

int ensure32(sbp_handle h) {
    if (h->have32) return SBP_OK;
    const size_t n = h->ragged ? h->rtot : h->L[0].n * h->nprob;
    HIPCHK(hipMalloc(&h->r32, n * sizeof(float)));
    HIPCHK(hipMalloc(&h->p32, n * sizeof(float)));
    HIPCHK(hipMalloc(&h->p32b, n * sizeof(float)));
    HIPCHK(hipMalloc(&h->Ap32, n * sizeof(float)));
    HIPCHK(hipMalloc(&h->z32, n * sizeof(float)));
    HIPCHK(hipMalloc(&h->d32, n * sizeof(float)));
    if (h->mg) STCHK(alloc_level_buffers(h, 1));
    h->have32 = true;
    return SBP_OK;
}

// Handle-owned Krylov work vectors of one precision.
template <typename T>
struct KVecs {
    T* x;      // iterate
    T* r;      // residual
    T* p[2];   // search directions (ping-pong: iteration k writes p[k % 2])
    T* Ap;
    T* z;      // preconditioned residual
};

template <typename T> KVecs<T> kvecs(sbp_handle h);
template <> KVecs<double> kvecs<double>(sbp_handle h) {
    KVecs<double> v{h->x, h->r, {h->pv, h->pv2}, h->Ap, h->z};
    return v;
}
template <> KVecs<float> kvecs<float>(sbp_handle h) {
    KVecs<float> v{h->d32, h->r32, {h->p32, h->p32b}, h->Ap32, h->z32};
    return v;
}

// Ragged handles: the Krylov vectors of the problems whose finest level is l
// (ranks j0[l] .. np[l]-1).  x, p[2] and Ap are the handle's ragged arrays
// (those problems from element koff[l]); r and z are their slots of level
// l's right-hand-side and output stacks, which the V-cycle reads and writes
// (r is the V-cycle's right-hand side on that level, z its result).
template <typename T>
KVecs<T> rkvecs(sbp_handle h, int l) {
    const int pi = prec_index<T>();
    const Level& L = h->L[l];
    const size_t off = h->koff[l];
    const size_t slot = (size_t)h->j0[l] * L.n;
    const KVecs<T> v0 = kvecs<T>(h);
    KVecs<T> v;
    v.x = v0.x + off;
    v.p[0] = v0.p[0] + off;
    v.p[1] = v0.p[1] + off;
    v.Ap = v0.Ap + off;
    v.r = (T*)L.b[pi] + slot;
    v.z = (T*)L.xo[pi] + slot;
    return v;
}

inline int joiners(const sbp_handle h, int l) { return h->np[l] - h->j0[l]; }
inline double* rpart(sbp_handle h, int l) { return h->partial + h->roff[l]; }

// The V-cycle of every problem of a ragged handle (all start on the level of
// their own side; each level's last pass leaves r'z partials of the problems
// that joined there).
template <typename T>
int vcycle_ragged(sbp_handle h, Scalars* sc) {
    const int pi = prec_index<T>();
    Level& L0 = h->L[0];
    const bool rz0 = joiners(h, 0) > 0;
    return vcycle_level<T>(h, 0, (const T*)L0.b[pi], (T*)L0.xo[pi], rz0 ? (const T*)L0.b[pi] : nullptr, sc,
                           rz0 ? rpart(h, 0) : nullptr);
}

// The variable-coefficient operator's pass and V-cycle (defined with the
// rest of that operator below); Krylov loops of such a handle call them in
// place of the constant-coefficient ones.
template <typename T, class IN, class OUT>
int launch_var_pass(sbp_handle h, int level, const IN& in, const OUT& out, const Scalars* sc, int* nparts);
template <typename T>
int var_vcycle(sbp_handle h, int l, const T* b, T* x, const Scalars* sc, int* np_rz);

// z = V(r) with the r'z partials in h->partial (*nparts of them per problem).
template <typename T>
int krylov_vcycle(sbp_handle h, const KVecs<T>& v, Scalars* sc, int* nparts) {
    if constexpr (std::is_same<T, double>::value) {
        if (h->var) {
            int nrz = 0;
            STCHK(var_vcycle<double>(h, 0, v.r, v.z, sc, &nrz));
            if (nrz == 0) {   // no post-smoothing: r'z by its own kernel
                const size_t n = h->L[0].n;
                hipLaunchKernelGGL((dot_partial<double, kEwThreads>), dim3(ew_blocks(n), 1), dim3(kEwThreads), 0, h->st,
                                   (const double*)v.r, (const double*)v.z, n, h->partial);
                HIPCHK(hipGetLastError());
                nrz = (int)ew_blocks(n);
            }
            *nparts = nrz;
            return SBP_OK;
        }
    }
    STCHK(vcycle_level<T>(h, 0, v.r, v.z, v.r, sc, h->partial));
    *nparts = vcycle_nparts<T>(h);
    return SBP_OK;
}

// One Krylov iteration of a ragged handle: per level with joiners the fused
// search-direction pass, alpha, the residual update and the convergence test
// (one launch each over that level's joiners), then one V-cycle over every
// problem and beta per level.
template <typename T>
int issue_iteration_rag(sbp_handle h, Scalars* sc, bool use_mg, int par) {
    hipStream_t st = h->st;
    for (int l = 0; l < h->nlev; ++l) {
        const int nj = joiners(h, l);
        if (nj <= 0) continue;
        const KVecs<T> v = rkvecs<T>(h, l);
        Scalars* s = sc + h->j0[l];
        double* part = rpart(h, l);
        const size_t n = h->L[l].n;
        T* pold = v.p[par ^ 1];
        T* pnew = v.p[par];
        const T* zz = use_mg ? v.z : v.r;
        int np = 0;
        STCHK(launch_pass<T>(h, l, InAxpy<T>{zz, pold, &s->beta, T(0)},
                             OutPUpd<T>{pnew, v.Ap, v.x, pold, &s->alpha, T(0)}, s, part, &np, nj));
        hipLaunchKernelGGL(krylov_alpha, dim3(nj), dim3(kRedThreads), 0, st, part, np, s, use_mg ? 1 : 0);
        const dim3 g(ew_blocks(n), nj);
        hipLaunchKernelGGL((update_r<T, kEwThreads>), g, dim3(kEwThreads), 0, st, v.r, (const T*)v.Ap,
                           (const Scalars*)s, n, part);
        hipLaunchKernelGGL(krylov_check, dim3(nj), dim3(kRedThreads), 0, st, part, (int)g.x, s, use_mg ? 0 : 1);
        HIPCHK(hipGetLastError());
    }
    if (use_mg) {
        STCHK(vcycle_ragged<T>(h, sc));
        for (int l = 0; l < h->nlev; ++l) {
            const int nj = joiners(h, l);
            if (nj <= 0) continue;
            const int nvp = vcycle_nparts<T>(h, l);
            hipLaunchKernelGGL(krylov_beta, dim3(nj), dim3(kRedThreads), 0, st, rpart(h, l) + (size_t)h->j0[l] * nvp,
                               nvp, sc + h->j0[l]);
        }
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

// Iteration k (1-based, k % 2 == par): p_k = z_{k-1} + beta p_{k-1} formed on
// the fly inside the apply, which also writes A p_k, accumulates p_k'A p_k and
// applies the pending x += alpha_{k-1} p_{k-1}; then alpha, r -= alpha A p_k
// with r'r, the convergence test, and (MG) z = V(r) with r'z and beta.
// Every kernel tests its problem's "done" flag.
template <typename T>
int issue_iteration(sbp_handle h, const KVecs<T>& v, Scalars* sc, bool use_mg, int par) {
    if (h->ragged) return issue_iteration_rag<T>(h, sc, use_mg, par);
    hipStream_t st = h->st;
    const size_t n = h->L[0].n;
    T* pold = v.p[par ^ 1];
    T* pnew = v.p[par];
    const T* zz = use_mg ? v.z : v.r;
    int np = 0;
    if constexpr (std::is_same<T, double>::value) {
        if (h->var)
            STCHK(launch_var_pass<double>(h, 0, VarInAxpy{zz, pold, &sc->beta, 0.0},
                                          VarOutPUpd{pnew, v.Ap, v.x, pold, &sc->alpha, 0.0}, sc, &np));
    }
    if (!h->var)
        STCHK(launch_pass<T>(h, 0, InAxpy<T>{zz, pold, &sc->beta, T(0)},
                             OutPUpd<T>{pnew, v.Ap, v.x, pold, &sc->alpha, T(0)}, sc, h->partial, &np));
    hipLaunchKernelGGL(krylov_alpha, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, np, sc, use_mg ? 1 : 0);
    const dim3 g = ew_grid(h, n);
    hipLaunchKernelGGL((update_r<T, kEwThreads>), g, dim3(kEwThreads), 0, st, v.r, (const T*)v.Ap,
                       (const Scalars*)sc, n, h->partial);
    hipLaunchKernelGGL(krylov_check, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc,
                       use_mg ? 0 : 1);
    HIPCHK(hipGetLastError());
    if (use_mg) {
        int nvp = 0;
        STCHK(krylov_vcycle<T>(h, v, sc, &nvp));
        hipLaunchKernelGGL(krylov_beta, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, nvp, sc);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

// Captures one iteration of each parity once per (precision, solver) into a
// hipGraph on the handle's own capture stream; replays go to the handle's
// stream, alternating the parities (the search directions ping-pong).
template <typename T>
int iteration_graph(sbp_handle h, const KVecs<T>& v, Scalars* sc, bool use_mg, int par, hipGraphExec_t* out) {
    const int a = prec_index<T>(), b = use_mg ? 1 : 0;
    if (!h->gexec[a][b][par]) {
        if (!h->cap) HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
        hipStream_t saved = h->st;
        h->st = h->cap;
        const int rc = issue_iteration<T>(h, v, sc, use_mg, par);
        h->st = saved;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->cap, &g);
        if (rc != SBP_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        HIPCHK(e);
        hipGraphExec_t ex = nullptr;
        if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            return SBP_ERR_HIP;
        }
        h->graph[a][b][par] = g;
        h->gexec[a][b][par] = ex;
    }
    *out = h->gexec[a][b][par];
    return SBP_OK;
}

// Iterations until the next host poll (poll_every = 0): from each unfinished
// problem's residual history, ||r_k|| = ||r_0|| rho^k, the number still needed
// to reach its tolerance, at least 1 and at most kPollAhead; the largest
// over the problems.
constexpr int kPollAhead = 16;
int poll_ahead(const sbp_handle h) {
    int need = 1;
    for (int q = 0; q < h->nprob; ++q) {
        const Scalars& s = h->sc_host[q];
        if (s.done) continue;
        int k = kPollAhead;
        if (s.iters > 0 && s.resnorm > 0.0 && s.resnorm < s.rnorm0 && s.abstol > 0.0) {
            const double lrho = std::log(s.resnorm / s.rnorm0) / s.iters;   // < 0
            // nearest iteration: the rate of a multigrid-preconditioned solve
            // improves as it runs, so rounding up would overshoot
            const double est = std::ceil(std::log(s.abstol / s.resnorm) / lrho - 0.5);
            k = est < 1.0 ? 1 : (est > kPollAhead ? kPollAhead : (int)est);
        }
        need = std::max(need, k);
    }
    return need;
}

// Issues up to maxiter Krylov iterations (issue(k) queues iteration k, from
// 0) with the host polls between them: every poll_every iterations, or
// (poll_every = 0) after two iterations and then where the measured
// convergence rate predicts the last problem converges, so that few launches
// run for finished problems.  The first adaptive poll only measures the
// rate: one more iteration is queued before the host waits for the scalars,
// so the device does not idle through the round trip (that iteration's
// kernels fall through on the done flags if every problem has already
// converged).  Returns once every problem is done or maxiter iterations are
// queued; the caller's last poll drains the stream.
template <class F>
int krylov_iterate(sbp_handle h, const Scalars* sc, int maxiter, F&& issue) {
    const bool adaptive = h->prm.poll_every == 0;
    int next = adaptive ? 2 : h->prm.poll_every;
    bool first = true, all = false;
    for (int it = 0; it < maxiter && !all;) {
        STCHK(issue(it));
        ++it;
        while (it >= next && it < maxiter) {
            const int seen = it;   // iterations the copied scalars reflect
            STCHK(poll_start(h, sc));
            if (adaptive && first) {
                STCHK(issue(it));
                ++it;
            }
            first = false;
            STCHK(poll_wait(h, &all));
            if (all) break;
            next = seen + (adaptive ? poll_ahead(h) : h->prm.poll_every);
        }
    }
    return SBP_OK;
}

// Preconditioned CG body shared by FP64 and FP32: solves A x = b for every
// problem given the initial residuals in v.r (scalar blocks already set by the
// caller's init kernel); the preconditioner is the V-cycle when use_mg, else
// the identity.  The host polls now and then (poll_ahead) and stops issuing
// once every problem is done.  Leaves the scalar blocks in h->sc_host.
template <typename T>
int pcg_loop(sbp_handle h, const KVecs<T>& v, Scalars* sc, int maxiter, bool use_mg, bool final_poll = true) {
    hipStream_t st = h->st;
    const size_t n = h->L[0].n;
    const T* z0 = v.r;
    if (h->ragged) {
        if (use_mg) {
            STCHK(vcycle_ragged<T>(h, sc));
            for (int l = 0; l < h->nlev; ++l) {
                const int nj = joiners(h, l);
                if (nj <= 0) continue;
                const int nvp = vcycle_nparts<T>(h, l);
                hipLaunchKernelGGL(krylov_set_rz, dim3(nj), dim3(kRedThreads), 0, st,
                                   rpart(h, l) + (size_t)h->j0[l] * nvp, nvp, sc + h->j0[l]);
            }
        }
        for (int l = 0; l < h->nlev; ++l) {
            const int nj = joiners(h, l);
            if (nj <= 0) continue;
            const KVecs<T> vl = rkvecs<T>(h, l);
            hipLaunchKernelGGL((copy_vec<T, kEwThreads>), dim3(ew_blocks(h->L[l].n), nj), dim3(kEwThreads), 0, st,
                               (const T*)(use_mg ? vl.z : vl.r), vl.p[0], h->L[l].n, (const Scalars*)(sc + h->j0[l]));
        }
        HIPCHK(hipGetLastError());
    } else {
        if (use_mg) {
            int nvp = 0;
            STCHK(krylov_vcycle<T>(h, v, sc, &nvp));
            hipLaunchKernelGGL(krylov_set_rz, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, nvp, sc);
            HIPCHK(hipGetLastError());
            z0 = v.z;
        }
        // p_0 := z_0 (any finite values serve: alpha = beta = 0 in the first pass)
        hipLaunchKernelGGL((copy_vec<T, kEwThreads>), ew_grid(h, n), dim3(kEwThreads), 0, st, z0, v.p[0], n,
                           (const Scalars*)sc);
        HIPCHK(hipGetLastError());
    }
    hipGraphExec_t ex[2] = {nullptr, nullptr};
    // (a nested coarse solve synchronises with the host: no capture)
    if (h->prm.use_graph && !(use_mg && h->coarse_nested)) {
        STCHK(iteration_graph<T>(h, v, sc, use_mg, 1, &ex[1]));
        STCHK(iteration_graph<T>(h, v, sc, use_mg, 0, &ex[0]));
    }
    STCHK(krylov_iterate(h, sc, maxiter, [&](int k) -> int {
        const int par = (k & 1) ? 0 : 1;   // iteration 0 has parity 1
        if (ex[par]) {
            HIPCHK(hipGraphLaunch(ex[par], st));
            return SBP_OK;
        }
        return issue_iteration<T>(h, v, sc, use_mg, par);
    }));
    if (h->ragged) {
        for (int l = 0; l < h->nlev; ++l) {
            const int nj = joiners(h, l);
            if (nj <= 0) continue;
            const KVecs<T> vl = rkvecs<T>(h, l);
            hipLaunchKernelGGL((flush_x<T, kEwThreads>), dim3(ew_blocks(h->L[l].n), nj), dim3(kEwThreads), 0, st,
                               vl.x, (const T*)vl.p[0], (const T*)vl.p[1], (const Scalars*)(sc + h->j0[l]),
                               h->L[l].n);
        }
    } else {
        hipLaunchKernelGGL((flush_x<T, kEwThreads>), ew_grid(h, n), dim3(kEwThreads), 0, st, v.x, (const T*)v.p[0],
                           (const T*)v.p[1], (const Scalars*)sc, n);
    }
    HIPCHK(hipGetLastError());
    return final_poll ? poll(h, sc) : SBP_OK;
}

// Per-problem tolerances to the device (through pinned staging, stream-ordered).
int put_abstol(sbp_handle h, const double* abstol) {
    HIPCHK(hipStreamSynchronize(h->st));   // the staging buffer may still feed an earlier copy
    for (int q = 0; q < h->nprob; ++q) h->habs[q] = abstol[q];
    HIPCHK(hipMemcpyAsync(h->dabs, h->habs, sizeof(double) * h->nprob, hipMemcpyHostToDevice, h->st));
    return SBP_OK;
}

// Per-problem results from the pinned mirror (seconds: device wall clock from
// the solve's start to the moment the problem stopped, -1 when it never did).
void get_results(sbp_handle h, int* iters, double* resnorm, double* seconds) {
    for (int q = 0; q < h->nprob; ++q) {
        const Scalars& s = h->sc_host[q];
        if (iters) iters[q] = s.iters;
        if (resnorm) resnorm[q] = s.resnorm;
        if (seconds)
            seconds[q] = (s.t1 > s.t0 && h->clock_khz > 0) ? (double)(s.t1 - s.t0) / (h->clock_khz * 1e3) : -1.0;
    }
}

// FP64 (P)CG on the internal iterate: x -> h->x, init, iterate, h->x -> x.
int krylov64(sbp_handle h, const double* b, double* x, const double* abstol, int maxiter, double* hist,
             bool use_mg) {
    const size_t n = h->L[0].n;
    hipStream_t st = h->st;
    STCHK(put_abstol(h, abstol));
    hipLaunchKernelGGL(stamp_start, dim3(1), dim3(256), 0, st, h->sc, h->nprob);
    // (ragged: x and h->x are ragged arrays of rtot elements)
    const dim3 gx = h->ragged ? dim3(ew_blocks(h->rtot), 1) : ew_grid(h, n);
    const size_t nx = h->ragged ? h->rtot : n;
    hipLaunchKernelGGL((copy_vec<double, kEwThreads>), gx, dim3(kEwThreads), 0, st, x, h->x, nx,
                       (const Scalars*)nullptr);
    HIPCHK(hipGetLastError());
    if (h->ragged) {
        for (int l = 0; l < h->nlev; ++l) {
            const int nj = joiners(h, l);
            if (nj <= 0) continue;
            const KVecs<double> vl = rkvecs<double>(h, l);
            int np = 0;
            STCHK(launch_pass<double>(h, l, InPlain<double>{vl.x}, OutResid<double, double>{b + h->koff[l], vl.r},
                                      nullptr, rpart(h, l), &np, nj));
            hipLaunchKernelGGL(krylov_init, dim3(nj), dim3(kRedThreads), 0, st, rpart(h, l), np, h->sc + h->j0[l],
                               (const double*)h->dabs + h->j0[l], (double*)nullptr, maxiter);
        }
    } else {
        int np = 0;
        if (h->var)
            STCHK(launch_var_pass<double>(h, 0, VarIn<double>{h->x}, VarOutResid<double>{b, h->r}, nullptr, &np));
        else
            STCHK(launch_pass<double>(h, 0, InPlain<double>{h->x}, OutResid<double, double>{b, h->r}, nullptr,
                                      h->partial, &np));
        hipLaunchKernelGGL(krylov_init, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, np, h->sc,
                           (const double*)h->dabs, hist, maxiter);
    }
    HIPCHK(hipGetLastError());
    STCHK(pcg_loop<double>(h, kvecs<double>(h), h->sc, maxiter, use_mg));
    hipLaunchKernelGGL((copy_vec<double, kEwThreads>), gx, dim3(kEwThreads), 0, st, (const double*)h->x,
                       x, nx, (const Scalars*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return SBP_OK;
}

// Outer refinement check: resnorm = ||r||; converged -> both done; else arm
// the inner solver with abstol = inner_rtol * ||r|| and its iteration cap.
__global__ __launch_bounds__(1024) void ir_check(const double* partial, int np, Scalars* sob, Scalars* sib,
                                                double inner_rtol, int inner_maxiter) {
    Scalars* so = sob + blockIdx.x;
    Scalars* si = sib + blockIdx.x;
    double rs = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        si->iters = 0;
        si->status = 0;
        si->alpha = si->beta = 0.0;
        si->hist = nullptr;
        si->maxiter = inner_maxiter;
        if (so->done) { si->done = 1; return; }
        const double nr = sqrt(rs);
        so->resnorm = nr;
        so->rr = rs;
        si->rr = rs;
        si->resnorm = nr;
        si->rnorm0 = nr;
        if (!isfinite(rs)) { so->status = 5; mark_done(so); si->done = 1; return; }
        if (nr < so->abstol) { mark_done(so); si->done = 1; return; }
        si->abstol = inner_rtol * nr;
        si->done = 0;
    }
}

__global__ void ir_init(Scalars* so, const double* abstol, int nprob) {
    const unsigned long long t = wall_clock64();
    for (int q = threadIdx.x; q < nprob; q += blockDim.x) {
        so[q].abstol = abstol[q];
        so[q].done = 0;
        so[q].status = 0;
        so[q].iters = 0;
        so[q].dprev = 0.0;
        so[q].est = 0.0;
        so[q].inner = 0;
        so[q].t0 = t;
        so[q].t1 = 0;
    }
}

// x64 += (double) d32 with block partials of ||d||^2 (slot 2b) and ||x||^2
// (slot 2b+1).  Problems whose refinement has finished are left alone; a
// problem whose inner solve failed (non-finite or breakdown) takes that status
// and stops without the update.
template <int NT>
__global__ __launch_bounds__(NT) void ir_update(double* __restrict__ x, const float* __restrict__ d, size_t n,
                                               Scalars* sob, const Scalars* sib, double* partial) {
    const int q = blockIdx.y;
    if (sob[q].done) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) sob[q].inner += sib[q].iters;
    const int ist = sib[q].status;
    if (ist == 5 || ist == 6) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sob[q].status = ist;
            mark_done(sob + q);
        }
        return;
    }
    x += q * n;
    d += q * n;
    double dd = 0.0, xx = 0.0;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) {
        const double dv = (double)d[k];
        const double xv = x[k] + dv;
        x[k] = xv;
        dd += dv * dv;
        xx += xv * xv;
    }
    const double sd = block_sum<double, NT>(dd);
    __syncthreads();
    const double sx = block_sum<double, NT>(xx);
    if (threadIdx.x == 0) {
        double* pq = partial + (size_t)q * 2 * gridDim.x;
        pq[2 * blockIdx.x] = sd;
        pq[2 * blockIdx.x + 1] = sx;
    }
}

// Error estimate after correction k: ||e_k|| ~ ||d_k|| * min(1, ||d_k|| / ||d_{k-1}||)
// (the observed contraction of the refinement).  Stops that problem's
// refinement when est <= ir_rtol * ||x|| (ir_rtol > 0).
__global__ __launch_bounds__(1024) void ir_estimate(const double* partial, int nb, Scalars* sob, double ir_rtol) {
    Scalars* so = sob + blockIdx.x;
    if (so->done) return;
    const double* pq = partial + (size_t)blockIdx.x * 2 * nb;
    __shared__ double red[2][16];
    double sd = 0.0, sx = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) { sd += pq[2 * k]; sx += pq[2 * k + 1]; }
    sd = wave_sum(sd);
    sx = wave_sum(sx);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[0][w] = sd; red[1][w] = sx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double td = 0.0, tx = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) { td += red[0][k]; tx += red[1][k]; }
        const double dn = sqrt(td), xn = sqrt(tx);
        so->iters += 1;
        const double ratio = so->dprev > 0.0 ? fmin(1.0, dn / so->dprev) : 1.0;
        so->est = dn * ratio;
        so->dprev = dn;
        if (ir_rtol > 0.0 && so->iters >= 2 && so->est <= ir_rtol * xn) mark_done(so);
    }
}

// Mixed-precision refinement over every problem.  inner_tot (host, nprob)
// receives each problem's total inner iterations.  The final scalar blocks
// are left in h->sc_host.
int mixed_solve(sbp_handle h, const double* b, double* x, const double* abstol, int max_outer, int* inner_tot) {
    STCHK(ensure32(h));
    hipStream_t st = h->st;
    const size_t n = h->L[0].n;
    const int ub = std::min<int>(ew_blocks(n), (int)(h->npp / 2));
    STCHK(put_abstol(h, abstol));
    hipLaunchKernelGGL(ir_init, dim3(1), dim3(256), 0, st, h->sc, (const double*)h->dabs, h->nprob);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < h->nprob; ++q) inner_tot[q] = 0;
    for (int o = 0; o <= max_outer; ++o) {
        // FP64 residual r = b - A x, cast to FP32 for the inner solve.
        if (h->ragged) {
            for (int l = 0; l < h->nlev; ++l) {
                const int nj = joiners(h, l);
                if (nj <= 0) continue;
                int np = 0;
                STCHK(launch_pass<double>(h, l, InPlain<double>{x + h->koff[l]},
                                          OutResid<double, float>{b + h->koff[l], rkvecs<float>(h, l).r},
                                          h->sc + h->j0[l], rpart(h, l), &np, nj));
                hipLaunchKernelGGL(ir_check, dim3(nj), dim3(kRedThreads), 0, st, rpart(h, l), np, h->sc + h->j0[l],
                                   h->sc_in + h->j0[l], h->prm.inner_rtol, h->prm.inner_maxiter);
            }
        } else {
            int np = 0;
            STCHK(launch_pass<double>(h, 0, InPlain<double>{x}, OutResid<double, float>{b, h->r32}, h->sc,
                                      h->partial, &np));
            hipLaunchKernelGGL(ir_check, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, np, h->sc, h->sc_in,
                               h->prm.inner_rtol, h->prm.inner_maxiter);
        }
        HIPCHK(hipGetLastError());
        // The first residual is not checked from the host: a problem that is
        // already converged makes the inner solve's kernels fall through.
        bool all = false;
        if (o > 0 || o == max_outer) STCHK(poll(h, h->sc, &all));
        if (all || o == max_outer) break;
        // Inner FP32 MG-PCG on A d = r from d = 0.
        if (h->ragged) {
            for (int l = 0; l < h->nlev; ++l) {
                const int nj = joiners(h, l);
                if (nj <= 0) continue;
                hipLaunchKernelGGL((fill<float, kEwThreads>), dim3(ew_blocks(h->L[l].n), nj), dim3(kEwThreads), 0, st,
                                   h->d32 + h->koff[l], 0.0f, h->L[l].n, (const Scalars*)(h->sc_in + h->j0[l]));
            }
        } else {
            hipLaunchKernelGGL((fill<float, kEwThreads>), ew_grid(h, n), dim3(kEwThreads), 0, st, h->d32, 0.0f, n,
                               (const Scalars*)h->sc_in);
        }
        HIPCHK(hipGetLastError());
        // (no host wait after the inner solve: its iterations are added up on
        // the device by ir_update)
        STCHK(pcg_loop<float>(h, kvecs<float>(h), h->sc_in, h->prm.inner_maxiter, true, false));
        // FP64 update and the error estimate from the correction norms.
        if (h->ragged) {
            for (int l = 0; l < h->nlev; ++l) {
                const int nj = joiners(h, l);
                if (nj <= 0) continue;
                const int ubl = std::min<int>(ew_blocks(h->L[l].n), (int)(h->npl[l] / 2));
                hipLaunchKernelGGL((ir_update<kEwThreads>), dim3(ubl, nj), dim3(kEwThreads), 0, st, x + h->koff[l],
                                   (const float*)h->d32 + h->koff[l], h->L[l].n, h->sc + h->j0[l],
                                   (const Scalars*)(h->sc_in + h->j0[l]), rpart(h, l));
                hipLaunchKernelGGL(ir_estimate, dim3(nj), dim3(kRedThreads), 0, st, rpart(h, l), ubl, h->sc + h->j0[l],
                                   h->prm.ir_rtol);
            }
        } else {
            hipLaunchKernelGGL((ir_update<kEwThreads>), dim3(ub, h->nprob), dim3(kEwThreads), 0, st, x,
                               (const float*)h->d32, n, h->sc, (const Scalars*)h->sc_in, h->partial);
            hipLaunchKernelGGL(ir_estimate, dim3(h->nprob), dim3(kRedThreads), 0, st, h->partial, ub, h->sc,
                               h->prm.ir_rtol);
        }
        HIPCHK(hipGetLastError());
    }
    // Every exit from the loop follows a poll with nothing issued since, so
    // h->sc_host is current and the stream has drained.
    for (int q = 0; q < h->nprob; ++q) inner_tot[q] = h->sc_host[q].inner;
    return SBP_OK;
}

// Status of one finished refinement.
int mixed_status_of(const Scalars& s) {
    if (s.status == 5) return SBP_ERR_NONFINITE;
    if (s.status == 6) return SBP_ERR_BREAKDOWN;
    return s.done ? SBP_OK : SBP_ERR_NOT_CONVERGED;
}

// Worst refinement status over the problems (errors first, then not-converged).
int mixed_status(sbp_handle h) {
    int worst = SBP_OK;
    for (int q = 0; q < h->nprob; ++q) {
        const int st = mixed_status_of(h->sc_host[q]);
        if (st == SBP_ERR_NONFINITE || st == SBP_ERR_BREAKDOWN) return st;
        if (st != SBP_OK) worst = st;
    }
    return worst;
}

// Smoothing sweeps on level l for every problem (ping-pong through a level
// scratch buffer; the finest level borrows Ap / Ap32).
template <typename T>
int smooth_t(sbp_handle h, int level, const T* b, T* x, int nsweeps) {
    const int pi = prec_index<T>();
    Level& L = h->L[level];
    T* tmp;
    if (level < h->nlev - 1 && L.xb[pi]) tmp = (T*)L.xb[pi];
    else if (level == 0) tmp = pi == 0 ? (T*)(void*)h->Ap : (T*)(void*)h->Ap32;
    else tmp = (T*)L.xo[pi];
    if (!tmp || tmp == x || (const T*)tmp == b) return SBP_ERR_INVALID_ARG;
    T* cur = x;
    for (int s = 0; s < nsweeps; ++s) {
        T* dst = (cur == x) ? tmp : x;
        STCHK(smooth_pass<T>(h, level, InPlain<T>{cur}, b, dst, nullptr, nullptr, nullptr));
        cur = dst;
    }
    if (cur != x) {
        hipLaunchKernelGGL((copy_vec<T, kEwThreads>), ew_grid(h, L.n), dim3(kEwThreads), 0, h->st, (const T*)cur, x,
                           L.n, (const Scalars*)nullptr);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

// Exclusive scan of the CSR row counts into row pointers, three passes:
// per-block scans of kScanPer elements, a scan of the block totals in one
// block, and the block offsets added back (rp[n] = nnz).
constexpr int kScanThreads = 256, kScanItems = 8, kScanPer = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void scan_rows(const int32_t* counts, int32_t* rp, int n, int64_t* bsum) {
    __shared__ int64_t sh[kScanThreads];
    const int tid = threadIdx.x;
    const int base = blockIdx.x * kScanPer + tid * kScanItems;
    int64_t loc[kScanItems];
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        loc[k] = acc;
        acc += base + k < n ? counts[base + k] : 0;
    }
    sh[tid] = acc;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int64_t v = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += v;
        __syncthreads();
    }
    const int64_t excl = sh[tid] - acc;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) rp[base + k] = (int32_t)(excl + loc[k]);
    if (tid == kScanThreads - 1) bsum[blockIdx.x] = sh[tid];
}

__global__ __launch_bounds__(1024) void scan_block_sums(int64_t* bsum, int nb, int64_t* total) {
    __shared__ int64_t sh[1024];
    const int tid = threadIdx.x;
    int64_t carry = 0;
    for (int start = 0; start < nb; start += 1024) {
        const int idx = start + tid;
        const int64_t v = idx < nb ? bsum[idx] : 0;
        sh[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int64_t u = tid >= off ? sh[tid - off] : 0;
            __syncthreads();
            sh[tid] += u;
            __syncthreads();
        }
        if (idx < nb) bsum[idx] = carry + sh[tid] - v;
        carry += sh[1023];
        __syncthreads();
    }
    if (tid == 0) *total = carry;
}

__global__ __launch_bounds__(256) void scan_add(int32_t* rp, const int64_t* bsum, int n, const int64_t* total) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < n) rp[idx] += (int32_t)bsum[idx / kScanPer];
    if (idx == 0) rp[n] = (int32_t)*total;
}

// r = b - y with block partials of r'r (one problem).
template <int NT>
__global__ __launch_bounds__(NT) void axpby_dot(const double* __restrict__ b, const double* __restrict__ y,
                                               double* __restrict__ r, size_t n, double* partial) {
    double acc = 0.0;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) {
        const double v = b[k] - y[k];
        r[k] = v;
        acc += v * v;
    }
    const double s = block_sum<double, NT>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_vec16(const uint4* __restrict__ s, uint4* __restrict__ d, size_t n16) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n16; q += stride) d[q] = s[q];
}
__global__ __launch_bounds__(256) void copy_u8(const unsigned char* __restrict__ s, unsigned char* __restrict__ d,
                                              size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) d[q] = s[q];
}

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of the no-allocation entry point (sbp_apply_dot_noalloc):
// the stencil pass's block partials and the per-problem results live in a
// file-scope device array, so that call needs no device memory of its own.
// Each GPU holds its own copy until the process ends.  Host code never uses
// the array's name as an address: its device address is resolved once per
// device with hipGetSymbolAddress (scratch_ptr below) and only that pointer
// is passed to kernels.
// ---------------------------------------------------------------------------
constexpr size_t kScratchDoubles = size_t(1) << 17;   // 1 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved on first use).
int scratch_ptr(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return SBP_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return SBP_OK;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* sbp_version(void) { return "sc_2022_amd 0.2 (gfx950)"; }

const char* sbp_status_string(int s) {
    switch (s) {
        case SBP_OK: return "ok";
        case SBP_ERR_INVALID_SIZE: return "invalid size";
        case SBP_ERR_INVALID_ORDER: return "invalid SBP order (p must be 2, 4, 6, 8 or 10; 2, 4 or 6 with variable coefficients)";
        case SBP_ERR_INVALID_ARG: return "invalid argument";
        case SBP_ERR_HIP: return "HIP runtime error";
        case SBP_ERR_NONFINITE: return "non-finite value in residual or inner product";
        case SBP_ERR_BREAKDOWN: return "breakdown: p'Ap <= 0";
        case SBP_ERR_NOT_CONVERGED: return "not converged within maxiter";
        case SBP_ERR_NO_DEVICE: return "no HIP device";
        case SBP_ERR_SCRATCH: return "problem too large for the library's static scratch array (no-allocation entry point)";
        case SBP_ERR_VENDOR: return "rocSPARSE call failed";
        default: return "unknown status";
    }
}

void sbp_default_params(sbp_params* prm) {
    if (!prm) return;
    std::memset(prm, 0, sizeof(*prm));
    prm->nu_pre = 2;
    prm->nu_post = 2;
    prm->omega = 2.0 / 3.0;
    prm->coarse_n = 17;
    prm->restrict_scale = 4.0;
    prm->smoother = 0;
    prm->richardson_omega = 0.15;
    prm->poll_every = 0;
    prm->inner_rtol = 1e-5;
    prm->inner_maxiter = 50;
    prm->max_outer = 10;
    prm->tile_i = 0;
    prm->tile_j = 0;
    prm->use_graph = 1;
    prm->ir_rtol = 0.0;
    prm->bottom_n = 0;
    prm->fuse_n = 1025;
}

int sbp_destroy(sbp_handle h) {
    if (!h) return SBP_ERR_INVALID_ARG;
    for (int l = 0; l < kMaxLevels; ++l) {
        Level& L = h->L[l];
        if (L.dtab) (void)hipFree(L.dtab);
        for (int q = 0; q < 2; ++q) {
            if (L.b[q] && (l > 0 || h->ragged)) (void)hipFree(L.b[q]);
            if (L.xa[q]) (void)hipFree(L.xa[q]);
            if (L.xb[q]) (void)hipFree(L.xb[q]);
            if (L.xo[q]) (void)hipFree(L.xo[q]);
            if (L.r[q]) (void)hipFree(L.r[q]);
        }
    }
    drop_graphs(h);
    if (h->vs_x) (void)rocsparse_destroy_dnvec_descr(h->vs_x);
    if (h->vs_y) (void)rocsparse_destroy_dnvec_descr(h->vs_y);
    if (h->vs_mat) (void)rocsparse_destroy_spmat_descr(h->vs_mat);
    if (h->vs) (void)rocsparse_destroy_handle(h->vs);
    if (h->vs_buf) (void)hipFree(h->vs_buf);
    if (h->cap) (void)hipStreamDestroy(h->cap);
    if (h->pev) (void)hipEventDestroy(h->pev);
    if (h->child) (void)sbp_destroy(h->child);
    if (h->cexact) (void)sbp_destroy(h->cexact);
    if (h->vtab) (void)hipFree(h->vtab);
    if (h->vband) (void)hipFree(h->vband);
    for (size_t l = 0; l < h->vl.size(); ++l) {
        auto& V = h->vl[l];
        if (l > 0) {
            void* own[] = {V.dtab, V.band, V.b, V.x, V.t, V.b32, V.x32, V.t32};
            for (void* q : own)
                if (q) (void)hipFree(q);
        }
    }
    void* ptrs[] = {h->cbuf, h->xc, h->w, h->lam, h->Ainv, h->r, h->pv, h->pv2, h->Ap, h->z, h->x, h->r32, h->p32, h->p32b, h->Ap32,
                    h->z32, h->d32, h->partial, h->sc, h->sc_in, h->dtmp, h->dabs};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (h->sc_host) (void)hipHostFree(h->sc_host);
    if (h->habs) (void)hipHostFree(h->habs);
    delete h;
    return SBP_OK;
}

namespace {
int create_impl(int N, int p, int nprob, const sbp_params* prm, bool no_mg, sbp_handle* out,
                const int* sides = nullptr);
}

int sbp_create_batch(int N, int p, int nprob, const sbp_params* prm, sbp_handle* out) {
    return create_impl(N, p, nprob, prm, false, out);
}

namespace {
// sides (nullable): a ragged handle, problem q of side sides[q] (non-increasing,
// sides[0] = N, each a level side of N's hierarchy).
int create_impl(int N, int p, int nprob, const sbp_params* prm, bool no_mg, sbp_handle* out, const int* sides) {
    if (!out) return SBP_ERR_INVALID_ARG;
    *out = nullptr;
    if (p != 2 && p != 4 && p != 6 && p != 8 && p != 10) return SBP_ERR_INVALID_ORDER;
    {
        OpTab t0;
        if (!build_optab(p, N, &t0, nullptr)) return SBP_ERR_INVALID_SIZE;
        if (N > 46340) return SBP_ERR_INVALID_SIZE;   // N*N must fit a 32-bit block index space
    }
    if (nprob < 1 || nprob > 65535) return SBP_ERR_INVALID_ARG;   // problems ride grid.y of the elementwise kernels
    if ((long long)N * N * nprob > (1ll << 40)) return SBP_ERR_INVALID_SIZE;
    if (prm && (prm->nu_pre < 0 || prm->nu_post < 0 || prm->poll_every < 0 || prm->max_outer < 1 ||
                prm->inner_maxiter < 1 || !(prm->omega > 0) || !(prm->restrict_scale > 0) ||
                (prm->smoother != 0 && prm->smoother != 1) || prm->bottom_n < 0 ||
                prm->fuse_n < 0 || (prm->drop_coarse_rhs != 0 && prm->drop_coarse_rhs != 1)))
        return SBP_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SBP_ERR_NO_DEVICE;
    sbp_handle h = new (std::nothrow) sbp_handle_s();
    if (!h) return SBP_ERR_HIP;
    h->N = N;
    h->p = p;
    h->nprob = nprob;
    h->ragged = sides != nullptr;
    if (prm) h->prm = *prm; else sbp_default_params(&h->prm);
    // p = 6, 8, 10 closures need at least 2 C = 18 / 24 / 32 points per side
    // on every level: their coarsest grid is 33 x 33 at the least.
    if (p >= 6 && h->prm.coarse_n < 33) h->prm.coarse_n = 33;
    if (!build_optab(p, N, &h->L[0].tb, &h->rt)) { delete h; return SBP_ERR_INVALID_SIZE; }
    h->L[0].N = N;
    h->L[0].n = (size_t)N * N;
    // Multigrid hierarchy when N = 2^k+1 and coarse_n = 2^c+1 <= N.
    int nlev = 1;
    const int cn = h->prm.coarse_n;
    auto pow2p1 = [](int v) { return v >= 3 && ((v - 1) & (v - 2)) == 0; };
    if (!no_mg && pow2p1(N) && pow2p1(cn) && cn <= N && (cn <= 33 || nprob == 1)) {
        int n = N;
        while (n > cn && nlev < kMaxLevels) {
            n = (n + 1) / 2;
            Level& L = h->L[nlev];
            L.N = n;
            L.n = (size_t)n * n;
            if (!build_optab(p, n, &L.tb, nullptr)) break;
            ++nlev;
        }
        if (h->L[nlev - 1].N == cn) h->mg = true;
        else nlev = 1;
    }
    h->nlev = nlev;
    if (h->ragged) {
        // every side must be one of the hierarchy's levels
        if (!h->mg || h->L[nlev - 1].N > 33) { delete h; return SBP_ERR_INVALID_SIZE; }
        size_t tot = 0;
        int q = 0;
        for (int l = 0; l < nlev; ++l) {
            h->j0[l] = q;
            h->koff[l] = tot;
            while (q < nprob && sides[q] == h->L[l].N) { tot += h->L[l].n; ++q; }
            h->np[l] = q;
        }
        if (q != nprob) { delete h; return SBP_ERR_INVALID_SIZE; }   // a side off the hierarchy
        h->rtot = tot;
    } else {
        for (int l = 0; l < nlev; ++l) {
            h->np[l] = nprob;
            h->j0[l] = l == 0 ? 0 : nprob;
        }
    }
    // Bottom kernels: the levels from the first of side <= bottom_n down to
    // the coarsest, when they fit the kernels' LDS pool and prefetch registers.
    if (h->mg && p <= 4 && !h->ragged && !h->prm.drop_coarse_rhs) {
        const int bn = std::min(h->prm.bottom_n, kBottomN);
        int l0 = -1;
        for (int l = 0; l < nlev - 1; ++l)
            if (h->L[l].N <= bn) { l0 = l; break; }
        if (l0 >= 0 && nlev - 1 - l0 <= kBottomMaxLevels) {
            const long long nc = (long long)h->L[nlev - 1].n;
            long long down = 0, up = nc, pre = nc;
            bool ok = (long long)h->L[l0].n <= (long long)kBottomDownPre * kBottomThreads;
            for (int l = l0; l < nlev - 1; ++l) {
                const long long n = (long long)h->L[l].n;
                ok = ok && h->L[l].N > 2 * h->L[l].tb.R;
                down += 3 * n;
                up += (l == l0 ? 3 : 4) * n;
                pre += 2 * n;
            }
            if (ok && down <= kBottomPool && up <= kBottomPool && pre <= (long long)kBottomUpPre * kBottomThreads)
                h->bottom_l0 = l0;
        }
    }
    for (int l = 0; l < nlev; ++l) {
        Level& L = h->L[l];
        for (int q = 0; q < 2; ++q) {
            if (l == 0 && h->prm.tile_i > 0) {
                if (!tile_ok(p, h->prm.tile_i, h->prm.tile_j)) { delete h; return SBP_ERR_INVALID_ARG; }
                L.ti[q] = h->prm.tile_i;
                L.tj[q] = h->prm.tile_j;
                L.ta[q] = 0;
            } else {
                default_tile(L.N, p, q, &L.ti[q], &L.tj[q], &L.ta[q]);
            }
            // fused two-sweep blocks: small blocks keep the per-block chain
            // short; from 257^2 points up 64 x 8 (half the blocks; 257^2: 8.0
            // against 8.9 us per pre-smoothing launch, 513^2: 10.2 against 11.7)
            L.fti[q] = 64;
            L.ftj[q] = (long long)L.n * h->np[l] >= 257ll * 257 ? 8 : 4;
        }
    }
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess)
        h->clock_khz = (double)khz;
    h->dev = dev;
    int rc = SBP_OK;
    do {
        for (int l = 0; l < nlev && rc == SBP_OK; ++l) {
            Level& L = h->L[l];
            if (hipMalloc(&L.dtab, sizeof(OpTab)) != hipSuccess) { rc = SBP_ERR_HIP; break; }
            L.tb.dev = L.dtab;
            if (hipMemcpy(L.dtab, &L.tb, sizeof(OpTab), hipMemcpyHostToDevice) != hipSuccess) rc = SBP_ERR_HIP;
        }
        if (rc != SBP_OK) break;
        const size_t n = h->ragged ? h->rtot : h->L[0].n * (size_t)nprob;
        if (hipMalloc(&h->r, n * 8) != hipSuccess || hipMalloc(&h->pv, n * 8) != hipSuccess ||
            hipMalloc(&h->Ap, n * 8) != hipSuccess || hipMalloc(&h->z, n * 8) != hipSuccess ||
            hipMalloc(&h->x, n * 8) != hipSuccess || hipMalloc(&h->pv2, n * 8) != hipSuccess) { rc = SBP_ERR_HIP; break; }
        // Partial slots per problem: the finest pass at any compiled tile shape,
        // and the elementwise kernels (two per block in the refinement update).
        h->npp = std::max<size_t>(pass_blocks(N, 64, 4), 2 * 4096) + 64;
        size_t nslots = h->npp * nprob;
        if (h->ragged) {   // one region per level: np[l] problems x npl[l] slots
            nslots = 0;
            for (int l = 0; l < nlev; ++l) {
                h->npl[l] = std::max<size_t>(pass_blocks(h->L[l].N, 64, 4), 2 * 4096) + 64;
                h->roff[l] = nslots;
                nslots += h->npl[l] * h->np[l];
            }
        }
        if (hipMalloc(&h->partial, nslots * sizeof(double)) != hipSuccess) { rc = SBP_ERR_HIP; break; }
        if (hipMalloc(&h->sc, sizeof(Scalars) * nprob) != hipSuccess ||
            hipMalloc(&h->sc_in, sizeof(Scalars) * nprob) != hipSuccess ||
            hipMalloc(&h->dtmp, nprob * sizeof(double)) != hipSuccess ||
            hipMalloc(&h->dabs, nprob * sizeof(double)) != hipSuccess) { rc = SBP_ERR_HIP; break; }
        if (hipHostMalloc((void**)&h->sc_host, sizeof(Scalars) * nprob, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&h->habs, sizeof(double) * nprob, hipHostMallocDefault) != hipSuccess) {
            rc = SBP_ERR_HIP;
            break;
        }
        if (hipMemset(h->sc, 0, sizeof(Scalars) * nprob) != hipSuccess ||
            hipMemset(h->sc_in, 0, sizeof(Scalars) * nprob) != hipSuccess) { rc = SBP_ERR_HIP; break; }
        if (h->mg) {
            if ((rc = alloc_level_buffers(h, 0)) != SBP_OK) break;
            // Coarsest level: FP64 inverse of the rediscretised operator via Cholesky.
            const Level& Lc = h->L[nlev - 1];
            if (Lc.N > 33) {   // nested exact solve instead of a dense inverse
                h->coarse_nested = true;
                h->nc = (int)Lc.n;
                break;
            }
            std::vector<double> A;
            dense_operator(Lc.tb, A);
            const int nc = (int)Lc.n;
            if (!spd_inverse(A, nc)) {
                // Not positive definite (p = 8 with the reference's penalty):
                // no multigrid; the apply and plain CG stay available.
                h->mg = false;
            } else {
                h->nc = nc;
                if (hipMalloc(&h->Ainv, (size_t)nc * nc * sizeof(double)) != hipSuccess) { rc = SBP_ERR_HIP; break; }
                if (hipMemcpy(h->Ainv, A.data(), (size_t)nc * nc * sizeof(double), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                    rc = SBP_ERR_HIP;
                    break;
                }
            }
        }
    } while (0);
    if (rc != SBP_OK) { sbp_destroy(h); return rc; }
    *out = h;
    return SBP_OK;
}
}  // namespace

int sbp_create(int N, int p, const sbp_params* prm, sbp_handle* out) { return sbp_create_batch(N, p, 1, prm, out); }

int sbp_create_ragged(int p, int nprob, const int* sides, const sbp_params* prm, sbp_handle* out) {
    if (!out || !sides) return SBP_ERR_INVALID_ARG;
    *out = nullptr;
    if (nprob < 1 || nprob > 65535) return SBP_ERR_INVALID_ARG;
    long long tot = 0;
    for (int q = 0; q < nprob; ++q) {
        if (sides[q] < 3 || (q > 0 && sides[q] > sides[q - 1])) return SBP_ERR_INVALID_ARG;   // sorted, largest first
        tot += (long long)sides[q] * sides[q];
    }
    if (tot > (1ll << 40)) return SBP_ERR_INVALID_SIZE;
    return create_impl(sides[0], p, nprob, prm, false, out, sides);
}

namespace {
// Multigrid hierarchy of a variable-coefficient handle: levels of side
// (N+1)/2, ... down to the coarse side (coarse_n, at least 33 for p = 6,
// whose closures need 26 points), each the operator rediscretised with the
// profiles injected (b_l[i] = b[2^l i]); the coarsest level's dense inverse.
// Leaves vmg false (apply and CG only) when N is not 2^k + 1, the coarse side
// is not reachable or above 33, or the coarsest operator is not SPD.
int var_hierarchy(sbp_handle h, const VarTab& t0, const double* bx, const double* by) {
    const int N = h->N, p = h->p;
    auto pow2p1 = [](int v) { return v >= 3 && ((v - 1) & (v - 2)) == 0; };
    int cn = h->prm.coarse_n;
    if (p == 6 && cn < 33) cn = 33;
    if (!pow2p1(N) || !pow2p1(cn) || cn > 33 || cn > N) return SBP_OK;
    std::vector<int> sides{N};
    while (sides.back() > cn) sides.push_back((sides.back() + 1) / 2);
    if (sides.back() != cn) return SBP_OK;
    const int nlev = (int)sides.size();
    h->vl.resize(nlev);
    std::vector<double> bxl(bx, bx + N), byl(by, by + N), kb, mb, A;
    for (int l = 0; l < nlev; ++l) {
        auto& V = h->vl[l];
        V.N = sides[l];
        V.n = (size_t)V.N * V.N;
        VarTab t = t0;
        if (l == 0) {
            V.dtab = h->vtab;
            V.band = h->vband;
            V.x = h->z;
            V.t = h->Ap;   // free while a Krylov iteration runs its V-cycle (p ping-pongs over pv / pv2)
        } else {
            std::vector<double> cx(V.N), cy(V.N);
            for (int k = 0; k < V.N; ++k) { cx[k] = bxl[2 * k]; cy[k] = byl[2 * k]; }
            bxl.swap(cx);
            byl.swap(cy);
            OpTab ot;
            if (!build_optab(p, V.N, &ot, nullptr)) return SBP_OK;
            t = VarTab{};
            if (!build_var_axis(p, V.N, bxl.data(), true, &t.k, kb) || !build_var_axis(p, V.N, byl.data(), false, &t.m, mb))
                return SBP_OK;
            t.N = V.N;
            t.H = p / 2;
            t.R = var_R(p);
            t.C = var_C(p);
            t.RH = ot.RH;
            for (int r = 0; r < kMaxR; ++r) t.hw[r] = ot.hw[r];
            const size_t nb = kb.size();
            HIPCHK(hipMalloc(&V.band, 2 * nb * sizeof(double)));
            HIPCHK(hipMalloc(&V.dtab, sizeof(VarTab)));
            t.k.band = V.band;
            t.m.band = V.band + nb;
            HIPCHK(hipMemcpy(V.band, kb.data(), nb * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(V.band + nb, mb.data(), nb * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(V.dtab, &t, sizeof(VarTab), hipMemcpyHostToDevice));
            HIPCHK(hipMalloc(&V.b, V.n * sizeof(double)));
            HIPCHK(hipMalloc(&V.x, V.n * sizeof(double)));
            HIPCHK(hipMalloc(&V.t, V.n * sizeof(double)));
        }
        if (l == nlev - 1) {
            if (l == 0) return SBP_OK;   // N is the coarse side: nothing to precondition
            VarTab th = t;                // host bands for the dense build
            dense_operator_var(th, kb.data(), mb.data(), A);
            const int nc = (int)V.n;
            if (!spd_inverse(A, nc)) return SBP_OK;
            HIPCHK(hipMalloc(&h->Ainv, (size_t)nc * nc * sizeof(double)));
            HIPCHK(hipMemcpy(h->Ainv, A.data(), (size_t)nc * nc * sizeof(double), hipMemcpyHostToDevice));
            h->nc = nc;
        }
    }
    HIPCHK(hipStreamSynchronize(h->st));
    h->vmg = true;
    return SBP_OK;
}
}  // namespace

int sbp_create_var(int N, int p, const double* bx, const double* by, const sbp_params* prm, sbp_handle* out) {
    if (!out) return SBP_ERR_INVALID_ARG;
    *out = nullptr;
    if (p != 2 && p != 4 && p != 6) return SBP_ERR_INVALID_ORDER;
    if (!bx || !by) return SBP_ERR_INVALID_ARG;
    for (int k = 0; k < N; ++k)
        if (!(bx[k] > 0.0) || !(by[k] > 0.0) || !std::isfinite(bx[k]) || !std::isfinite(by[k]))
            return SBP_ERR_INVALID_ARG;
    VarTab t{};
    std::vector<double> bk, bm;
    if (N < 2 || !build_var_axis(p, N, bx, true, &t.k, bk) || !build_var_axis(p, N, by, false, &t.m, bm))
        return SBP_ERR_INVALID_SIZE;
    sbp_handle h = nullptr;
    STCHK(create_impl(N, p, 1, prm, true, &h));
    const OpTab& tb = h->L[0].tb;
    t.N = N;
    t.H = p / 2;
    t.R = var_R(p);
    t.C = var_C(p);
    t.RH = tb.RH;
    for (int r = 0; r < kMaxR; ++r) t.hw[r] = tb.hw[r];
    const size_t nb = bk.size();
    int rc = SBP_OK;
    if (hipMalloc(&h->vband, 2 * nb * sizeof(double)) != hipSuccess ||
        hipMalloc(&h->vtab, sizeof(VarTab)) != hipSuccess) {
        rc = SBP_ERR_HIP;
    } else {
        t.k.band = h->vband;
        t.m.band = h->vband + nb;
        if (hipMemcpy(h->vband, bk.data(), nb * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->vband + nb, bm.data(), nb * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->vtab, &t, sizeof(VarTab), hipMemcpyHostToDevice) != hipSuccess)
            rc = SBP_ERR_HIP;
    }
    if (rc != SBP_OK) { sbp_destroy(h); return rc; }
    h->var = true;
    h->rt.bW = bx[0];
    h->rt.bE = bx[N - 1];
    if ((rc = var_hierarchy(h, t, bx, by)) != SBP_OK) { sbp_destroy(h); return rc; }
    *out = h;
    return SBP_OK;
}

int sbp_set_stream(sbp_handle h, void* s) {
    if (!h) return SBP_ERR_INVALID_ARG;
    h->st = (hipStream_t)s;
    return SBP_OK;
}

int sbp_get_info(sbp_handle h, int* N, int* p, int* nlevels) {
    if (!h) return SBP_ERR_INVALID_ARG;
    if (N) *N = h->N;
    if (p) *p = h->p;
    if (nlevels) *nlevels = h->var ? (h->vmg ? (int)h->vl.size() : 0) : (h->mg ? h->nlev : 0);
    return SBP_OK;
}

int sbp_get_batch(sbp_handle h, int* nprob) {
    if (!h || !nprob) return SBP_ERR_INVALID_ARG;
    *nprob = h->nprob;
    return SBP_OK;
}

int sbp_set_tiles(sbp_handle h, int level, int dtype, int ti, int tj) {
    if (!h || level < 0 || level >= h->nlev) return SBP_ERR_INVALID_ARG;
    const bool fused = (dtype & SBP_FUSED) != 0;
    const int ta = (dtype & SBP_ALIGNED) ? 1 : 0;
    const int d = dtype & ~(SBP_FUSED | SBP_ALIGNED);
    if ((d != SBP_F64 && d != SBP_F32) || (fused && ta) ||
        (fused ? !fused_shape_ok(ti, tj, h->p) : !tile_ok(h->p, ti, tj, ta)))
        return SBP_ERR_INVALID_ARG;
    Level& L = h->L[level];
    int* pti = fused ? &L.fti[d] : &L.ti[d];
    int* ptj = fused ? &L.ftj[d] : &L.tj[d];
    if (*pti != ti || *ptj != tj || (!fused && L.ta[d] != ta)) drop_graphs(h);   // launch shapes change
    *pti = ti;
    *ptj = tj;
    if (!fused) L.ta[d] = ta;
    return SBP_OK;
}

namespace {
extern "C++" {   // templates inside the C-linkage block
// One pass of the variable-coefficient operator on level `level` with
// operand policy IN and output policy OUT (sbp_var.h); *nparts receives the
// number of block partials.
template <typename T, class IN, class OUT>
int launch_var_pass(sbp_handle h, int level, const IN& in, const OUT& out, const Scalars* sc, int* nparts) {
    const int N = level == 0 ? h->N : h->vl[level].N;
    const VarTab* tab = level == 0 ? h->vtab : h->vl[level].dtab;
    const double* kb = level == 0 ? h->vband : h->vl[level].band;   // k band, then m band
    const double* mb = kb + (size_t)(h->p + 1) * N;
    const int nbx = (N + kVarThreads - 1) / kVarThreads, nby = (N + kVarRows - 1) / kVarRows;
    if ((size_t)nbx * nby > h->npp) return SBP_ERR_INVALID_SIZE;
    const dim3 g(nbx * nby);
    if (h->p == 2)
        hipLaunchKernelGGL((var_pass<2, T, IN, OUT>), g, dim3(kVarThreads), 0, h->st, in, out, tab, kb, mb, sc, h->partial, nbx, nby);
    else if (h->p == 4)
        hipLaunchKernelGGL((var_pass<4, T, IN, OUT>), g, dim3(kVarThreads), 0, h->st, in, out, tab, kb, mb, sc, h->partial, nbx, nby);
    else
        hipLaunchKernelGGL((var_pass<6, T, IN, OUT>), g, dim3(kVarThreads), 0, h->st, in, out, tab, kb, mb, sc, h->partial, nbx, nby);
    HIPCHK(hipGetLastError());
    if (nparts) *nparts = nbx * nby;
    return SBP_OK;
}

// The plain-operand passes by mode: 0 out = sign A u; 1 out = A u with
// partials of aux' out; 2 out = aux - A u with partials of out' out; 3 the
// damped-Jacobi sweep out = u + w (aux - A u) / diag(A); 4 the same with
// partials of aux' out.
template <typename T = double>
int launch_var(sbp_handle h, int mode, const T* u, T* out, const T* aux, double sign,
               const Scalars* sc, int* nparts, int level = 0, double omega = 0.0) {
    const VarIn<T> in{u};
    switch (mode) {
    case 0: return launch_var_pass<T>(h, level, in, VarOutStore<T>{out, sign}, sc, nparts);
    case 1: return launch_var_pass<T>(h, level, in, VarOutStoreDot<T>{out, aux}, sc, nparts);
    case 2: return launch_var_pass<T>(h, level, in, VarOutResid<T>{aux, out}, sc, nparts);
    case 3: return launch_var_pass<T>(h, level, in, VarOutSmooth<T>{aux, out, omega}, sc, nparts);
    default: return launch_var_pass<T>(h, level, in, VarOutSmoothDot<T>{aux, out, omega}, sc, nparts);
    }
}

// rc = restrict_scale * FW(b - A u) from level l to l + 1, the fine residual on chip.
template <typename T>
int launch_var_rr(sbp_handle h, int l, const T* u, const T* b, T* rc, const Scalars* sc) {
    const auto& V = h->vl[l];
    const int Nc = h->vl[l + 1].N;
    constexpr int CI = kVarThreads / 2 - 1, CJ = kVarRows / 2;
    const int nbx = (Nc + CI - 1) / CI, nby = (Nc + CJ - 1) / CJ;
    const dim3 g(nbx * nby);
    const double sc_r = h->prm.restrict_scale;
    const double* kb = V.band;
    const double* mb = kb + (size_t)(h->p + 1) * V.N;
    if (h->p == 2)
        hipLaunchKernelGGL((var_resid_restrict<2, T>), g, dim3(kVarThreads), 0, h->st, u, b, rc, sc_r, V.dtab, kb, mb, Nc, nbx, nby, sc);
    else if (h->p == 4)
        hipLaunchKernelGGL((var_resid_restrict<4, T>), g, dim3(kVarThreads), 0, h->st, u, b, rc, sc_r, V.dtab, kb, mb, Nc, nbx, nby, sc);
    else
        hipLaunchKernelGGL((var_resid_restrict<6, T>), g, dim3(kVarThreads), 0, h->st, u, b, rc, sc_r, V.dtab, kb, mb, Nc, nbx, nby, sc);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

// One V-cycle x ~ A_l^-1 b on level l of the variable-coefficient hierarchy
// from a zero initial guess: nu_pre damped-Jacobi sweeps (the first one
// x = w D^-1 b), residual, full-weighting restriction (times restrict_scale),
// the coarser cycle, bilinear prolongation, nu_post sweeps; the coarsest
// level is the dense inverse.  Sweeps ping-pong between x and t, starting so
// that the last one lands in x; the residual goes to whichever is free.
// T is the storage type of the level vectors (float: the b32 / x32 / t32
// copies, level 0 sweeping through Ap32); the kernels compute in FP64.
template <typename T>
int var_vcycle(sbp_handle h, int l, const T* b, T* x, const Scalars* sc, int* np_rz) {
    constexpr bool f32 = std::is_same<T, float>::value;
    auto& V = h->vl[l];
    const int nlev = (int)h->vl.size();
    hipStream_t st = h->st;
    if (l == nlev - 1) {
        hipLaunchKernelGGL((coarse_solve<T>), dim3((h->nc + kCoarseRows - 1) / kCoarseRows, 1), dim3(256), 0, st,
                           (const double*)h->Ainv, b, x, h->nc, sc);
        HIPCHK(hipGetLastError());
        return SBP_OK;
    }
    const int nu1 = h->prm.nu_pre, nu2 = h->prm.nu_post;
    const double om = h->prm.omega;
    T* tl;   // this level's sweep buffer
    if constexpr (f32) tl = l == 0 ? h->Ap32 : V.t32;
    else tl = V.t;
    auto& C = h->vl[l + 1];
    T* cb;
    T* cx;
    if constexpr (f32) { cb = C.b32; cx = C.x32; }
    else { cb = C.b; cx = C.x; }
    // Small levels (at most fuse_n^2 points) with 2 + 2 sweeps: the two
    // temporally blocked kernels, x2 = S(S(0)) with the restricted residual,
    // then x = S(S(x2 + P e)) (and the b'x partials).
    if (nu1 == 2 && nu2 == 2 && (long long)V.n <= (long long)h->prm.fuse_n * h->prm.fuse_n) {
        const int ti = 64, tj = V.n >= (size_t)257 * 257 ? 8 : 4;   // as the constant-coefficient levels
        const size_t nb = (size_t)(h->p + 1) * V.N;
        const VarFusedArgs a{V.N, V.dtab, V.band, V.band + nb, om, h->prm.restrict_scale, sc ? &sc->done : nullptr};
        if (!var_fused_shape_ok(ti, tj)) return SBP_ERR_INVALID_ARG;
        STCHK(launch_var_smooth2_restrict<T>(h->p, ti, tj, a, b, tl, cb, st));
        STCHK(var_vcycle<T>(h, l + 1, (const T*)cb, cx, sc, nullptr));
        STCHK(launch_var_prolong_smooth2<T>(h->p, ti, tj, a, (const T*)tl, (const T*)cx, b, x, np_rz ? b : nullptr,
                                            h->partial, st));
        if (np_rz) *np_rz = fused_post_blocks(V.N, ti, tj);
        return SBP_OK;
    }
    T* bufs[2] = {x, tl};
    // Buffer writes: nu1 - 1 for the pre-smoothing (the first two sweeps are
    // one pass; one write for nu1 = 0 or 1) and nu2 for the post-smoothing
    // (the prolongation is folded into its first sweep).  The first write goes
    // to the buffer that makes the last one land in x.
    const int writes = (nu1 >= 2 ? nu1 - 1 : 1) + nu2;
    int cur = (writes - 1) % 2;
    if (nu1 >= 2) {
        // x2 = S(S(0)): the first sweep x1 = w D^-1 b formed on the fly
        STCHK(launch_var_pass<T>(h, l, VarInSmooth0<T>{b, om}, VarOutSmooth<T>{b, bufs[cur], om}, sc, nullptr));
        for (int s = 2; s < nu1; ++s) {
            STCHK(launch_var(h, 3, bufs[cur], bufs[1 - cur], b, 1.0, sc, nullptr, l, om));
            cur = 1 - cur;
        }
    } else if (nu1 == 1) {
        hipLaunchKernelGGL((var_jacobi0<T>), dim3((V.N + 255) / 256, V.N), dim3(256), 0, st, bufs[cur], b,
                           (const VarTab*)V.dtab, om, sc);
        HIPCHK(hipGetLastError());
    } else {
        hipLaunchKernelGGL((fill<T, kEwThreads>), dim3(ew_blocks(V.n), 1), dim3(kEwThreads), 0, st, bufs[cur], T(0),
                           V.n, sc);
        HIPCHK(hipGetLastError());
    }
    // residual and full-weighting restriction in one pass (the fine residual stays on chip)
    STCHK(launch_var_rr<T>(h, l, (const T*)bufs[cur], b, cb, sc));
    STCHK(var_vcycle<T>(h, l + 1, (const T*)cb, cx, sc, nullptr));
    if (nu2 == 0) {
        hipLaunchKernelGGL((prolong_add<T>), dim3((V.N + 255) / 256, V.N, 1), dim3(256), 0, st, (const T*)cx,
                           bufs[cur], V.N, C.N, sc);
        HIPCHK(hipGetLastError());
    } else {
        // the first post-smoothing sweep on x + P e; with np_rz the last sweep
        // also leaves the block partials of b'x
        const VarInProlong<T> inp{bufs[cur], cx, C.N};
        if (nu2 == 1 && np_rz)
            STCHK(launch_var_pass<T>(h, l, inp, VarOutSmoothDot<T>{b, bufs[1 - cur], om}, sc, np_rz));
        else
            STCHK(launch_var_pass<T>(h, l, inp, VarOutSmooth<T>{b, bufs[1 - cur], om}, sc, nullptr));
        cur = 1 - cur;
        for (int s = 1; s < nu2; ++s) {
            const bool last = s == nu2 - 1 && np_rz;
            STCHK(launch_var(h, last ? 4 : 3, bufs[cur], bufs[1 - cur], b, 1.0, sc, last ? np_rz : nullptr, l, om));
            cur = 1 - cur;
        }
    }
    if (bufs[cur] != x) {
        hipLaunchKernelGGL((copy_vec<T, kEwThreads>), dim3(ew_blocks(V.n), 1), dim3(kEwThreads), 0, st,
                           (const T*)bufs[cur], x, V.n, sc);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

// FP32 copies of the variable-coefficient V-cycle's level buffers (made on
// first use; level 0 borrows r32 / z32 / Ap32).
int ensure_var32(sbp_handle h) {
    STCHK(ensure32(h));
    for (size_t l = 1; l < h->vl.size(); ++l) {
        auto& V = h->vl[l];
        if (V.b32) continue;
        HIPCHK(hipMalloc(&V.b32, V.n * sizeof(float)));
        HIPCHK(hipMalloc(&V.x32, V.n * sizeof(float)));
        HIPCHK(hipMalloc(&V.t32, V.n * sizeof(float)));
    }
    return SBP_OK;
}

// FP64 MG-PCG on the variable-coefficient operator with the V-cycle on FP32
// copies of the level vectors (sbp_mgpcg_mixed on such a handle): r rounded
// to FP32 by the residual update, z read in FP32 by r'z and the
// search-direction update; the Krylov vectors, the operator products and the
// inner products stay FP64.  Stops on ||r||_2 < abstol.  (The FP64 V-cycle
// runs in the shared Krylov loop, krylov64.)
int var_pcg_f32vc(sbp_handle h, const double* b, double* x, double abstol, int maxiter) {
    hipStream_t st = h->st;
    const size_t n = h->L[0].n;
    Scalars* sc = h->sc;
    STCHK(put_abstol(h, &abstol));
    hipLaunchKernelGGL(stamp_start, dim3(1), dim3(256), 0, st, sc, 1);
    int np = 0;
    STCHK(launch_var(h, 2, x, h->r, b, 1.0, nullptr, &np));
    hipLaunchKernelGGL(krylov_init, dim3(1), dim3(kRedThreads), 0, st, h->partial, np, sc, (const double*)h->dabs,
                       (double*)nullptr, maxiter);
    HIPCHK(hipGetLastError());
    const dim3 g(ew_blocks(n), 1);
    // z = M r with the r'z partials (g.x of them); r32 holds r rounded to
    // FP32 once the residual update has written it
    bool r32_ready = false;
    auto precondition = [&]() -> int {
        if (!r32_ready)
            hipLaunchKernelGGL((cast_vec<double, float, kEwThreads>), g, dim3(kEwThreads), 0, st,
                               (const double*)h->r, h->r32, n);
        STCHK(var_vcycle<float>(h, 0, (const float*)h->r32, h->z32, sc, nullptr));
        hipLaunchKernelGGL((dot_partial<double, kEwThreads, float>), g, dim3(kEwThreads), 0, st,
                           (const double*)h->r, (const float*)h->z32, n, h->partial);
        HIPCHK(hipGetLastError());
        return SBP_OK;
    };
    STCHK(precondition());
    hipLaunchKernelGGL(krylov_set_rz, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc);
    hipLaunchKernelGGL((cast_vec<float, double, kEwThreads>), g, dim3(kEwThreads), 0, st, (const float*)h->z32,
                       h->pv, n);
    HIPCHK(hipGetLastError());
    r32_ready = true;
    STCHK(krylov_iterate(h, sc, maxiter, [&](int) -> int {
        STCHK(launch_var(h, 1, h->pv, h->Ap, h->pv, 1.0, sc, &np));
        hipLaunchKernelGGL(krylov_alpha, dim3(1), dim3(kRedThreads), 0, st, h->partial, np, sc, 1);
        hipLaunchKernelGGL((update_xr_r32<kEwThreads>), g, dim3(kEwThreads), 0, st, x, h->r, h->r32,
                           (const double*)h->pv, (const double*)h->Ap, (const Scalars*)sc, n, h->partial);
        hipLaunchKernelGGL(krylov_check, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc, 0);
        HIPCHK(hipGetLastError());
        STCHK(precondition());
        hipLaunchKernelGGL(krylov_beta, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc);
        hipLaunchKernelGGL((update_p<double, kEwThreads, float>), g, dim3(kEwThreads), 0, st, h->pv,
                           (const float*)h->z32, (const Scalars*)sc, n);
        HIPCHK(hipGetLastError());
        return SBP_OK;
    }));
    return poll(h, sc);
}
}  // extern "C++"
}  // namespace

int sbp_apply(sbp_handle h, int level, const void* u, void* Au, int sign, int dtype) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !u || !Au || (sign != 1 && sign != -1) || level < 0 || level >= h->nlev) return SBP_ERR_INVALID_ARG;
    if (h->var) {   // variable coefficients: FP64, level 0
        if (dtype != SBP_F64 || level != 0) return SBP_ERR_INVALID_ARG;
        return launch_var<double>(h, 0, (const double*)u, (double*)Au, nullptr, (double)sign, nullptr, nullptr);
    }
    if (dtype == SBP_F64)
        return launch_pass<double>(h, level, InPlain<double>{(const double*)u},
                                   OutStore<double>{(double*)Au, (double)sign}, nullptr, nullptr, nullptr);
    if (dtype == SBP_F32)
        return launch_pass<float>(h, level, InPlain<float>{(const float*)u}, OutStore<float>{(float*)Au, (float)sign},
                                  nullptr, nullptr, nullptr);
    return SBP_ERR_INVALID_ARG;
}

int sbp_apply_dot(sbp_handle h, const double* u, double* Au, double* dot) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !u || !Au || !dot) return SBP_ERR_INVALID_ARG;
    int np = 0;
    if (h->var) STCHK(launch_var(h, 1, u, Au, u, 1.0, nullptr, &np));
    else STCHK(launch_pass<double>(h, 0, InPlain<double>{u}, OutStoreDot<double>{Au}, nullptr, h->partial, &np));
    hipLaunchKernelGGL(reduce_to, dim3(h->nprob), dim3(kRedThreads), 0, h->st, h->partial, np, h->dtmp, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dot, h->dtmp, sizeof(double) * h->nprob, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SBP_OK;
}

int sbp_apply_dot_noalloc(sbp_handle h, const double* u, double* Au, double* dot) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !u || !Au || !dot || h->var) return SBP_ERR_INVALID_ARG;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != h->dev) return SBP_ERR_INVALID_ARG;   // the handle's device must be current
    const Level& L = h->L[0];
    const size_t np = pass_nblocks<double, OutStoreDot<double>>(h->p, L);
    // block partials of every problem, then one result per problem
    if (np * (size_t)h->nprob + (size_t)h->nprob > kScratchDoubles) return SBP_ERR_SCRATCH;
    double* s = nullptr;
    STCHK(scratch_ptr(&s));
    int npi = 0;
    STCHK(launch_pass<double>(h, 0, InPlain<double>{u}, OutStoreDot<double>{Au}, nullptr, s, &npi));
    double* res = s + np * (size_t)h->nprob;
    hipLaunchKernelGGL(reduce_to, dim3(h->nprob), dim3(kRedThreads), 0, h->st, (const double*)s, npi, res, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dot, res, sizeof(double) * h->nprob, hipMemcpyDeviceToDevice, h->st));
    return SBP_OK;
}

int sbp_diag(sbp_handle h, double* d) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !d || h->var) return SBP_ERR_INVALID_ARG;
    hipLaunchKernelGGL(diag_kernel, dim3(ew_blocks(h->L[0].n)), dim3(256), 0, h->st, d, h->L[0].tb);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_assemble_rhs(sbp_handle h, const double* f, const double* gW, const double* gE, const double* gS,
                     const double* gN, double* b) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !b) return SBP_ERR_INVALID_ARG;
    hipLaunchKernelGGL(assemble_rhs, dim3(ew_blocks(h->L[0].n)), dim3(256), 0, h->st, f, gW, gE, gS, gN, b,
                       h->L[0].tb, h->rt);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_rhs_transpose(sbp_handle h, const double* lam, double* df, double* dgW, double* dgE, double* dgS,
                      double* dgN) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !lam) return SBP_ERR_INVALID_ARG;
    if (df) {
        hipLaunchKernelGGL(rhs_t_f, dim3(ew_blocks(h->L[0].n)), dim3(256), 0, h->st, lam, df, h->L[0].tb, h->rt);
        HIPCHK(hipGetLastError());
    }
    if (dgW || dgE || dgS || dgN) {
        hipLaunchKernelGGL(rhs_t_g, dim3((h->N + 255) / 256), dim3(256), 0, h->st, lam, dgW, dgE, dgS, dgN, h->L[0].tb,
                           h->rt);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

int sbp_hnorm(sbp_handle h, const double* v, double* out) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !v || !out) return SBP_ERR_INVALID_ARG;
    const int nb = ew_blocks(h->L[0].n);
    const double hh = 1.0 / (h->N - 1);
    hipLaunchKernelGGL((hnorm_partial<256>), dim3(nb), dim3(256), 0, h->st, v, h->L[0].tb, hh * hh, h->partial);
    hipLaunchKernelGGL(reduce_to, dim3(1), dim3(kRedThreads), 0, h->st, h->partial, nb, h->dtmp, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->dtmp, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SBP_OK;
}

int sbp_h_weights(sbp_handle h, double* hw) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !hw) return SBP_ERR_INVALID_ARG;
    const OpTab& t = h->L[0].tb;
    const int N = h->N;
    for (int k = 0; k < N; ++k) {
        const int r = k < N - 1 - k ? k : N - 1 - k;
        hw[k] = r < t.RH ? t.hw[r] : 1.0;
    }
    return SBP_OK;
}

int sbp_dot(sbp_handle h, const double* u, const double* v, double* out) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !u || !v || !out) return SBP_ERR_INVALID_ARG;
    const int nb = ew_blocks(h->L[0].n);
    hipLaunchKernelGGL((dot_partial<double, 256>), dim3(nb, 1), dim3(256), 0, h->st, u, v, h->L[0].n, h->partial);
    hipLaunchKernelGGL(reduce_to, dim3(1), dim3(kRedThreads), 0, h->st, h->partial, nb, h->dtmp, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, h->dtmp, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SBP_OK;
}

int sbp_restrict(sbp_handle h, int level, const void* rf, void* rc, int dtype) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !rf || !rc || level < 0 || level + 1 >= h->nlev) return SBP_ERR_INVALID_ARG;
    const int Nf = h->L[level].N, Nc = h->L[level + 1].N;
    const dim3 g((Nc + 255) / 256, Nc, h->nprob);
    if (dtype == SBP_F64)
        hipLaunchKernelGGL((restrict_fw<double>), g, dim3(256), 0, h->st, (const double*)rf, (double*)rc, Nf, Nc,
                           h->prm.restrict_scale, (const Scalars*)nullptr);
    else if (dtype == SBP_F32)
        hipLaunchKernelGGL((restrict_fw<float>), g, dim3(256), 0, h->st, (const float*)rf, (float*)rc, Nf, Nc,
                           (float)h->prm.restrict_scale, (const Scalars*)nullptr);
    else return SBP_ERR_INVALID_ARG;
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_prolong_add(sbp_handle h, int level, const void* ec, void* xf, int dtype) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !ec || !xf || level < 0 || level + 1 >= h->nlev) return SBP_ERR_INVALID_ARG;
    const int Nf = h->L[level].N, Nc = h->L[level + 1].N;
    const dim3 g((Nf + 255) / 256, Nf, h->nprob);
    if (dtype == SBP_F64)
        hipLaunchKernelGGL((prolong_add<double>), g, dim3(256), 0, h->st, (const double*)ec, (double*)xf, Nf, Nc,
                           (const Scalars*)nullptr);
    else if (dtype == SBP_F32)
        hipLaunchKernelGGL((prolong_add<float>), g, dim3(256), 0, h->st, (const float*)ec, (float*)xf, Nf, Nc,
                           (const Scalars*)nullptr);
    else return SBP_ERR_INVALID_ARG;
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_smooth(sbp_handle h, int level, const void* b, void* x, int nsweeps, int dtype) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !b || !x || level < 0 || level >= h->nlev || nsweeps < 0 || h->var) return SBP_ERR_INVALID_ARG;
    if (dtype == SBP_F64) return smooth_t<double>(h, level, (const double*)b, (double*)x, nsweeps);
    if (dtype == SBP_F32) {
        STCHK(ensure32(h));
        return smooth_t<float>(h, level, (const float*)b, (float*)x, nsweeps);
    }
    return SBP_ERR_INVALID_ARG;
}

int sbp_coarse_solve(sbp_handle h, const double* b, double* x) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !b || !x || !h->mg) return SBP_ERR_INVALID_ARG;
    if (h->coarse_nested) return coarse_exact<double>(h, b, x, nullptr, nullptr);
    hipLaunchKernelGGL((coarse_solve<double>), dim3((h->nc + kCoarseRows - 1) / kCoarseRows, h->nprob), dim3(256), 0,
                       h->st, h->Ainv, b, x, h->nc, (const Scalars*)nullptr);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_vcycle(sbp_handle h, const void* r, void* z, int dtype) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (h && h->var) {
        if (!r || !z || !h->vmg || dtype != SBP_F64) return SBP_ERR_INVALID_ARG;
        return var_vcycle<double>(h, 0, (const double*)r, (double*)z, nullptr, nullptr);
    }
    if (!h || !r || !z || !h->mg) return SBP_ERR_INVALID_ARG;
    if (dtype == SBP_F64) return vcycle_level<double>(h, 0, (const double*)r, (double*)z, nullptr, nullptr, nullptr);
    if (dtype == SBP_F32) {
        STCHK(ensure32(h));
        return vcycle_level<float>(h, 0, (const float*)r, (float*)z, nullptr, nullptr, nullptr);
    }
    return SBP_ERR_INVALID_ARG;
}

int sbp_cg(sbp_handle h, const double* b, double* x, double abstol, int maxiter, int* iters, double* resnorm,
           double* hist) {
    if (!h || !b || !x || maxiter < 0 || !(abstol >= 0) || h->nprob != 1) return SBP_ERR_INVALID_ARG;
    STCHK(krylov64(h, b, x, &abstol, maxiter, hist, false));   // (either operator)
    get_results(h, iters, resnorm, nullptr);
    return batch_status(h);
}

int sbp_mgpcg(sbp_handle h, const double* b, double* x, double abstol, int maxiter, int* iters, double* resnorm,
              double* hist) {
    if (!h || !b || !x || maxiter < 0 || !(abstol >= 0) || h->nprob != 1) return SBP_ERR_INVALID_ARG;
    if (h->var ? !h->vmg : !h->mg) return SBP_ERR_INVALID_SIZE;
    STCHK(krylov64(h, b, x, &abstol, maxiter, hist, true));
    get_results(h, iters, resnorm, nullptr);
    return batch_status(h);
}

int sbp_mgpcg_batch(sbp_handle h, const double* b, double* x, const double* abstol, int maxiter, int* iters,
                    double* resnorm, double* seconds, int* status) {
    if (!h || !b || !x || !abstol || maxiter < 0) return SBP_ERR_INVALID_ARG;
    for (int q = 0; q < h->nprob; ++q)
        if (!(abstol[q] >= 0)) return SBP_ERR_INVALID_ARG;
    if (!h->mg) return SBP_ERR_INVALID_SIZE;
    STCHK(krylov64(h, b, x, abstol, maxiter, nullptr, true));
    get_results(h, iters, resnorm, seconds);
    if (status)
        for (int q = 0; q < h->nprob; ++q) status[q] = status_of(h->sc_host[q]);
    return batch_status(h);
}

int sbp_mgpcg_mixed(sbp_handle h, const double* b, double* x, double abstol, int max_outer, int* outer_iters,
                    int* inner_iters, double* resnorm) {
    if (!h || !b || !x || max_outer < 0 || !(abstol >= 0) || h->nprob != 1) return SBP_ERR_INVALID_ARG;
    if (h->var) {   // FP64 MG-PCG with the V-cycle in FP32 (no refinement loop)
        if (!h->vmg) return SBP_ERR_INVALID_SIZE;
        STCHK(ensure_var32(h));
        const int cap = h->prm.inner_maxiter * std::max(max_outer, 1);
        STCHK(var_pcg_f32vc(h, b, x, abstol, cap));
        int it = 0;
        get_results(h, &it, resnorm, nullptr);
        if (outer_iters) *outer_iters = it;
        if (inner_iters) *inner_iters = it;
        return batch_status(h);
    }
    if (!h->mg) return SBP_ERR_INVALID_SIZE;
    int inner = 0;
    STCHK(mixed_solve(h, b, x, &abstol, max_outer, &inner));
    get_results(h, outer_iters, resnorm, nullptr);
    if (inner_iters) *inner_iters = inner;
    return mixed_status(h);
}

int sbp_mgpcg_mixed_batch(sbp_handle h, const double* b, double* x, const double* abstol, int max_outer,
                          int* outer_iters, int* inner_iters, double* resnorm, double* seconds, int* status) {
    if (!h || !b || !x || !abstol || max_outer < 0) return SBP_ERR_INVALID_ARG;
    for (int q = 0; q < h->nprob; ++q)
        if (!(abstol[q] >= 0)) return SBP_ERR_INVALID_ARG;
    if (!h->mg) return SBP_ERR_INVALID_SIZE;
    std::vector<int> inner(h->nprob, 0);
    STCHK(mixed_solve(h, b, x, abstol, max_outer, inner.data()));
    get_results(h, outer_iters, resnorm, seconds);
    if (inner_iters)
        for (int q = 0; q < h->nprob; ++q) inner_iters[q] = inner[q];
    if (status)
        for (int q = 0; q < h->nprob; ++q) status[q] = mixed_status_of(h->sc_host[q]);
    return mixed_status(h);
}

// Adjoint gradient: L = 1/2 (u - u_obs)' H~ (u - u_obs), u = A^-1 b(f, g).
int sbp_adjoint_grad(sbp_handle h, const double* u, const double* u_obs, double abstol, int maxiter, double* loss,
                     double* lam, double* df, double* dgW, double* dgE, double* dgS, double* dgN, int* iters) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !u || !u_obs || maxiter < 0 || !(abstol >= 0) || h->nprob != 1) return SBP_ERR_INVALID_ARG;
    if (!h->mg) return SBP_ERR_INVALID_SIZE;
    const size_t n = h->L[0].n;
    if (!h->w) HIPCHK(hipMalloc(&h->w, n * sizeof(double)));
    double* lm = lam;
    if (!lm) {
        if (!h->lam) HIPCHK(hipMalloc(&h->lam, n * sizeof(double)));
        lm = h->lam;
    }
    const int nb = ew_blocks(n);
    const double hh = 1.0 / (h->N - 1);
    hipLaunchKernelGGL(adjoint_source, dim3(nb), dim3(256), 0, h->st, u, u_obs, h->w, h->L[0].tb, hh * hh, h->partial);
    hipLaunchKernelGGL(reduce_to, dim3(1), dim3(kRedThreads), 0, h->st, h->partial, nb, h->dtmp, 0);
    hipLaunchKernelGGL((fill<double, kEwThreads>), dim3(nb, 1), dim3(kEwThreads), 0, h->st, lm, 0.0, n,
                       (const Scalars*)nullptr);
    HIPCHK(hipGetLastError());
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, h->dtmp, sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        *loss *= 0.5;
    }
    // A is symmetric, so the adjoint solve A^T lambda = w is the forward MG-PCG.
    STCHK(krylov64(h, h->w, lm, &abstol, maxiter, nullptr, true));
    get_results(h, iters, nullptr, nullptr);
    const int rc = batch_status(h);
    if (rc != SBP_OK) return rc;
    return sbp_rhs_transpose(h, lm, df, dgW, dgE, dgS, dgN);
}

// ---------------------------------------------------------------------------
// Initial-guess interpolation CG (nested iteration, one problem)
// ---------------------------------------------------------------------------
namespace {

// Handle of the next coarser grid ((N+1)/2, same order and parameters, level
// l keeping h's tile shapes for level l+1), made on first use; it runs on h's
// stream.
int coarser(sbp_handle h, sbp_handle* out) {
    if (h->nprob != 1 || !h->mg || h->nlev < 2) return SBP_ERR_INVALID_SIZE;
    if (!h->child) {
        sbp_handle c = nullptr;
        STCHK(sbp_create(h->L[1].N, h->p, &h->prm, &c));
        for (int l = 0; l < c->nlev && l + 1 < h->nlev; ++l)
            for (int q = 0; q < 2; ++q) {
                c->L[l].ti[q] = h->L[l + 1].ti[q];
                c->L[l].tj[q] = h->L[l + 1].tj[q];
                c->L[l].ta[q] = h->L[l + 1].ta[q];
                c->L[l].fti[q] = h->L[l + 1].fti[q];
                c->L[l].ftj[q] = h->L[l + 1].ftj[q];
            }
        if (hipMalloc(&h->xc, c->L[0].n * sizeof(double)) != hipSuccess) {
            (void)sbp_destroy(c);
            return SBP_ERR_HIP;
        }
        h->child = c;
    }
    h->child->st = h->st;
    *out = h->child;
    return SBP_OK;
}

// bs[0] is this grid's right-hand side, bs[1] the next coarser grid's, ...
// The coarsest of the nlev grids is solved exactly; every finer one starts
// CG from the bilinear prolongation of the solution one level down.
// iters / resnorm [nlev - 2] receive this grid's CG results.
int nested_cg(sbp_handle h, int nlev, const double* const* bs, double* x, double abstol, int maxiter, int* iters,
              double* resnorm, double* hist) {
    sbp_handle c = nullptr;
    STCHK(coarser(h, &c));
    if (nlev == 2) {
        STCHK(exact_solve(c, bs[1], h->xc));
    } else {
        const int rc = nested_cg(c, nlev - 1, bs + 1, h->xc, abstol, maxiter, iters, resnorm, nullptr);
        if (rc != SBP_OK && rc != SBP_ERR_NOT_CONVERGED) return rc;
    }
    HIPCHK(hipMemsetAsync(x, 0, h->L[0].n * sizeof(double), h->st));
    STCHK(sbp_prolong_add(h, 0, h->xc, x, SBP_F64));
    int it = 0;
    double rn = 0.0;
    const int rc = sbp_cg(h, bs[0], x, abstol, maxiter, &it, &rn, hist);
    if (iters) iters[nlev - 2] = it;
    if (resnorm) resnorm[nlev - 2] = rn;
    return rc;
}

}  // namespace

int sbp_cg_interp(sbp_handle h, int nlev, const double* b, const double* b2h, const double* b4h, double* x,
                  double abstol, int maxiter, int* iters, double* resnorm, double* hist) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !b || !b2h || !x || (nlev != 2 && nlev != 3) || (nlev == 3 && !b4h) || maxiter < 0 ||
        !(abstol >= 0) || h->nprob != 1)
        return SBP_ERR_INVALID_ARG;
    const double* bs[3] = {b, b2h, b4h};
    return nested_cg(h, nlev, bs, x, abstol, maxiter, iters, resnorm, hist);
}

// ---------------------------------------------------------------------------
// CSR comparison path (one problem)
// ---------------------------------------------------------------------------
int sbp_csr_assemble(sbp_handle h, int32_t* row_ptr, int32_t* col, double* val, int64_t* nnz) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !nnz || h->var) return SBP_ERR_INVALID_ARG;
    const int n = (int)h->L[0].n;
    int32_t* counts = nullptr;
    int32_t* rp = row_ptr;
    int64_t* tot = nullptr;
    int64_t* bsum = nullptr;
    const int nsb = (n + kScanPer - 1) / kScanPer;
    HIPCHK(hipMalloc(&counts, (size_t)n * sizeof(int32_t)));
    HIPCHK(hipMalloc(&tot, sizeof(int64_t)));
    HIPCHK(hipMalloc(&bsum, (size_t)nsb * sizeof(int64_t)));
    if (!rp) HIPCHK(hipMalloc(&rp, (size_t)(n + 1) * sizeof(int32_t)));
    const int nb = (n + 255) / 256;
    hipLaunchKernelGGL(csr_fill, dim3(nb), dim3(256), 0, h->st, rp, col, val, h->L[0].tb, 1, counts);
    hipLaunchKernelGGL(scan_rows, dim3(nsb), dim3(kScanThreads), 0, h->st, (const int32_t*)counts, rp, n, bsum);
    hipLaunchKernelGGL(scan_block_sums, dim3(1), dim3(1024), 0, h->st, bsum, nsb, tot);
    hipLaunchKernelGGL(scan_add, dim3(nb), dim3(256), 0, h->st, rp, (const int64_t*)bsum, n, (const int64_t*)tot);
    if (col && val && row_ptr)
        hipLaunchKernelGGL(csr_fill, dim3(nb), dim3(256), 0, h->st, rp, col, val, h->L[0].tb, 0, counts);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(nnz, tot, sizeof(int64_t), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    (void)hipFree(counts);
    (void)hipFree(tot);
    (void)hipFree(bsum);
    if (!row_ptr) (void)hipFree(rp);
    return SBP_OK;
}

int sbp_csr_spmv(sbp_handle h, const int32_t* rp, const int32_t* col, const double* val, const double* x,
                 double* y) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !rp || !col || !val || !x || !y) return SBP_ERR_INVALID_ARG;
    const int n = (int)h->L[0].n;
    const int nb = std::min(kCsrMaxBlocks, (n + 256 / kCsrLanes - 1) / (256 / kCsrLanes));
    hipLaunchKernelGGL(csr_spmv, dim3(nb), dim3(256), 0, h->st, rp, col, val, x, y, n, nullptr,
                       (const Scalars*)nullptr);
    HIPCHK(hipGetLastError());
    return SBP_OK;
}

int sbp_csr_cg(sbp_handle h, const int32_t* rp, const int32_t* col, const double* val, const double* b, double* x,
               double abstol, int maxiter, int* iters, double* resnorm) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !rp || !col || !val || !b || !x || maxiter < 0 || !(abstol >= 0) || h->nprob != 1)
        return SBP_ERR_INVALID_ARG;
    hipStream_t st = h->st;
    const int n = (int)h->L[0].n;
    const int nbs = (n + 255) / 256;
    if ((size_t)nbs > h->npp) return SBP_ERR_INVALID_SIZE;
    Scalars* sc = h->sc;
    STCHK(put_abstol(h, &abstol));
    hipLaunchKernelGGL(stamp_start, dim3(1), dim3(256), 0, st, sc, 1);
    // r = b - A x with block partials of r'r
    hipLaunchKernelGGL(csr_resid, dim3(nbs), dim3(256), 0, st, rp, col, val, x, b, h->r, n, h->partial);
    hipLaunchKernelGGL(krylov_init, dim3(1), dim3(kRedThreads), 0, st, h->partial, nbs, sc, (const double*)h->dabs,
                       (double*)nullptr, maxiter);
    hipLaunchKernelGGL((copy_vec<double, kEwThreads>), dim3(ew_blocks(n), 1), dim3(kEwThreads), 0, st,
                       (const double*)h->r, h->pv, (size_t)n, (const Scalars*)sc);
    HIPCHK(hipGetLastError());
    const dim3 g(ew_blocks(n), 1);
    const int nbv = std::min(kCsrMaxBlocks, (n + 256 / kCsrLanes - 1) / (256 / kCsrLanes));
    STCHK(krylov_iterate(h, sc, maxiter, [&](int) -> int {
        hipLaunchKernelGGL(csr_spmv, dim3(nbv), dim3(256), 0, st, rp, col, val, (const double*)h->pv, h->Ap, n,
                           h->partial, (const Scalars*)sc);
        hipLaunchKernelGGL(krylov_alpha, dim3(1), dim3(kRedThreads), 0, st, h->partial, nbv, sc, 0);
        hipLaunchKernelGGL((update_xr<double, kEwThreads>), g, dim3(kEwThreads), 0, st, x, h->r, (const double*)h->pv,
                           (const double*)h->Ap, (const Scalars*)sc, (size_t)n, h->partial);
        hipLaunchKernelGGL(krylov_check, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc, 1);
        hipLaunchKernelGGL((update_p<double, kEwThreads>), g, dim3(kEwThreads), 0, st, h->pv, (const double*)h->r,
                           (const Scalars*)sc, (size_t)n);
        HIPCHK(hipGetLastError());
        return SBP_OK;
    }));
    STCHK(poll(h, sc));
    get_results(h, iters, resnorm, nullptr);
    return batch_status(h);
}

// ---- the CSR operator through rocSPARSE -----------------------------------
namespace {
#define VSCHK(x)                                             \
    do {                                                     \
        if ((x) != rocsparse_status_success) return SBP_ERR_VENDOR; \
    } while (0)

// y = A x with the attached matrix (descriptors repointed, compute stage).
int vendor_spmv(sbp_handle h, const double* x, double* y) {
    const double one = 1.0, zero = 0.0;
    VSCHK(rocsparse_set_stream(h->vs, h->st));
    VSCHK(rocsparse_dnvec_set_values(h->vs_x, (void*)x));
    VSCHK(rocsparse_dnvec_set_values(h->vs_y, (void*)y));
    size_t bs = h->vs_bufsz;
    VSCHK(rocsparse_spmv(h->vs, rocsparse_operation_none, &one, h->vs_mat, h->vs_x, &zero, h->vs_y,
                         rocsparse_datatype_f64_r, h->vs_alg, rocsparse_spmv_stage_compute, &bs, h->vs_buf));
    return SBP_OK;
}
}  // namespace

int sbp_csr_vendor_attach(sbp_handle h, const int32_t* rp, const int32_t* col, const double* val, int64_t nnz,
                          int alg) {
    if (h && h->ragged) return SBP_ERR_INVALID_ARG;   // ragged handles serve the batched solvers only
    if (!h || !rp || !col || !val || nnz <= 0 || h->nprob != 1 || (alg != 0 && alg != 1)) return SBP_ERR_INVALID_ARG;
    const int64_t n = (int64_t)h->L[0].n;
    if (h->vs_mat) {   // re-attach: drop the previous matrix and its analysis
        (void)rocsparse_destroy_spmat_descr(h->vs_mat);
        h->vs_mat = nullptr;
        if (h->vs_buf) (void)hipFree(h->vs_buf);
        h->vs_buf = nullptr;
        h->vs_bufsz = 0;
    }
    if (!h->vs) {
        VSCHK(rocsparse_create_handle(&h->vs));
        VSCHK(rocsparse_create_dnvec_descr(&h->vs_x, n, (void*)h->pv, rocsparse_datatype_f64_r));
        VSCHK(rocsparse_create_dnvec_descr(&h->vs_y, n, (void*)h->Ap, rocsparse_datatype_f64_r));
    }
    VSCHK(rocsparse_set_stream(h->vs, h->st));
    h->vs_alg = alg == 0 ? rocsparse_spmv_alg_csr_adaptive : rocsparse_spmv_alg_csr_rowsplit;
    VSCHK(rocsparse_create_csr_descr(&h->vs_mat, n, n, nnz, (void*)rp, (void*)col, (void*)val,
                                     rocsparse_indextype_i32, rocsparse_indextype_i32, rocsparse_index_base_zero,
                                     rocsparse_datatype_f64_r));
    const double one = 1.0, zero = 0.0;
    VSCHK(rocsparse_dnvec_set_values(h->vs_x, (void*)h->pv));
    VSCHK(rocsparse_dnvec_set_values(h->vs_y, (void*)h->Ap));
    size_t bs = 0;
    VSCHK(rocsparse_spmv(h->vs, rocsparse_operation_none, &one, h->vs_mat, h->vs_x, &zero, h->vs_y,
                         rocsparse_datatype_f64_r, h->vs_alg, rocsparse_spmv_stage_buffer_size, &bs, nullptr));
    if (bs > 0) HIPCHK(hipMalloc(&h->vs_buf, bs));
    h->vs_bufsz = bs;
    VSCHK(rocsparse_spmv(h->vs, rocsparse_operation_none, &one, h->vs_mat, h->vs_x, &zero, h->vs_y,
                         rocsparse_datatype_f64_r, h->vs_alg, rocsparse_spmv_stage_preprocess, &bs, h->vs_buf));
    HIPCHK(hipStreamSynchronize(h->st));
    return SBP_OK;
}

int sbp_csr_vendor_spmv(sbp_handle h, const double* x, double* y) {
    if (!h || !x || !y || !h->vs_mat) return SBP_ERR_INVALID_ARG;
    return vendor_spmv(h, x, y);
}

int sbp_csr_vendor_cg(sbp_handle h, const double* b, double* x, double abstol, int maxiter, int* iters,
                      double* resnorm) {
    if (!h || !b || !x || !h->vs_mat || maxiter < 0 || !(abstol >= 0) || h->nprob != 1) return SBP_ERR_INVALID_ARG;
    hipStream_t st = h->st;
    const size_t n = h->L[0].n;
    const dim3 g(ew_blocks(n), 1);
    if ((size_t)g.x > h->npp) return SBP_ERR_INVALID_SIZE;
    Scalars* sc = h->sc;
    STCHK(put_abstol(h, &abstol));
    hipLaunchKernelGGL(stamp_start, dim3(1), dim3(256), 0, st, sc, 1);
    // r = b - A x: Ap = A x, then r = b - Ap with partials of r'r
    STCHK(vendor_spmv(h, x, h->Ap));
    hipLaunchKernelGGL((axpby_dot<kEwThreads>), g, dim3(kEwThreads), 0, st, b, (const double*)h->Ap, h->r, n,
                       h->partial);
    hipLaunchKernelGGL(krylov_init, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc,
                       (const double*)h->dabs, (double*)nullptr, maxiter);
    hipLaunchKernelGGL((copy_vec<double, kEwThreads>), g, dim3(kEwThreads), 0, st, (const double*)h->r, h->pv, n,
                       (const Scalars*)sc);
    HIPCHK(hipGetLastError());
    STCHK(krylov_iterate(h, sc, maxiter, [&](int) -> int {
        STCHK(vendor_spmv(h, h->pv, h->Ap));
        hipLaunchKernelGGL((dot_partial<double, kEwThreads>), g, dim3(kEwThreads), 0, st, (const double*)h->pv,
                           (const double*)h->Ap, n, h->partial);
        hipLaunchKernelGGL(krylov_alpha, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc, 0);
        hipLaunchKernelGGL((update_xr<double, kEwThreads>), g, dim3(kEwThreads), 0, st, x, h->r, (const double*)h->pv,
                           (const double*)h->Ap, (const Scalars*)sc, n, h->partial);
        hipLaunchKernelGGL(krylov_check, dim3(1), dim3(kRedThreads), 0, st, h->partial, (int)g.x, sc, 1);
        hipLaunchKernelGGL((update_p<double, kEwThreads>), g, dim3(kEwThreads), 0, st, h->pv, (const double*)h->r,
                           (const Scalars*)sc, n);
        HIPCHK(hipGetLastError());
        return SBP_OK;
    }));
    STCHK(poll(h, sc));
    get_results(h, iters, resnorm, nullptr);
    return batch_status(h);
}

// Streaming copy: 16-byte vector body when both ends are 16-byte aligned,
// bytes for the tail (or everything when unaligned).
int sbp_copy(sbp_handle h, const void* src, void* dst, size_t bytes) {
    if (!h || !src || !dst) return SBP_ERR_INVALID_ARG;
    if (bytes == 0) return SBP_OK;
    const bool aligned = (((uintptr_t)src | (uintptr_t)dst) & 15) == 0;
    const size_t n16 = aligned ? bytes / 16 : 0;
    if (n16) {
        const int nb = (int)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 8192));
        hipLaunchKernelGGL(copy_vec16, dim3(nb), dim3(256), 0, h->st, (const uint4*)src, (uint4*)dst, n16);
        HIPCHK(hipGetLastError());
    }
    const size_t done = n16 * 16, rest = bytes - done;
    if (rest) {
        const int nb = (int)std::max<size_t>(1, std::min<size_t>((rest + 255) / 256, 8192));
        hipLaunchKernelGGL(copy_u8, dim3(nb), dim3(256), 0, h->st, (const unsigned char*)src + done,
                           (unsigned char*)dst + done, rest);
        HIPCHK(hipGetLastError());
    }
    return SBP_OK;
}

}  // extern "C"
