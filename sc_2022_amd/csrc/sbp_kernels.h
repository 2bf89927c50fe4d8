// sbp_kernels.h -- gfx950 kernels of the Krylov loops, transfers, coarse
// solve, right-hand side and CSR comparison path (the stencil pass itself is
// in sbp_stencil.h).  Included by exactly one translation unit.
#pragma once
#include "sbp_stencil.h"

namespace sbp {

// ---------------------------------------------------------------------------
// Device-resident Krylov state, one block per stacked problem
// ---------------------------------------------------------------------------
struct Scalars {
    double rr;             // r'r (CG) or latest ||r||^2 (PCG)
    double rz;             // r'z (PCG)
    double pAp;
    double alpha, beta;
    double abstol;
    double resnorm;        // latest ||r||_2
    double rnorm0;         // ||r_0||_2 of the current solve (host-side convergence prediction)
    double bnorm;
    double* hist;          // nullable residual-norm history, hist[k] = ||r_k||_2
    double dprev;          // refinement: norm of the previous correction
    double est;            // refinement: a-posteriori error estimate
    unsigned long long t0; // wall clock at the solve's start
    unsigned long long t1; // wall clock when this problem converged
    int maxiter;           // iteration cap (tested on the device)
    int iters;             // iterations completed (advanced on the device)
    int done;              // convergence mask: kernels of a done problem fall through
    int status;
    int inner;             // refinement: inner iterations over the corrections so far
};
constexpr int kScalarInts = (int)(sizeof(Scalars) / sizeof(int));
constexpr int kScalarDoubles = (int)(sizeof(Scalars) / sizeof(double));

__device__ __forceinline__ bool skip(const Scalars* sc, int q) { return sc && sc[q].done; }

// ---------------------------------------------------------------------------
// Transfers (blockIdx.z = problem)
// ---------------------------------------------------------------------------
// rc(I,J) = scale * sum_{a,b in -1..1} w_a w_b rf(2I+a, 2J+b) / 4, w_0 = 1, w_+-1 = 1/2,
// fine points outside the grid count as zero.
template <typename T>
__global__ __launch_bounds__(256) void restrict_fw(const T* __restrict__ rf, T* __restrict__ rc, int Nf, int Nc,
                                                  T scale, const Scalars* sc) {
    const int q = blockIdx.z;
    if (skip(sc, q)) return;
    rf += (size_t)q * Nf * Nf;
    rc += (size_t)q * Nc * Nc;
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int Jc = blockIdx.y;
    if (I >= Nc) return;
    const int fi = 2 * I, fj = 2 * Jc;
    T s = T(0);
#pragma unroll
    for (int b = -1; b <= 1; ++b) {
        const int jj = fj + b;
        if (jj < 0 || jj >= Nf) continue;
        const T wb = b == 0 ? T(1) : T(0.5);
        const T* row = rf + (size_t)jj * Nf;
        T rs = row[fi];
        if (fi - 1 >= 0) rs += T(0.5) * row[fi - 1];
        if (fi + 1 < Nf) rs += T(0.5) * row[fi + 1];
        s += wb * rs;
    }
    rc[(size_t)Jc * Nc + I] = scale * T(0.25) * s;
}

// xf += P ec
template <typename T>
__global__ __launch_bounds__(256) void prolong_add(const T* __restrict__ ec, T* __restrict__ xf, int Nf, int Nc,
                                                  const Scalars* sc) {
    const int q = blockIdx.z;
    if (skip(sc, q)) return;
    ec += (size_t)q * Nc * Nc;
    xf += (size_t)q * Nf * Nf;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= Nf) return;
    const int ic = i >> 1, jc = j >> 1;
    const T* e = ec + (size_t)jc * Nc + ic;
    T v;
    if (!(i & 1)) v = (!(j & 1)) ? e[0] : (e[0] + e[Nc]) * T(0.5);
    else v = (!(j & 1)) ? (e[0] + e[1]) * T(0.5) : (e[0] + e[1] + e[Nc] + e[Nc + 1]) * T(0.25);
    xf[(size_t)j * Nf + i] += v;
}

// Coarsest-level exact solve x = Ainv b with the FP64 inverse (symmetric,
// stored n x n); blockIdx.y = problem.  Blocks of 4 waves own kCoarseRows
// consecutive rows; each wave forms its rows' dot products across its 64
// lanes (coalesced row reads, all loads of a row issued together).
constexpr int kCoarseRows = 8;
template <typename T>
__global__ __launch_bounds__(256) void coarse_solve(const double* __restrict__ Ainv, const T* __restrict__ b,
                                                   T* __restrict__ x, int n, const Scalars* sc) {
    const int q = blockIdx.y;
    if (skip(sc, q)) return;
    b += (size_t)q * n;
    x += (size_t)q * n;
    __shared__ double bs[1089];   // up to 33 x 33
    for (int k = threadIdx.x; k < n; k += blockDim.x) bs[k] = (double)b[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int rr = 0; rr < kCoarseRows / 4; ++rr) {
        const int r = blockIdx.x * kCoarseRows + rr * 4 + wv;
        if (r >= n) break;
        const double* row = Ainv + (size_t)r * n;
        double s = 0.0;
        for (int k = lane; k < n; k += 64) s += row[k] * bs[k];
        s = wave_sum(s);
        if (lane == 0) x[r] = (T)s;
    }
}

// ---------------------------------------------------------------------------
// Reductions: one workgroup per problem (blockIdx.x = problem), each summing
// that problem's np block partials at partial[q * np ...]
// ---------------------------------------------------------------------------
// Deterministic sum (fixed order: thread t sums t, t+NT, ... with four
// independent accumulators, then a tree across lanes and waves).  Launched
// with kRedThreads threads.
constexpr int kRedThreads = 1024;
__device__ __forceinline__ double sum_partials(const double* partial, int np) {
    __shared__ double red[kRedThreads / 64];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = threadIdx.x;
    const int st = blockDim.x;
    for (; k + 3 * st < np; k += 4 * st) {
        a0 += partial[k];
        a1 += partial[k + st];
        a2 += partial[k + 2 * st];
        a3 += partial[k + 3 * st];
    }
    for (; k < np; k += st) a0 += partial[k];
    double s = (a0 + a1) + (a2 + a3);
    s = wave_sum(s);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = s;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) tot += red[q];
    return tot;   // valid on thread 0
}

__device__ __forceinline__ void mark_done(Scalars* s) {
    s->done = 1;
    s->t1 = wall_clock64();
}

// out[q] = sum (or sqrt of it) of problem q's partials.
__global__ __launch_bounds__(1024) void reduce_to(const double* partial, int np, double* out, int take_sqrt) {
    const int q = blockIdx.x;
    double s = sum_partials(partial + (size_t)q * np, np);
    if (threadIdx.x == 0) out[q] = take_sqrt ? sqrt(s) : s;
}

// Wall-clock stamp of the solve's start for every problem.
__global__ void stamp_start(Scalars* sc, int nprob) {
    const unsigned long long t = wall_clock64();
    for (int q = threadIdx.x; q < nprob; q += blockDim.x) { sc[q].t0 = t; sc[q].t1 = 0; }
}

// alpha = num / pAp where num = rr (CG, use_rz = 0) or rz (PCG, use_rz = 1).
__global__ __launch_bounds__(1024) void krylov_alpha(const double* partial, int np, Scalars* scb, int use_rz) {
    Scalars* sc = scb + blockIdx.x;
    if (sc->done) return;
    double pAp = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        sc->pAp = pAp;
        if (!(pAp > 0.0) || !isfinite(pAp)) {
            sc->status = isfinite(pAp) ? 6 : 5;
            mark_done(sc);
            return;
        }
        sc->alpha = (use_rz ? sc->rz : sc->rr) / pAp;
    }
}

// rs = ||r||^2; iteration counter and history advance on the device (so a
// captured iteration replays unchanged); converge test ||r|| < abstol; for CG
// also beta.
__global__ __launch_bounds__(1024) void krylov_check(const double* partial, int np, Scalars* scb, int cg_beta) {
    Scalars* sc = scb + blockIdx.x;
    if (sc->done) return;
    double rs = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        const double nr = sqrt(rs);
        const int k = sc->iters + 1;
        sc->iters = k;
        sc->resnorm = nr;
        if (sc->hist) sc->hist[k] = nr;
        if (!isfinite(rs)) { sc->status = 5; mark_done(sc); return; }
        if (nr < sc->abstol) { mark_done(sc); return; }
        if (k >= sc->maxiter) { sc->status = 7; mark_done(sc); return; }
        if (cg_beta) { sc->beta = rs / sc->rr; }
        sc->rr = rs;
    }
}

// rz_new = r'z; beta = rz_new / rz.
__global__ __launch_bounds__(1024) void krylov_beta(const double* partial, int np, Scalars* scb) {
    Scalars* sc = scb + blockIdx.x;
    if (sc->done) return;
    double rz = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        if (!isfinite(rz)) { sc->status = 5; mark_done(sc); return; }
        sc->beta = rz / sc->rz;
        sc->rz = rz;
    }
}

// Initial state: rr = sum partial; test ||r0|| < abstol[q] (or maxiter = 0).
__global__ __launch_bounds__(1024) void krylov_init(const double* partial, int np, Scalars* scb,
                                                   const double* abstol, double* hist, int maxiter) {
    Scalars* sc = scb + blockIdx.x;
    double rs = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        sc->abstol = abstol[blockIdx.x];
        sc->maxiter = maxiter;
        sc->rr = rs;
        sc->resnorm = sqrt(rs);
        sc->rnorm0 = sqrt(rs);
        sc->hist = hist;
        sc->iters = 0;
        sc->status = 0;
        sc->done = 0;
        sc->alpha = sc->beta = 0.0;
        if (hist) hist[0] = sqrt(rs);
        if (!isfinite(rs)) { sc->status = 5; mark_done(sc); return; }
        if (sqrt(rs) < sc->abstol) { mark_done(sc); return; }
        if (maxiter <= 0) { sc->status = 7; mark_done(sc); }
    }
}

// rz0 = r'z (stage 2) without touching done.
__global__ __launch_bounds__(1024) void krylov_set_rz(const double* partial, int np, Scalars* scb) {
    Scalars* sc = scb + blockIdx.x;
    if (sc->done) return;
    double rz = sum_partials(partial + (size_t)blockIdx.x * np, np);
    if (threadIdx.x == 0) {
        if (!isfinite(rz)) { sc->status = 5; mark_done(sc); return; }
        sc->rz = rz;
    }
}

// ---------------------------------------------------------------------------
// Elementwise kernels: grid (blocks per problem, problems); problem q owns
// elements [q n, (q+1) n) and partial slots [q gridDim.x, (q+1) gridDim.x).
// ---------------------------------------------------------------------------
// r -= alpha Ap ; block partials of r'r.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void update_r(T* __restrict__ r, const T* __restrict__ Ap, const Scalars* scb,
                                              size_t n, double* partial) {
    const int q = blockIdx.y;
    const Scalars* sc = scb + q;
    if (sc->done) return;
    r += q * n;
    Ap += q * n;
    const T a = (T)sc->alpha;
    T acc = T(0);
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) {
        const T rv = r[k] - a * Ap[k];
        r[k] = rv;
        acc += rv * rv;
    }
    T s = block_sum<T, NT>(acc);
    if (threadIdx.x == 0) partial[(size_t)q * gridDim.x + blockIdx.x] = (double)s;
}

// Unfused CG updates (CSR comparison path): x += alpha p ; r -= alpha Ap with
// block partials of r'r, then p = r + beta p.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void update_xr(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                                               const T* __restrict__ Ap, const Scalars* scb, size_t n,
                                               double* partial) {
    const int q = blockIdx.y;
    const Scalars* sc = scb + q;
    if (sc->done) return;
    x += q * n;
    r += q * n;
    p += q * n;
    Ap += q * n;
    const T a = (T)sc->alpha;
    T acc = T(0);
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) {
        x[k] += a * p[k];
        const T rv = r[k] - a * Ap[k];
        r[k] = rv;
        acc += rv * rv;
    }
    T s = block_sum<T, NT>(acc);
    if (threadIdx.x == 0) partial[(size_t)q * gridDim.x + blockIdx.x] = (double)s;
}

// update_xr in FP64 that also leaves the new residual rounded to FP32 in r32
// (the input of a single-precision preconditioner; single problem).
template <int NT>
__global__ __launch_bounds__(NT) void update_xr_r32(double* __restrict__ x, double* __restrict__ r,
                                                   float* __restrict__ r32, const double* __restrict__ p,
                                                   const double* __restrict__ Ap, const Scalars* sc, size_t n,
                                                   double* partial) {
    if (sc->done) return;
    const double a = sc->alpha;
    double acc = 0.0;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) {
        x[k] += a * p[k];
        const double rv = r[k] - a * Ap[k];
        r[k] = rv;
        r32[k] = (float)rv;
        acc += rv * rv;
    }
    const double s = block_sum<double, NT>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// p = z + beta p; z may be stored in a narrower type Z than p.
template <typename T, int NT, typename Z = T>
__global__ __launch_bounds__(NT) void update_p(T* __restrict__ p, const Z* __restrict__ r, const Scalars* scb,
                                              size_t n) {
    const int q = blockIdx.y;
    const Scalars* sc = scb + q;
    if (sc->done) return;
    p += q * n;
    r += q * n;
    const T b = (T)sc->beta;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) p[k] = (T)r[k] + b * p[k];
}

// The update x += alpha_K p_K still pending after the last iteration K (the
// fused search-direction pass of iteration K+1 would have applied it).
// p_K sits in pbuf[K % 2]; nothing is pending after a breakdown or when no
// iteration ran.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void flush_x(T* __restrict__ x, const T* __restrict__ p0, const T* __restrict__ p1,
                                             const Scalars* scb, size_t n) {
    const int q = blockIdx.y;
    const Scalars* sc = scb + q;
    const int K = sc->iters;
    if (K < 1 || sc->status == 5 || sc->status == 6) return;
    const T* pk = ((K & 1) ? p1 : p0) + q * n;
    x += q * n;
    const T a = (T)sc->alpha;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) x[k] += a * pk[k];
}

// block partials of u'v (v may equal u).
// v may be stored in a narrower type V than u (the sum is in T).
template <typename T, int NT, typename V = T>
__global__ __launch_bounds__(NT) void dot_partial(const T* __restrict__ u, const V* __restrict__ v, size_t n,
                                                 double* partial) {
    const int q = blockIdx.y;
    u += q * n;
    v += q * n;
    T acc = T(0);
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) acc += u[k] * (T)v[k];
    T s = block_sum<T, NT>(acc);
    if (threadIdx.x == 0) partial[(size_t)q * gridDim.x + blockIdx.x] = (double)s;
}

// block partials of v' H~ v with H~(i,j) = h^2 hw[i] hw[j] (single problem).
template <int NT>
__global__ __launch_bounds__(NT) void hnorm_partial(const double* __restrict__ v, OpTab t, double h2,
                                                   double* partial) {
    const size_t n = (size_t)t.N * t.N;
    double acc = 0.0;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t q = (size_t)blockIdx.x * NT + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % t.N), j = (int)(q / t.N);
        acc += hw_of<double>(t, i) * hw_of<double>(t, j) * v[q] * v[q];
    }
    double s = block_sum<double, NT>(acc) * h2;
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void fill(T* __restrict__ x, T v, size_t n, const Scalars* sc) {
    const int q = blockIdx.y;
    if (skip(sc, q)) return;
    x += q * n;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) x[k] = v;
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void copy_vec(const T* __restrict__ x, T* __restrict__ y, size_t n,
                                              const Scalars* sc) {
    const int q = blockIdx.y;
    if (skip(sc, q)) return;
    x += q * n;
    y += q * n;
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) y[k] = x[k];
}

// y = (D) x, one problem
template <typename S, typename D, int NT>
__global__ __launch_bounds__(NT) void cast_vec(const S* __restrict__ x, D* __restrict__ y, size_t n) {
    const size_t stride = (size_t)gridDim.x * NT;
    for (size_t k = (size_t)blockIdx.x * NT + threadIdx.x; k < n; k += stride) y[k] = (D)x[k];
}

// diag(A)
__global__ __launch_bounds__(256) void diag_kernel(double* d, OpTab t) {
    const size_t n = (size_t)t.N * t.N;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % t.N), j = (int)(q / t.N);
        d[q] = diag_of<double>(t, i, j);
    }
}

// ---------------------------------------------------------------------------
// Right-hand side assembly and its transpose (closed form of
// b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN)):
//   b(i,j) = -h^2 hw_i hw_j f(i,j)
//          + [j < nbs]        hw_i (13 d_{j,0} - bs_j) gW(i)
//          + [N-1-j < nbs]    hw_i (13 d_{N-1-j,0} - bs_{N-1-j}) gE(i)
//          + [i == 0]         h hw_j gS(j)
//          + [i == N-1]       h hw_j gN(j)
// ---------------------------------------------------------------------------
struct RhsTab {
    double bs[kMaxR];   // boundary-derivative row (dimensionless, h * BS)
    int nbs;
    double h;
    double bW, bE;      // Dirichlet-face coefficients (variable-coefficient operator; 1 otherwise)
};

__device__ __forceinline__ double dir_w(const RhsTab& r, int k) {
    if (k >= r.nbs) return 0.0;
    return (k == 0 ? 13.0 : 0.0) - r.bs[k];
}

__global__ __launch_bounds__(256) void assemble_rhs(const double* f, const double* gW, const double* gE,
                                                   const double* gS, const double* gN, double* b, OpTab t,
                                                   RhsTab rt) {
    const int N = t.N;
    const size_t n = (size_t)N * N;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % N), j = (int)(q / N);
        const double hi = hw_of<double>(t, i), hj = hw_of<double>(t, j);
        double v = f ? -rt.h * rt.h * hi * hj * f[q] : 0.0;
        if (gW && j < rt.nbs) v += rt.bW * hi * dir_w(rt, j) * gW[i];
        if (gE && N - 1 - j < rt.nbs) v += rt.bE * hi * dir_w(rt, N - 1 - j) * gE[i];
        if (gS && i == 0) v += rt.h * hj * gS[j];
        if (gN && i == N - 1) v += rt.h * hj * gN[j];
        b[q] = v;
    }
}

// Adjoint source w = H~ (u - u_obs) and block partials of (u - u_obs)' H~ (u - u_obs).
__global__ __launch_bounds__(256) void adjoint_source(const double* __restrict__ u, const double* __restrict__ uo,
                                                     double* __restrict__ w, OpTab t, double h2, double* partial) {
    const size_t n = (size_t)t.N * t.N;
    double acc = 0.0;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % t.N), j = (int)(q / t.N);
        const double e = u[q] - uo[q];
        const double we = h2 * hw_of<double>(t, i) * hw_of<double>(t, j) * e;
        w[q] = we;
        acc += we * e;
    }
    const double s = block_sum<double, 256>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// df = -h^2 hw_i hw_j lam
__global__ __launch_bounds__(256) void rhs_t_f(const double* lam, double* df, OpTab t, RhsTab rt) {
    const int N = t.N;
    const size_t n = (size_t)N * N;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const int i = (int)(q % N), j = (int)(q / N);
        df[q] = -rt.h * rt.h * hw_of<double>(t, i) * hw_of<double>(t, j) * lam[q];
    }
}

// dgW(i) = hw_i sum_k dir_w(k) lam(i,k); dgE mirrored; dgS(j) = h hw_j lam(0,j); dgN(j) = h hw_j lam(N-1,j)
__global__ __launch_bounds__(256) void rhs_t_g(const double* lam, double* dgW, double* dgE, double* dgS,
                                              double* dgN, OpTab t, RhsTab rt) {
    const int N = t.N;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= N) return;
    const double hq = hw_of<double>(t, q);
    if (dgW) {
        double s = 0.0;
        for (int k = 0; k < rt.nbs; ++k) s += dir_w(rt, k) * lam[gidx(q, k, N)];
        dgW[q] = rt.bW * hq * s;
    }
    if (dgE) {
        double s = 0.0;
        for (int k = 0; k < rt.nbs; ++k) s += dir_w(rt, k) * lam[gidx(q, N - 1 - k, N)];
        dgE[q] = rt.bE * hq * s;
    }
    if (dgS) dgS[q] = rt.h * hq * lam[gidx(0, q, N)];
    if (dgN) dgN[q] = rt.h * hq * lam[gidx(N - 1, q, N)];
}

// ---------------------------------------------------------------------------
// CSR comparison path
// ---------------------------------------------------------------------------
// kCsrLanes lanes per row (rows hold at most 11 nonzeros), so a wave reads
// 8 rows' values and column indices as contiguous runs; a grid-stride loop
// over rows keeps the number of blocks (and block partials of x'Ax) bounded.
constexpr int kCsrLanes = 8;
constexpr int kCsrMaxBlocks = 4096;
__global__ __launch_bounds__(256) void csr_spmv(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                               const double* __restrict__ val, const double* __restrict__ x,
                                               double* __restrict__ y, int nrows, double* partial, const Scalars* sc) {
    if (sc && sc->done) return;
    const int lane = threadIdx.x & (kCsrLanes - 1);
    const int ng = gridDim.x * (256 / kCsrLanes);
    double acc = 0.0;
    for (int r = (blockIdx.x * 256 + threadIdx.x) / kCsrLanes; r < nrows; r += ng) {
        const int k1 = rp[r + 1];
        double s = 0.0;
        for (int k = rp[r] + lane; k < k1; k += kCsrLanes) s += val[k] * x[col[k]];
#pragma unroll
        for (int off = kCsrLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kCsrLanes);
        if (lane == 0) {
            y[r] = s;
            acc += x[r] * s;
        }
    }
    if (partial) {
        double bsum = block_sum<double, 256>(acc);
        if (threadIdx.x == 0) partial[blockIdx.x] = bsum;
    }
}

// r = b - A x with block partials of r'r.
__global__ __launch_bounds__(256) void csr_resid(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ x,
                                                const double* __restrict__ b, double* __restrict__ r, int nrows,
                                                double* partial) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (q < nrows) {
        double s = 0.0;
        for (int k = rp[q]; k < rp[q + 1]; ++k) s += val[k] * x[col[k]];
        const double rv = b[q] - s;
        r[q] = rv;
        acc = rv * rv;
    }
    double bsum = block_sum<double, 256>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = bsum;
}

// Fills the CSR arrays of A row by row from the tensor form (rows in natural order).
__global__ __launch_bounds__(256) void csr_fill(int32_t* rp, int32_t* col, double* val, OpTab t, int count_only,
                                               int32_t* counts) {
    const int N = t.N;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= N * N) return;
    const int i = q % N, j = q / N;
    // Collect (column, value) pairs for row (i,j): k-part along j, m-part along i.
    int cj[kMaxC + 5]; double vj[kMaxC + 5]; int nj = 0;
    int ci[kMaxC + 5]; double vi[kMaxC + 5]; int ni = 0;
    auto row1d = [&](int r, int* cc, double* vv, int& nn, bool is_k) {
        if (r < t.R) {
            for (int c = 0; c < t.C; ++c) { double v = is_k ? t.dev->kb[r][c] : t.dev->mb[r][c]; if (v != 0.0) { cc[nn] = c; vv[nn++] = v; } }
        } else if (r >= N - t.R) {
            const int rr = N - 1 - r;
            for (int c = t.C - 1; c >= 0; --c) { double v = is_k ? t.dev->kb[rr][c] : t.dev->mb[rr][c]; if (v != 0.0) { cc[nn] = N - 1 - c; vv[nn++] = v; } }
        } else {
            for (int s = -t.H; s <= t.H; ++s) {
                cc[nn] = r + s;
                vv[nn++] = t.cint[s < 0 ? -s : s];
            }
        }
    };
    row1d(j, cj, vj, nj, true);
    row1d(i, ci, vi, ni, false);
    const double hwi = hw_of<double>(t, i), hwj = hw_of<double>(t, j);
    // Merge into sorted global columns: entries (i, cj[k]) -> col = i + cj*N; (ci[k], j) -> ci + j*N.
    int cnt = 0;
    int outc[2 * (kMaxC + 5)]; double outv[2 * (kMaxC + 5)];
    // k-part columns with cj < j come first (they are < j*N), then row j block, then cj > j.
    for (int a = 0; a < nj; ++a) {
        if (cj[a] < j) { outc[cnt] = i + cj[a] * N; outv[cnt++] = hwi * vj[a]; }
    }
    double diagk = 0.0;
    for (int a = 0; a < nj; ++a) if (cj[a] == j) diagk = hwi * vj[a];
    for (int a = 0; a < ni; ++a) {
        double v = hwj * vi[a];
        if (ci[a] == i) v += diagk;
        outc[cnt] = ci[a] + j * N; outv[cnt++] = v;
    }
    for (int a = 0; a < nj; ++a) {
        if (cj[a] > j) { outc[cnt] = i + cj[a] * N; outv[cnt++] = hwi * vj[a]; }
    }
    if (count_only) { counts[q] = cnt; return; }
    const int base = rp[q];
    for (int a = 0; a < cnt; ++a) { col[base + a] = outc[a]; val[base + a] = outv[a]; }
}

}  // namespace sbp
