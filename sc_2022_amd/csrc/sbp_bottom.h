// sbp_bottom.h -- the coarse end of the V-cycle in one workgroup.
//
// On the small levels of the hierarchy (sides <= 65) a grid function is at
// most 33 KB, and a stencil pass over it is a few microseconds of dependent
// memory round trips plus the launch: at these sizes the cost of a level is
// the number of kernels, not the bytes.  This kernel runs everything below a
// given level -- pre-smoothing, residual, full-weighting restriction, the
// coarsest direct solve, prolongation and post-smoothing -- in a single
// workgroup per problem, with every level's right-hand side and iterates held
// in LDS (160 KB per CU on gfx950), so the whole bottom of the V-cycle is one
// launch.  The arithmetic per point is the same as the tiled passes' general
// path (difference-form rows, division by the diagonal).
#pragma once
#include "sbp_kernels.h"

namespace sbp {

constexpr int kBottomN = 65;          // largest side handled here
constexpr int kBottomThreads = 1024;
constexpr int kBottomMaxLevels = 5;   // 65, 33, 17, 9, 5 (the LDS budget: five tables and the pool)
constexpr int kBottomMaxCoarse = 289; // coarsest system size handled here (17 x 17)
constexpr int kCoarseChunks = 3;      // the coarse product is split three ways over k
// Three grid functions per level (b and two iterates) for a 65 -> 5 chain.
constexpr int kBottomPool = 3 * (65 * 65 + 33 * 33 + 17 * 17 + 9 * 9 + 5 * 5);

struct BottomCtl {
    const OpTab* tab[kBottomMaxLevels];   // device tables, finest first
    int nlev;                             // levels handled; the last one is solved directly
    const double* Ainv;                   // coarsest inverse, n_c x n_c
    int nu1, nu2, use_diag;
    double omega, scale;
    const int* done;                      // nullable: per-problem flags, stride ds ints
    int ds;
    long long sb;                         // elements between problems on the first level
    long long ss;                         // doubles between problems' partial slots (1 used)
    double* partial;                      // r'z partial of the last pass (when dotv is given)
};

// Operand read from an LDS grid function.
template <typename T> struct InLds {
    const T* u;
    using Raw = T;
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab& t, int i, int j) const { return u[gidx(i, j, t.N)]; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab&, int, int, const Raw& r) const { return r; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab& t, int i, int j) const { return u[gidx(i, j, t.N)]; }
};

// Calls f(i, j, k = i + j N) for every point of an N x N level (N <= 128):
// wave w takes rows j = w, w + 16, ..., lane l columns i = l and l + 64.
template <class F>
__device__ __forceinline__ void for_points(int N, F&& f) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < N; j += kBottomThreads / 64)
        for (int i = lane; i < N; i += 64) f(i, j, j * N + i);
}

// dst = u + w D^-1 (b - A u) (or u + w (b - A u)), then a barrier.
template <int P, typename T>
__device__ __forceinline__ void bottom_sweep(const OpTab& t, int N, const T* u, const T* b, T* dst, T omega,
                                             int use_diag) {
    const InLds<T> in{u};
    for_points(N, [&](int i, int j, int k) {
        const T Au = au_point<P, T, true>(t, in, i, j);
        T c = omega * (b[k] - Au);
        if (use_diag) c = c / diag_of<T, true>(t, i, j);
        dst[k] = u[k] + c;
    });
    __syncthreads();
}

// Side of level l below a first level of side N0 (N = 2^k + 1 halves each level).
__device__ __forceinline__ int bottom_side(int N0, int l) { return ((N0 - 1) >> l) + 1; }

// Offset of level l's three grid functions in the LDS pool.
__device__ __forceinline__ int bottom_offset(int N0, int l) {
    int o = 0;
    for (int m = 0; m < l; ++m) {
        const int Nm = bottom_side(N0, m);
        o += 3 * Nm * Nm;
    }
    return o;
}

template <int P, typename T>
__global__ __launch_bounds__(kBottomThreads) void vcycle_bottom(const T* __restrict__ bg, T* __restrict__ xg,
                                                               const T* __restrict__ dotv, BottomCtl c) {
    const int q = blockIdx.x;
    if (c.done && c.done[(size_t)q * c.ds]) return;
    bg += q * c.sb;
    xg += q * c.sb;
    if (dotv) dotv += q * c.sb;
    __shared__ OpTab tabs[kBottomMaxLevels];
    __shared__ T pool[kBottomPool];
    __shared__ double cpart[kCoarseChunks][kBottomMaxCoarse];
    const int tid = threadIdx.x;
    const T om = (T)c.omega;
    // Tables into LDS (constant level indices: no indexed access into the
    // argument segment); lookups then read the LDS copies directly.
#pragma unroll
    for (int l = 0; l < kBottomMaxLevels; ++l) {
        if (l < c.nlev) {
            const unsigned long long* src = reinterpret_cast<const unsigned long long*>(c.tab[l]);
            unsigned long long* dst = reinterpret_cast<unsigned long long*>(&tabs[l]);
            for (int k = tid; k < (int)(sizeof(OpTab) / 8); k += kBottomThreads) dst[k] = src[k];
        }
    }
    __syncthreads();
    const int NF = tabs[0].N;
    // Level l: B = pool + off, U0 = B + n, U1 = U0 + n.
    auto Bp = [&](int l) { return pool + bottom_offset(NF, l); };
    auto Up = [&](int l, int s) {
        const int Nl = bottom_side(NF, l);
        return pool + bottom_offset(NF, l) + (1 + s) * Nl * Nl;
    };
    int curmask = 0;   // bit l: which iterate buffer of level l is current
    {
        const int n0 = NF * NF;
        T* B0 = Bp(0);
        for (int k = tid; k < n0; k += kBottomThreads) B0[k] = bg[k];
        __syncthreads();
    }
    const int L = c.nlev - 1;   // coarsest level index
    // ---- down: pre-smoothing, residual, restriction ----
    for (int l = 0; l < L; ++l) {
        const OpTab& t = tabs[l];
        const int N = bottom_side(NF, l);
        const T* B = Bp(l);
        int s = 0;
        T* X = Up(l, 0);
        if (c.nu1 == 0) {
            for_points(N, [&](int, int, int k) { X[k] = T(0); });
        } else {
            for_points(N, [&](int i, int j, int k) {
                const T v = om * B[k];
                X[k] = c.use_diag ? v / diag_of<T, true>(t, i, j) : v;
            });
        }
        __syncthreads();
        for (int sw = 1; sw < c.nu1; ++sw) {
            bottom_sweep<P, T>(t, N, Up(l, s), B, Up(l, s ^ 1), om, c.use_diag);
            s ^= 1;
        }
        curmask |= s << l;
        // residual into the free buffer
        {
            const InLds<T> in{Up(l, s)};
            T* Rr = Up(l, s ^ 1);
            for_points(N, [&](int i, int j, int k) { Rr[k] = B[k] - au_point<P, T, true>(t, in, i, j); });
            __syncthreads();
        }
        // full weighting [1 2 1; 2 4 2; 1 2 1] / 16 (times scale) with zero padding
        {
            const T* Rr = Up(l, s ^ 1);
            const int Nc = bottom_side(NF, l + 1);
            T* Bc = Bp(l + 1);
            const T sc = (T)c.scale;
            auto rv = [&](int i, int j) { return (i >= 0 && i < N && j >= 0 && j < N) ? Rr[gidx(i, j, N)] : T(0); };
            for_points(Nc, [&](int I, int Jc, int o) {
                const int fi = 2 * I, fj = 2 * Jc;
                const T s0 = rv(fi, fj) + T(0.5) * (rv(fi - 1, fj) + rv(fi + 1, fj));
                const T sm = rv(fi, fj - 1) + T(0.5) * (rv(fi - 1, fj - 1) + rv(fi + 1, fj - 1));
                const T sp = rv(fi, fj + 1) + T(0.5) * (rv(fi - 1, fj + 1) + rv(fi + 1, fj + 1));
                Bc[o] = sc * T(0.25) * (s0 + T(0.5) * (sm + sp));
            });
            __syncthreads();
        }
    }
    // ---- coarsest level: x = Ainv b in FP64 ----
    // Ainv is symmetric, so x_r = sum_k Ainv[k][r] b_k: thread (r, chunk) runs
    // down column r over a third of k (loads coalesced across r, all
    // independent), then the three partial sums are added.
    {
        const int Ncs = bottom_side(NF, L);
        const int nc = Ncs * Ncs;
        const T* Bc = Bp(L);
        T* Xc = Up(L, 0);
        const int kc = (nc + kCoarseChunks - 1) / kCoarseChunks;
        for (int w = tid; w < kCoarseChunks * nc; w += kBottomThreads) {
            const int ch = w / nc, r = w - ch * nc;
            const int k0 = ch * kc, k1 = k0 + kc < nc ? k0 + kc : nc;
            double s = 0.0;
#pragma unroll 8
            for (int k = k0; k < k1; ++k) s += c.Ainv[(size_t)k * nc + r] * (double)Bc[k];
            cpart[ch][r] = s;
        }
        __syncthreads();
        for (int r = tid; r < nc; r += kBottomThreads) {
            double s = cpart[0][r];
#pragma unroll
            for (int ch = 1; ch < kCoarseChunks; ++ch) s += cpart[ch][r];
            Xc[r] = (T)s;
        }
        __syncthreads();
    }
    // ---- up: prolongation folded into the first post-smoothing sweep ----
    for (int l = L - 1; l >= 0; --l) {
        const OpTab& t = tabs[l];
        const int N = bottom_side(NF, l), Nc = bottom_side(NF, l + 1);
        const T* B = Bp(l);
        const T* E = Up(l + 1, (curmask >> (l + 1)) & 1);
        int s = (curmask >> l) & 1;
        // operand x + P e into the free buffer
        {
            const T* X = Up(l, s);
            T* Op = Up(l, s ^ 1);
            for_points(N, [&](int i, int j, int k) {
                const int ic = i >> 1, jc = j >> 1;
                const int ic1 = ic + 1 < Nc ? ic + 1 : ic, jc1 = jc + 1 < Nc ? jc + 1 : jc;
                const T e00 = E[gidx(ic, jc, Nc)], e10 = E[gidx(ic1, jc, Nc)];
                const T e01 = E[gidx(ic, jc1, Nc)], e11 = E[gidx(ic1, jc1, Nc)];
                T v;
                if (!(i & 1)) v = (!(j & 1)) ? e00 : (e00 + e01) * T(0.5);
                else v = (!(j & 1)) ? (e00 + e10) * T(0.5) : (e00 + e10 + e01 + e11) * T(0.25);
                Op[k] = X[k] + v;
            });
            __syncthreads();
            s ^= 1;
        }
        for (int sw = 0; sw < c.nu2; ++sw) {
            bottom_sweep<P, T>(t, N, Up(l, s), B, Up(l, s ^ 1), om, c.use_diag);
            s ^= 1;
        }
        curmask = (curmask & ~(1 << l)) | (s << l);
    }
    // ---- result (and dotv' x) ----
    {
        const int n0 = NF * NF;
        const T* X = Up(0, curmask & 1);
        T acc = T(0);
        for (int k = tid; k < n0; k += kBottomThreads) {
            const T v = X[k];
            xg[k] = v;
            if (dotv) acc += dotv[k] * v;
        }
        if (dotv) {
            const T s = block_sum<T, kBottomThreads>(acc);
            if (tid == 0) c.partial[(size_t)q * c.ss] = (double)s;
        }
    }
}

}  // namespace sbp
