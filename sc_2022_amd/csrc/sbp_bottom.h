// sbp_bottom.h -- the small levels of the V-cycle in one workgroup per problem.
//
// On the levels of side <= 65 a grid function is at most 34 KB and a stencil
// pass over it is a few microseconds of dependent memory round trips plus a
// launch: there the cost of a level is the number of kernels, not the bytes.
// These two kernels run the levels from a given one down to (not including)
// the coarsest in a single workgroup per problem, every level's data in LDS
// (160 KB per CU on gfx950):
//
//   vcycle_bottom_down  pre-smoothing, residual and full-weighting restriction
//                       of every level; stores each level's pre-smoothed
//                       iterate x2 and the restricted right-hand sides
//   vcycle_bottom_up    prolongation of the coarser correction and the
//                       post-smoothing of every level, finest last
//
// The coarsest direct solve between them stays the multi-block coarse_solve
// kernel (its dense inverse is read by many CUs at once; one CU pulling it
// from memory would take longer than the levels themselves).  So the levels
// of side <= 65 over a 17 x 17 coarsest level cost three launches instead of
// five.  Each pass visits the interior points (both indices off the closure
// rows: the plain cross, a constant inverse diagonal) and the closure frame in
// separate loops, so no wave mixes the two forms.  The arithmetic per point is
// that of the fused two-sweep kernels (au_interior / au_point, smooth_corr_at,
// the same restriction and prolongation formulas), so the results agree with
// them bit for bit.
#pragma once
#include "sbp_kernels.h"

namespace sbp {

constexpr int kBottomN = 65;            // largest side handled here
constexpr int kBottomThreads = 1024;
constexpr int kBottomMaxLevels = 3;     // levels above the coarsest: 65, 33, 17
// LDS pool (elements): the down kernel holds b and two iterates per level;
// the up kernel x2, b and a sweep buffer per level, the output of every level
// below the first, and the coarsest correction.  Sized for the levels 65, 33,
// 17 over a 9 x 9 coarsest level (the largest of the supported chains).
constexpr int kBottomPool = 3 * 65 * 65 + 4 * 33 * 33 + 4 * 17 * 17 + 9 * 9;
constexpr int kBottomDownPre = (65 * 65 + kBottomThreads - 1) / kBottomThreads;   // b of the first level
constexpr int kBottomUpPre =
    (2 * (65 * 65 + 33 * 33 + 17 * 17) + 9 * 9 + kBottomThreads - 1) / kBottomThreads;

struct BottomCtl {
    const OpTab* tab[kBottomMaxLevels];   // device tables of the levels, finest first
    int side[kBottomMaxLevels + 1];       // their sides, then the coarsest level's
    void* x2[kBottomMaxLevels];           // pre-smoothed iterate per level (down writes, up reads)
    void* b[kBottomMaxLevels + 1];        // right-hand sides: b[0] the caller's, b[nlev] the coarsest's
    const void* ec;                       // the coarsest level's solution (up reads)
    long long sz[kBottomMaxLevels + 1];   // elements between problems on each level
    int nlev;                             // levels handled (above the coarsest)
    int nu1, nu2, use_diag;
    double omega, scale;
    const int* done;                      // nullable: per-problem flags, stride ds ints
    int ds;
    long long ss;                         // doubles between problems' partial slots (1 used)
    double* partial;                      // dotv' out partial of the up kernel (when dotv is given)
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

// Calls f(std::true_type, i, j) for the interior points of an N x N level
// (i, j in [R, N - R)), then f(std::false_type, i, j) for the 4 R (N - R)
// points of the closure frame, each set spread over the workgroup.
template <class F>
__device__ __forceinline__ void bottom_points(int N, int R, F&& f) {
    const int NI = N - 2 * R;
    const float rNI = 1.0f / (float)NI, rN = 1.0f / (float)N, r2R = 1.0f / (float)(2 * R);
    for (int m = threadIdx.x; m < NI * NI; m += kBottomThreads) {
        const int r = div_small(m, NI, rNI);
        f(std::true_type{}, R + (m - r * NI), R + r);
    }
    const int nrow = 2 * R * N, nf = nrow + 2 * R * NI;
    for (int m = threadIdx.x; m < nf; m += kBottomThreads) {
        int i, j;
        if (m < nrow) {   // rows 0 .. R-1 and N-R .. N-1
            const int r = div_small(m, N, rN);
            i = m - r * N;
            j = r < R ? r : r + NI;
        } else {          // columns 0 .. R-1 and N-R .. N-1 of the rows between
            const int m2 = m - nrow;
            const int r = div_small(m2, 2 * R, r2R);
            const int c = m2 - r * 2 * R;
            j = R + r;
            i = c < R ? c : c + NI;
        }
        f(std::false_type{}, i, j);
    }
}

// The interior cross at point k of an LDS grid function of row stride N (the
// order of operations of au_interior).
template <int P, typename T>
__device__ __forceinline__ T au_cross(const T* a, int k, int N) {
    constexpr int H = Stencil<P>::H;
    const T uc = a[k];
    T kp = Stencil<P>::template c<T>(1) * ((a[k - N] - uc) + (a[k + N] - uc));
#pragma unroll
    for (int s = 2; s <= H; ++s) kp += Stencil<P>::template c<T>(s) * ((a[k - s * N] - uc) + (a[k + s * N] - uc));
    T mp = Stencil<P>::template c<T>(1) * ((a[k - 1] - uc) + (a[k + 1] - uc));
#pragma unroll
    for (int s = 2; s <= H; ++s) mp += Stencil<P>::template c<T>(s) * ((a[k - s] - uc) + (a[k + s] - uc));
    return kp + mp;
}

// (A u)(i, j) on an LDS grid function: the cross inside, the closure form on the frame.
template <int P, typename T, bool INT>
__device__ __forceinline__ T au_lds(const OpTab& t, const T* u, int i, int j) {
    if constexpr (INT) return au_cross<P, T>(u, j * t.N + i, t.N);
    else return au_point<P, T, true>(t, InLds<T>{u}, i, j);
}

// put(k, u + w D^-1 (b - A u)) at every point of level t (u, b in LDS; D^-1
// the constant interior inverse diagonal inside), then a barrier.
template <int P, typename T, class PUT>
__device__ __forceinline__ void bottom_sweep(const OpTab& t, const T* u, const T* b, T om, int use_diag, PUT&& put) {
    bottom_points(t.N, t.R, [&](auto in_, int i, int j) {
        constexpr bool INT = decltype(in_)::value;
        const int k = j * t.N + i;
        put(k, u[k] + smooth_corr_at<T, INT, true>(t, i, j, b[k] - au_lds<P, T, INT>(t, u, i, j), om, use_diag));
    });
    __syncthreads();
}

// Device tables of the handled levels into LDS (word copies, every thread).
__device__ __forceinline__ void bottom_tables(const BottomCtl& c, OpTab* tabs) {
    constexpr int W = (int)(sizeof(OpTab) / 8);
#pragma unroll
    for (int l = 0; l < kBottomMaxLevels; ++l) {
        if (l < c.nlev) {
            const unsigned long long* src = reinterpret_cast<const unsigned long long*>(c.tab[l]);
            unsigned long long* dst = reinterpret_cast<unsigned long long*>(&tabs[l]);
            for (int k = (int)threadIdx.x; k < W; k += kBottomThreads) dst[k] = src[k];
        }
    }
}

// Down: for each level d (side N), from b_d: x2 = S^nu1(0) (stored to c.x2[d]),
// the residual b_d - A x2 in place over b_d, and its full weighting (times
// scale) into b_{d+1} -- LDS and memory below the first level, memory only for
// the coarsest.
template <int P, typename T>
__global__ __launch_bounds__(kBottomThreads) void vcycle_bottom_down(BottomCtl c) {
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    __shared__ OpTab tabs[kBottomMaxLevels];
    __shared__ T pool[kBottomPool];
    const int dflag = c.done ? c.done[(size_t)q * c.ds] : 0;
    const T* bg = (const T*)c.b[0] + q * c.sz[0];
    {
        const int n0 = c.side[0] * c.side[0];
        T v[kBottomDownPre];
#pragma unroll
        for (int m = 0; m < kBottomDownPre; ++m) {
            const int k = tid + m * kBottomThreads;
            if (k < n0) v[m] = bg[k];
        }
        if (dflag) return;
        bottom_tables(c, tabs);
#pragma unroll
        for (int m = 0; m < kBottomDownPre; ++m) {
            const int k = tid + m * kBottomThreads;
            if (k < n0) pool[k] = v[m];
        }
    }
    __syncthreads();
    const T om = (T)c.omega;
    int off = 0;
    // (levels unrolled: the per-level fields of c are read at constant indices)
#pragma unroll
    for (int d = 0; d < kBottomMaxLevels; ++d) {
        if (d >= c.nlev) break;
        const OpTab& t = tabs[d];
        const int N = t.N, n = N * N;
        T* B = pool + off;
        T* X0 = B + n;
        T* X1 = B + 2 * n;
        int s = 0;
        if (c.nu1 == 0) {
            for (int k = tid; k < n; k += kBottomThreads) X0[k] = T(0);
        } else {
            bottom_points(N, t.R, [&](auto in_, int i, int j) {
                constexpr bool INT = decltype(in_)::value;
                const int k = j * N + i;
                X0[k] = smooth_corr_at<T, INT, true>(t, i, j, B[k], om, c.use_diag);
            });
        }
        __syncthreads();
        for (int sw = 1; sw < c.nu1; ++sw) {
            T* dst = s ? X0 : X1;
            bottom_sweep<P, T>(t, s ? X1 : X0, B, om, c.use_diag, [&](int k, T v) { dst[k] = v; });
            s ^= 1;
        }
        // x2 to memory; residual in place (each point reads b at itself only)
        {
            const T* x = s ? X1 : X0;
            T* xg = (T*)c.x2[d] + q * c.sz[d];
            bottom_points(N, t.R, [&](auto in_, int i, int j) {
                constexpr bool INT = decltype(in_)::value;
                const int k = j * N + i;
                xg[k] = x[k];
                B[k] = B[k] - au_lds<P, T, INT>(t, x, i, j);
            });
            __syncthreads();
        }
        // full weighting [1 2 1; 2 4 2; 1 2 1] / 16 (times scale) with zero padding
        {
            const int Nc = c.side[d + 1], nc = Nc * Nc;
            T* Bn = B + 3 * n;   // the next level's b (unused below the last level)
            const bool lds = d + 1 < c.nlev;
            T* bn = (T*)c.b[d + 1] + q * c.sz[d + 1];
            const T sc = (T)c.scale;
            const float rNc = 1.0f / (float)Nc;
            auto rv = [&](int i, int j) { return (i >= 0 && i < N && j >= 0 && j < N) ? B[j * N + i] : T(0); };
            for (int o = tid; o < nc; o += kBottomThreads) {
                const int Jc = div_small(o, Nc, rNc), I = o - Jc * Nc;
                const int fi = 2 * I, fj = 2 * Jc;
                const T s0 = rv(fi, fj) + T(0.5) * (rv(fi - 1, fj) + rv(fi + 1, fj));
                const T sm = rv(fi, fj - 1) + T(0.5) * (rv(fi - 1, fj - 1) + rv(fi + 1, fj - 1));
                const T sp = rv(fi, fj + 1) + T(0.5) * (rv(fi - 1, fj + 1) + rv(fi + 1, fj + 1));
                const T v = sc * T(0.25) * (s0 + T(0.5) * (sm + sp));
                bn[o] = v;
                if (lds) Bn[o] = v;
            }
            __syncthreads();
        }
        off += 3 * n;
    }
}

// Up: for each level d from the coarsest up, u = x2 + P e (e the coarser
// level's output, or the coarsest solution), then nu2 sweeps; the first
// level's result goes to out (with dotv: the partial dotv' out), the others'
// stay in LDS for the next finer level.  Every load from memory (x2 and b of
// all levels, the coarsest solution) is issued before the first use.
template <int P, typename T>
__global__ __launch_bounds__(kBottomThreads) void vcycle_bottom_up(T* __restrict__ out, const T* __restrict__ dotv,
                                                                  BottomCtl c) {
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    __shared__ OpTab tabs[kBottomMaxLevels];
    __shared__ T pool[kBottomPool];
    const int L = c.nlev;
    // LDS layout: level d at off[d]: X (x2, then the operand), B, V, and for
    // d > 0 O (the level's output); the coarsest solution at offE.  (Per-level
    // values are picked at constant indices: no indexed access to c.)
    int off[kBottomMaxLevels];
    int offE = 0, NE = c.side[0];
    long long szE = c.sz[0];
#pragma unroll
    for (int d = 0; d < kBottomMaxLevels; ++d) {
        off[d] = offE;
        if (d < L) offE += (d == 0 ? 3 : 4) * c.side[d] * c.side[d];
        if (d + 1 == L) { NE = c.side[d + 1]; szE = c.sz[d + 1]; }
    }
    {
        // segments: x2 and b of each level, then the coarsest solution
        const T* src[2 * kBottomMaxLevels + 1];
        int dst[2 * kBottomMaxLevels + 1], end[2 * kBottomMaxLevels + 1];
        int tot = 0;
#pragma unroll
        for (int d = 0; d < kBottomMaxLevels; ++d) {
            const int n = d < L ? c.side[d] * c.side[d] : 0;
            src[2 * d] = (const T*)c.x2[d] + (d < L ? q * c.sz[d] : 0);
            dst[2 * d] = off[d];
            end[2 * d] = tot += n;
            src[2 * d + 1] = (const T*)c.b[d] + (d < L ? q * c.sz[d] : 0);
            dst[2 * d + 1] = off[d] + n;
            end[2 * d + 1] = tot += n;
        }
        src[2 * kBottomMaxLevels] = (const T*)c.ec + q * szE;
        dst[2 * kBottomMaxLevels] = offE;
        end[2 * kBottomMaxLevels] = tot += NE * NE;
        const int dflag = c.done ? c.done[(size_t)q * c.ds] : 0;
        T v[kBottomUpPre];
        int at[kBottomUpPre];
#pragma unroll
        for (int m = 0; m < kBottomUpPre; ++m) {
            const int k = tid + m * kBottomThreads;
            at[m] = -1;
            if (k < tot) {
                const T* sp = src[0];
                int dp = dst[0], beg = 0;
#pragma unroll
                for (int e = 0; e < 2 * kBottomMaxLevels; ++e)
                    if (k >= end[e]) { sp = src[e + 1]; dp = dst[e + 1]; beg = end[e]; }
                v[m] = sp[k - beg];
                at[m] = dp + (k - beg);
            }
        }
        if (dflag) return;
        bottom_tables(c, tabs);
#pragma unroll
        for (int m = 0; m < kBottomUpPre; ++m)
            if (at[m] >= 0) pool[at[m]] = v[m];
    }
    __syncthreads();
    const T om = (T)c.omega;
    T acc = T(0);
    // (levels unrolled, coarsest first: the per-level fields of c and off are
    // read at constant indices)
#pragma unroll
    for (int dd = 0; dd < kBottomMaxLevels; ++dd) {
        const int d = kBottomMaxLevels - 1 - dd;
        if (d >= L) continue;
        const OpTab& t = tabs[d];
        const int N = t.N, n = N * N;
        const int Nc = c.side[d + 1];
        T* X = pool + off[d];
        T* B = X + n;
        T* V = B + n;
        T* O = V + n;                                    // d > 0
        const T* E = pool + (d + 1 < L ? off[d + 1 < kBottomMaxLevels ? d + 1 : d] + 3 * Nc * Nc : offE);
        T* og = out + q * c.sz[0];                       // d == 0
        const T* dv = dotv ? dotv + q * c.sz[0] : nullptr;
        auto put = [&](int k, T v) {
            if (d > 0) {
                O[k] = v;
            } else {
                og[k] = v;
                if (dv) acc += dv[k] * v;
            }
        };
        // operand x2 + P e in place over x2 (each point reads x2 at itself only)
        {
            const float rN = 1.0f / (float)N;
            for (int k = tid; k < n; k += kBottomThreads) {
                const int j = div_small(k, N, rN), i = k - j * N;
                const int ic = i >> 1, jc = j >> 1;
                const int ic1 = ic + 1 < Nc ? ic + 1 : ic, jc1 = jc + 1 < Nc ? jc + 1 : jc;
                const T e00 = E[jc * Nc + ic], e10 = E[jc * Nc + ic1];
                const T e01 = E[jc1 * Nc + ic], e11 = E[jc1 * Nc + ic1];
                T pv;
                if (!(i & 1)) pv = (!(j & 1)) ? e00 : (e00 + e01) * T(0.5);
                else pv = (!(j & 1)) ? (e00 + e10) * T(0.5) : (e00 + e10 + e01 + e11) * T(0.25);
                X[k] = X[k] + pv;
            }
            __syncthreads();
        }
        if (c.nu2 == 0) {
            for (int k = tid; k < n; k += kBottomThreads) put(k, X[k]);
            __syncthreads();
            continue;
        }
        int s = 0;
        for (int sw = 0; sw + 1 < c.nu2; ++sw) {
            T* dst = s ? X : V;
            bottom_sweep<P, T>(t, s ? V : X, B, om, c.use_diag, [&](int k, T v) { dst[k] = v; });
            s ^= 1;
        }
        bottom_sweep<P, T>(t, s ? V : X, B, om, c.use_diag, put);
    }
    if (dotv) {
        const T s = block_sum<T, kBottomThreads>(acc);
        if (tid == 0) c.partial[(size_t)q * c.ss] = (double)s;
    }
}

}  // namespace sbp
