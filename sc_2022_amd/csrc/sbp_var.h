// sbp_var.h -- the variable-coefficient operator, orders 2 and 4.
//
// d/dx (bx d/dx) + d/dy (by d/dy) with one coefficient profile per axis
// (bx along the slow, Dirichlet index j; by along the fast, Neumann index i),
// from the reference's variable_diagonal_sbp_D2 (diagonal_sbp.jl:470-740)
// used on both axes as its get_operators does (legacy_functions.jl:100-174).
// The tensor form of the constant case carries over:
//     (A u)(i,j) = hw[i] sum_j' k[j][j'] u(i,j') + hw[j] sum_i' m[i][i'] u(i',j)
// with m = h M(by) (the reference's M before its division by h) and
// k = m(bx) - (E bs + bs^T E) + 13 b_face E, bs the b_face-weighted boundary
// row and the Dirichlet penalty scaled by the face coefficient (the
// reference's tau = -13/h at b = 1).  The factors are no longer Toeplitz:
// every interior row has its own 2H+1 coefficients (stored offset-major,
// N per offset, so a wavefront reads them coalesced), and the two faces have
// different closure blocks.  H weights do not depend on b.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "sbp_kernels.h"

namespace sbp {

// One axis' factor on the device: closure rows (first / last R) and the band
// of the rows between.
struct VarAxis {
    double lo[kMaxR][kMaxC];   // rows r < R, columns c < C
    double hi[kMaxR][kMaxC];   // rows N-1-r, columns N-1-c
    const double* band;        // (2H+1) x N, offset s = -H..H at band[(s+H) N + r]
};

struct VarTab {
    int N, R, C, H, RH;
    double hw[kMaxR];
    VarAxis k, m;
};

// ---------------------------------------------------------------------------
// Host: the entries of M(b) (dimensionless, i.e. h M), restated from
// diagonal_sbp.jl:470-566 with 1-based indices as the reference writes them.
// ---------------------------------------------------------------------------
// p = 4 boundary block: entry (r, c), r <= c, is sum num/den * b_k; the
// integer literals are divided in double precision as the reference does.
struct VarTerm { double num, den; int k; };
struct VarEntry { int r, c; VarTerm t[6]; int nt; };

inline const std::vector<VarEntry>& var_p4_block() {
    static const std::vector<VarEntry> e = {
        {1, 1, {{12, 17, 1}, {59, 192, 2}, {27010400129.0, 345067064608.0, 3}, {69462376031.0, 2070402387648.0, 4}}, 4},
        {1, 2, {{-59, 68, 1}, {-6025413881.0, 21126554976.0, 3}, {-537416663.0, 7042184992.0, 4}}, 3},
        {1, 3, {{2, 17, 1}, {-59, 192, 2}, {213318005.0, 16049630912.0, 4}, {2083938599.0, 8024815456.0, 3}}, 4},
        {1, 4, {{3, 68, 1}, {-1244724001.0, 21126554976.0, 3}, {752806667.0, 21126554976.0, 4}}, 3},
        {1, 5, {{49579087.0, 10149031312.0, 3}, {-49579087.0, 10149031312.0, 4}}, 2},
        {1, 6, {{-1, 784, 4}, {1, 784, 3}}, 2},
        {2, 2, {{3481, 3264, 1}, {9258282831623875.0, 7669235228057664.0, 3}, {236024329996203.0, 1278205871342944.0, 4}}, 3},
        {2, 3, {{-59, 408, 1}, {-29294615794607.0, 29725717938208.0, 3}, {-2944673881023.0, 29725717938208.0, 4}}, 3},
        {2, 4, {{-59, 1088, 1}, {260297319232891.0, 2556411742685888.0, 3}, {-60834186813841.0, 1278205871342944.0, 4}}, 3},
        {2, 5, {{-1328188692663.0, 37594290333616.0, 3}, {1328188692663.0, 37594290333616.0, 4}}, 2},
        {2, 6, {{-8673, 2904112, 3}, {8673, 2904112, 4}}, 2},
        {3, 3, {{1, 51, 1}, {59, 192, 2}, {13777050223300597.0, 26218083221499456.0, 4}, {564461, 13384296, 5},
                {378288882302546512209.0, 270764341349677687456.0, 3}}, 5},
        {3, 4, {{1, 136, 1}, {-125059, 743572, 5}, {-4836340090442187227.0, 5525802884687299744.0, 3},
                {-17220493277981.0, 89177153814624.0, 4}}, 4},
        {3, 5, {{-10532412077335.0, 42840005263888.0, 4}, {1613976761032884305.0, 7963657098519931984.0, 3},
                {564461, 4461432, 5}}, 3},
        {3, 6, {{-960119, 1280713392, 4}, {-3391, 6692148, 5}, {33235054191.0, 26452850508784.0, 3}}, 3},
        {4, 4, {{3, 1088, 1}, {507284006600757858213.0, 475219048083107777984.0, 3}, {1869103, 2230716, 5}, {1, 24, 6},
                {1950062198436997.0, 3834617614028832.0, 4}}, 5},
        {4, 5, {{-4959271814984644613.0, 20965546238960637264.0, 3}, {-1, 6, 6}, {-15998714909649.0, 37594290333616.0, 4},
                {-375177, 743572, 5}}, 4},
        {4, 6, {{-368395, 2230716, 5}, {752806667.0, 539854092016.0, 3}, {1063649, 8712336, 4}, {1, 8, 6}}, 4},
        {5, 5, {{8386761355510099813.0, 128413970713633903242.0, 3}, {2224717261773437.0, 2763180339520776.0, 4},
                {5, 6, 6}, {1, 24, 7}, {280535, 371786, 5}}, 5},
        {5, 6, {{-35039615, 213452232, 4}, {-1, 6, 7}, {-13091810925.0, 13226425254392.0, 3}, {-1118749, 2230716, 5},
                {-1, 2, 6}}, 5},
        {6, 6, {{3290636, 80044587, 4}, {5580181, 6692148, 5}, {5, 6, 7}, {1, 24, 8},
                {660204843.0, 13226425254392.0, 3}, {3, 4, 6}}, 6},
    };
    return e;
}

// Calls put(row, col, value) (1-based) for every stored entry of h M(b)
// (duplicates add, as in the reference's sparse()).
template <class PUT>
inline void var_m_entries(int p, int Np, const double* B, PUT&& put) {
    const int N = Np - 1;
    auto b = [B](int k) { return B[k - 1]; };
    if (p == 2) {
        put(1, 1, (b(1) + b(2)) / 2);
        put(N + 1, N + 1, (b(N) + b(N + 1)) / 2);
        for (int i = 1; i <= N; ++i) put(i + 1, i, -(b(i) + b(i + 1)) / 2);
        for (int i = 2; i <= N; ++i) put(i, i, (b(i - 1) + 2 * b(i) + b(i + 1)) / 2);
        for (int i = 1; i <= N; ++i) put(i, i + 1, -(b(i) + b(i + 1)) / 2);
        return;
    }
    double blo[8], bhi[8];
    for (int k = 0; k < 8; ++k) { blo[k] = B[k]; bhi[k] = B[N - k]; }
    for (const VarEntry& e : var_p4_block()) {
        double vl = 0.0, vh = 0.0;
        for (int q = 0; q < e.nt; ++q) {
            const double w = e.t[q].num / e.t[q].den;
            vl += w * blo[e.t[q].k - 1];
            vh += w * bhi[e.t[q].k - 1];
        }
        put(e.r, e.c, vl);
        if (e.r != e.c) put(e.c, e.r, vl);
        put(N + 2 - e.r, N + 2 - e.c, vh);
        if (e.r != e.c) put(N + 2 - e.c, N + 2 - e.r, vh);
    }
    for (int J = 5; J <= N - 5; ++J) put(J + 2, J, (1.0 / 8) * b(J + 2) - (1.0 / 6) * b(J + 1) + (1.0 / 8) * b(J));
    for (int J = 6; J <= N - 5; ++J)
        put(J + 1, J, -(1.0 / 6) * b(J + 2) - (1.0 / 2) * b(J + 1) - (1.0 / 2) * b(J) - (1.0 / 6) * b(J - 1));
    for (int J = 7; J <= N - 5; ++J)
        put(J, J, (1.0 / 24) * b(J + 2) + (5.0 / 6) * b(J + 1) + (3.0 / 4) * b(J) + (5.0 / 6) * b(J - 1) +
                      (1.0 / 24) * b(J - 2));
    for (int J = 7; J <= N - 4; ++J)
        put(J - 1, J, -(1.0 / 6) * b(J + 1) - (1.0 / 2) * b(J) - (1.0 / 2) * b(J - 1) - (1.0 / 6) * b(J - 2));
    for (int J = 7; J <= N - 3; ++J) put(J - 2, J, (1.0 / 8) * b(J) - (1.0 / 6) * b(J - 1) + (1.0 / 8) * b(J - 2));
}

// Builds one axis' factor (m, or k when dirichlet) from the coefficient
// profile b (N values): closure blocks into ax, the band (2H+1) x N into
// band.  R / C: closure rows / columns (p = 2: 3 / 4, p = 4: 6 / 8).
// Returns false when N is too small or an entry falls outside the layout.
inline bool build_var_axis(int p, int N, const double* b, bool dirichlet, VarAxis* ax, std::vector<double>& band) {
    const int H = p / 2, R = p == 2 ? 3 : 6, C = p == 2 ? 4 : 8;
    if (N < 2 * C + 2) return false;
    std::memset(ax->lo, 0, sizeof(ax->lo));
    std::memset(ax->hi, 0, sizeof(ax->hi));
    band.assign((size_t)(2 * H + 1) * N, 0.0);
    bool ok = true;
    auto put0 = [&](int r, int c, double v) {   // 0-based
        if (r < R) {
            if (c >= C) { ok = false; return; }
            ax->lo[r][c] += v;
        } else if (r >= N - R) {
            const int rr = N - 1 - r, cc = N - 1 - c;
            if (cc < 0 || cc >= C) { ok = false; return; }
            ax->hi[rr][cc] += v;
        } else {
            const int s = c - r;
            if (s < -H || s > H) { ok = false; return; }
            band[(size_t)(s + H) * N + r] += v;
        }
    };
    var_m_entries(p, N, b, [&](int r, int c, double v) { put0(r - 1, c - 1, v); });
    if (dirichlet) {
        // k = m - (E bs + bs^T E) + 13 b_face E, bs = b_face * BS (dimensionless)
        const double bs2[3] = {3.0 / 2.0, -2.0, 1.0 / 2.0};
        const double bs4[4] = {11.0 / 6.0, -3.0, 3.0 / 2.0, -1.0 / 3.0};
        const double* bs = p == 2 ? bs2 : bs4;
        const int nbs = p == 2 ? 3 : 4;
        for (int c = 0; c < nbs; ++c) {
            put0(0, c, -b[0] * bs[c]);
            put0(c, 0, -b[0] * bs[c]);
            put0(N - 1, N - 1 - c, -b[N - 1] * bs[c]);
            put0(N - 1 - c, N - 1, -b[N - 1] * bs[c]);
        }
        put0(0, 0, 13.0 * b[0]);
        put0(N - 1, N - 1, 13.0 * b[N - 1]);
    }
    return ok;
}

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------
// sum_c a[r][c] u(c) along one axis for row r (load(c) gives u at index c).
template <int H, class LOAD>
__device__ __forceinline__ double var_row(const VarAxis& a, int R, int C, int N, int r, LOAD&& load) {
    double s = 0.0;
    if (r < R) {
        for (int c = 0; c < C; ++c) s += a.lo[r][c] * load(c);
    } else if (r >= N - R) {
        const int rr = N - 1 - r;
        for (int c = 0; c < C; ++c) s += a.hi[rr][c] * load(N - 1 - c);
    } else {
#pragma unroll
        for (int q = -H; q <= H; ++q) s += a.band[(size_t)(q + H) * N + r] * load(r + q);
    }
    return s;
}

// MODE 0: out = sign A u;  1: out = A u, partial += aux' out;  2: out = aux - A u, partial += out' out.
// A block: kVarThreads threads along i, kVarRows consecutive rows j.
constexpr int kVarThreads = 256;
constexpr int kVarRows = 8;

template <int MODE>
__device__ __forceinline__ void var_emit(double* out, const double* aux, double sign, size_t q, double Au,
                                         double& acc) {
    if constexpr (MODE == 0) {
        out[q] = sign * Au;
    } else if constexpr (MODE == 1) {
        out[q] = Au;
        acc += aux[q] * Au;
    } else {
        const double rv = aux[q] - Au;
        out[q] = rv;
        acc += rv * rv;
    }
}

// Tiled evaluation for blocks whose rows (with halo) avoid the closure rows:
// each thread keeps its column window u(i, j0-H .. j0+kVarRows-1+H) in
// registers for the j-direction, an LDS tile serves the i-direction, the
// i-band coefficients are loaded once per thread and the j-band ones once
// per row (uniform); H weights along j are 1 there.  With ICHK (the first
// and last column of blocks) each thread tests its column: out-of-grid
// threads idle, threads on the closure columns take the closure row of m
// (direct loads, table t staged in LDS) and their H weight.
template <int P, int MODE, bool ICHK>
__device__ __forceinline__ double var_tile(const double* __restrict__ u, double* __restrict__ out,
                                           const double* __restrict__ aux, double sign, const VarTab* __restrict__ tp,
                                           const VarTab& t, int i0, int j0,
                                           double (*tile)[kVarThreads + 2 * (P / 2)]) {
    constexpr int H = P / 2;
    constexpr int J = kVarRows;
    constexpr int W = J + 2 * H;
    const int N = tp->N, R = tp->R;
    const int tid = threadIdx.x;
    const int i = i0 + tid;
    const bool act = !ICHK || i < N;
    const bool iint = !ICHK || (i >= R && i < N - R);
    const double* kb = tp->k.band;
    const double* mb = tp->m.band;
    double w[W];
    double cm[2 * H + 1];
#pragma unroll
    for (int k = 0; k < W; ++k) w[k] = act ? u[gidx(i, j0 - H + k, N)] : 0.0;
#pragma unroll
    for (int q = 0; q <= 2 * H; ++q) cm[q] = iint ? mb[(size_t)q * N + i] : 0.0;
    if (tid < 2 * H * J) {   // halo columns: H left, H right, J rows
        const int side = tid / (H * J), rem = tid % (H * J), off = rem / J, r = rem % J;
        const int ii = side == 0 ? i0 - H + off : i0 + kVarThreads + off;
        tile[r][side == 0 ? off : kVarThreads + H + off] = (!ICHK || (ii >= 0 && ii < N)) ? u[gidx(ii, j0 + r, N)] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < J; ++r) tile[r][tid + H] = w[r + H];
    __syncthreads();
    double acc = 0.0;
    if (!act) return acc;
    double hwi = 1.0;
    if (ICHK) {
        const int ri = i < N - 1 - i ? i : N - 1 - i;
        hwi = ri < t.RH ? t.hw[ri] : 1.0;
    }
#pragma unroll
    for (int r = 0; r < J; ++r) {
        const int j = j0 + r;
        double kp = 0.0, mp = 0.0;
#pragma unroll
        for (int q = 0; q <= 2 * H; ++q) kp += kb[(size_t)q * N + j] * w[r + q];
        if (iint) {
#pragma unroll
            for (int q = 0; q <= 2 * H; ++q) mp += cm[q] * tile[r][tid + q];
        } else {
            mp = var_row<H>(t.m, t.R, t.C, N, i, [&](int c) { return u[gidx(c, j, N)]; });
        }
        var_emit<MODE>(out, aux, sign, gidx(i, j, N), hwi * kp + mp, acc);
    }
    return acc;
}

template <int P, int MODE>
__global__ __launch_bounds__(kVarThreads) void var_apply(const double* __restrict__ u, double* __restrict__ out,
                                                        const double* __restrict__ aux, double sign, double* partial,
                                                        const VarTab* __restrict__ tp, const Scalars* sc) {
    constexpr int H = P / 2;
    constexpr int J = kVarRows;
    if (sc && sc->done) return;   // the whole launch: a converged solve
    __shared__ VarTab t;
    __shared__ double tile[J][kVarThreads + 2 * H];
    const int N = tp->N, R = tp->R;
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * kVarThreads, j0 = blockIdx.y * J;
    const bool jin = j0 - H >= R && j0 + J + H <= N - R;
    const bool iin = i0 - H >= R && i0 + kVarThreads + H <= N - R;
    double acc = 0.0;
    if (jin && iin) {
        acc = var_tile<P, MODE, false>(u, out, aux, sign, tp, t, i0, j0, tile);
    } else {
        {
            const unsigned long long* src = reinterpret_cast<const unsigned long long*>(tp);
            unsigned long long* dst = reinterpret_cast<unsigned long long*>(&t);
            for (int k = tid; k < (int)(sizeof(VarTab) / 8); k += kVarThreads) dst[k] = src[k];
        }
        __syncthreads();
        if (jin) {
            acc = var_tile<P, MODE, true>(u, out, aux, sign, tp, t, i0, j0, tile);
        } else {
            const int i = i0 + tid;   // face rows: per point
            if (i < N) {
                const int ri = i < N - 1 - i ? i : N - 1 - i;
                const double hwi = ri < t.RH ? t.hw[ri] : 1.0;
                for (int r = 0; r < J; ++r) {
                    const int j = j0 + r;
                    if (j >= N) break;
                    const int rj = j < N - 1 - j ? j : N - 1 - j;
                    const double hwj = rj < t.RH ? t.hw[rj] : 1.0;
                    const double kp = var_row<H>(t.k, t.R, t.C, N, j, [&](int c) { return u[gidx(i, c, N)]; });
                    const double mp = var_row<H>(t.m, t.R, t.C, N, i, [&](int c) { return u[gidx(c, j, N)]; });
                    var_emit<MODE>(out, aux, sign, gidx(i, j, N), hwi * kp + hwj * mp, acc);
                }
            }
        }
    }
    if constexpr (MODE != 0) {
        __syncthreads();
        const double s = block_sum<double, kVarThreads>(acc);
        if (threadIdx.x == 0) partial[blockIdx.x + (size_t)blockIdx.y * gridDim.x] = s;
    }
}

}  // namespace sbp
