// sbp_var_fused.h -- the temporally blocked V-cycle passes (sbp_fused.h) for
// the variable-coefficient operator (sbp_var.h).
//
// The two bodies of sbp_fused.h are written against an operator class; this
// is the variable-coefficient one.  It evaluates in FP64 whatever the element
// type of the grid functions (as the variable-coefficient passes do), with
// the same order of operations as var_pass (sbp_var.h): per axis the band sum over
// offsets -H..H (interior rows) or the closure row of the face block, then
// hw_i * (k part) + hw_j * (m part); the Jacobi diagonal is
// hw_i k_jj + hw_j m_ii.  The band coefficients of the block's region A
// (k for its rows, m for its columns) are staged in LDS together with the
// table (closure blocks and H weights; face blocks only), issued with the
// block's data loads.
#pragma once
#include "sbp_fused.h"
#include "sbp_var_tab.h"

namespace sbp {

template <int P, int NT, int AW, int AH, bool INNER>
struct FusedOpV {
    using T = double;
    static constexpr int H = P / 2, R = VarShape<P>::R, C = VarShape<P>::C, NQ = 2 * H + 1;
    static constexpr bool kStage = true;
    static constexpr bool kInner = INNER;
    static constexpr int NW = (int)(sizeof(VarTab) / 8);       // table words
    static constexpr int KT = INNER ? 1 : (NW + NT - 1) / NT;
    static constexpr int NBD = NQ * (AH + AW);                   // staged band values
    static constexpr int KB = (NBD + NT - 1) / NT;
    static_assert(sizeof(VarTab) % 8 == 0, "table is copied in 8-byte words");
    const VarTab* tp;       // device table
    VarTab* lt;             // its LDS copy (face blocks)
    double* bd;             // LDS: k band [NQ][AH] for rows aj0.., then m band [NQ][AW] for columns ai0..
    const double* kband;    // (2H+1) x N, offset-major
    const double* mband;
    int N;
    int ai0 = 0, aj0 = 0;
    unsigned long long tv[KT];
    double bv[KB];
    __device__ FusedOpV(const VarTab* t, VarTab* l, double* b, const double* kb, const double* mb, int n)
        : tp(t), lt(l), bd(b), kband(kb), mband(mb), N(n) {}

    __device__ __forceinline__ void fetch(int a_i0, int a_j0) {
        ai0 = a_i0;
        aj0 = a_j0;
        const int tid = threadIdx.x;
        if constexpr (!INNER) {
            const unsigned long long* s = reinterpret_cast<const unsigned long long*>(tp);
#pragma unroll
            for (int m = 0; m < KT; ++m) {
                const int k = tid + m * NT;
                if (k < NW) tv[m] = s[k];
            }
        }
#pragma unroll
        for (int m = 0; m < KB; ++m) {
            const int k = tid + m * NT;
            double v = 0.0;
            if (k < NQ * AH) {
                const int q = k / AH, j = aj0 + (k - q * AH);
                if (j >= R && j < N - R) v = kband[(size_t)q * N + j];
            } else if (k < NBD) {
                const int kk = k - NQ * AH;
                const int q = kk / AW, i = ai0 + (kk - q * AW);
                if (i >= R && i < N - R) v = mband[(size_t)q * N + i];
            }
            bv[m] = v;
        }
    }
    __device__ __forceinline__ void store() {
        const int tid = threadIdx.x;
        if constexpr (!INNER) {
            unsigned long long* d = reinterpret_cast<unsigned long long*>(lt);
#pragma unroll
            for (int m = 0; m < KT; ++m) {
                const int k = tid + m * NT;
                if (k < NW) d[k] = tv[m];
            }
        }
#pragma unroll
        for (int m = 0; m < KB; ++m) {
            const int k = tid + m * NT;
            if (k < NBD) bd[k] = bv[m];
        }
    }

    __device__ __forceinline__ double hw(int r) const {
        const int rr = r < N - 1 - r ? r : N - 1 - r;
        return rr < lt->RH ? lt->hw[rr] : 1.0;
    }
    __device__ __forceinline__ double kdiag(int j) const {
        if (!INNER) {
            if (j < R) return lt->k.lo[j][j];
            if (j >= N - R) return lt->k.hi[N - 1 - j][N - 1 - j];
        }
        return bd[H * AH + (j - aj0)];
    }
    __device__ __forceinline__ double mdiag(int i) const {
        if (!INNER) {
            if (i < R) return lt->m.lo[i][i];
            if (i >= N - R) return lt->m.hi[N - 1 - i][N - 1 - i];
        }
        return bd[NQ * AH + H * AW + (i - ai0)];
    }
    // w (b - A u) / diag(A) at (i, j) (the variable-coefficient smoother is
    // always damped Jacobi).
    // (INT: the caller knows (i, j) is off the closure rows of both axes.)
    template <bool INT = false>
    __device__ __forceinline__ double corr(int i, int j, double r, double omega, int) const {
        const double d = (INNER || INT) ? bd[H * AH + (j - aj0)] + bd[NQ * AH + H * AW + (i - ai0)]
                                        : hw(i) * kdiag(j) + hw(j) * mdiag(i);
        return omega * r / d;
    }
    // (A u)(i, j), u in an LDS array holding columns [i0, i0 + W), rows [j0, ...).
    template <int W, bool INT = false>
    __device__ __forceinline__ double au(const double* a, int i0, int j0, int i, int j) const {
        const int k = (j - j0) * W + (i - i0);
        double kp = 0.0, mp = 0.0;
        if (INNER || INT || (j >= R && j < N - R)) {
            const double* c = bd + (j - aj0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) kp += c[q * AH] * a[k + (q - H) * W];
        } else {
            const bool lo = j < R;
            const int rr = lo ? j : N - 1 - j;
#pragma unroll
            for (int c = 0; c < C; ++c)
                kp += (lo ? lt->k.lo[rr][c] : lt->k.hi[rr][c]) * a[((lo ? c : N - 1 - c) - j0) * W + (i - i0)];
        }
        if (INNER || INT || (i >= R && i < N - R)) {
            const double* c = bd + NQ * AH + (i - ai0);
#pragma unroll
            for (int q = 0; q < NQ; ++q) mp += c[q * AW] * a[k + q - H];
        } else {
            const bool lo = i < R;
            const int rr = lo ? i : N - 1 - i;
#pragma unroll
            for (int c = 0; c < C; ++c)
                mp += (lo ? lt->m.lo[rr][c] : lt->m.hi[rr][c]) * a[(j - j0) * W + ((lo ? c : N - 1 - c) - i0)];
        }
        if (INNER || INT) return kp + mp;
        return hw(i) * kp + hw(j) * mp;
    }
};

// x2 = S(S(0)) and rc = scale * FW(b - A x2) (smooth2_restrict for the
// variable-coefficient operator; one problem per launch).
template <int P, int TI, int J, int NT, typename TG>
__global__ __launch_bounds__(NT) void var_smooth2_restrict(const TG* __restrict__ b, TG* __restrict__ x2g,
                                                          TG* __restrict__ rc, double omega, double scale,
                                                          const VarTab* __restrict__ tp, const double* kband,
                                                          const double* mband, int N, FusedCtl ctl) {
    constexpr int HA = FusedHaloR<VarShape<P>::R>::HA, HB = FusedHaloR<VarShape<P>::R>::HB;
    constexpr int JR = J + 1;
    constexpr int AW = TI + 2 * HA, AH = JR + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = JR + 2 * HB;
    constexpr int R = VarShape<P>::R;
    const int nb = ctl.nbx * ctl.nby;
    const int blk = xcd_remap((int)blockIdx.x, nb);
    const int bx = blk % ctl.nbx, by = blk / ctl.nbx;
    const int Nc = (N + 1) / 2;
    __shared__ double Bt[AW * AH];
    __shared__ double U[AW * AH];
    __shared__ double V[BW * BH];
    __shared__ double bd[(P + 1) * (AW + AH)];
    constexpr int CI = TI / 2 - 1, CJ = J / 2;
    int Ic0 = bx * CI < Nc - CI ? bx * CI : Nc - CI;
    int Jc0 = by * CJ < Nc - CJ ? by * CJ : Nc - CJ;
    if (Ic0 < 0) Ic0 = 0;
    if (Jc0 < 0) Jc0 = 0;
    const int ai0 = 2 * Ic0 - 1 - HA, aj0 = 2 * Jc0 - 1 - HA;
    const bool inner = ai0 >= R && ai0 + AW <= N - R && aj0 >= R && aj0 + AH <= N - R;
    if (inner) {
        FusedOpV<P, NT, AW, AH, true> op(tp, nullptr, bd, kband, mband, N);
        sr_body<TI, J, NT, double, TG>(op, N, b, x2g, rc, omega, 1, scale, ctl.done, bx, by, Nc, Bt, U, V);
    } else {
        __shared__ VarTab lt;
        FusedOpV<P, NT, AW, AH, false> op(tp, &lt, bd, kband, mband, N);
        sr_body<TI, J, NT, double, TG>(op, N, b, x2g, rc, omega, 1, scale, ctl.done, bx, by, Nc, Bt, U, V);
    }
}

// out = S(S(x + P e)), optionally with block partials of dotv' out
// (prolong_smooth2 for the variable-coefficient operator).
template <int P, int TI, int J, int NT, typename TG>
__global__ __launch_bounds__(NT) void var_prolong_smooth2(const TG* __restrict__ x, const TG* __restrict__ ec,
                                                         const TG* __restrict__ b, TG* __restrict__ out,
                                                         const TG* __restrict__ dotv, double omega,
                                                         const VarTab* __restrict__ tp, const double* kband,
                                                         const double* mband, int N, FusedCtl ctl) {
    constexpr int HA = FusedHaloR<VarShape<P>::R>::HA, HB = FusedHaloR<VarShape<P>::R>::HB;
    constexpr int AW = TI + 2 * HA, AH = J + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = J + 2 * HB;
    constexpr int EW = AW / 2 + 2, EH = AH / 2 + 2;
    constexpr int R = VarShape<P>::R;
    const int nb = ctl.nbx * ctl.nby;
    const int blk = xcd_remap((int)blockIdx.x, nb);
    const int bx = blk % ctl.nbx, by = blk / ctl.nbx;
    __shared__ double U[AW * AH];
    __shared__ double E[EW * EH];
    __shared__ double Bt[BW * BH];
    __shared__ double V[BW * BH];
    __shared__ double bd[(P + 1) * (AW + AH)];
    int i0 = bx * TI < N - TI ? bx * TI : N - TI;
    int j0 = by * J < N - J ? by * J : N - J;
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    const int ai0 = i0 - HA, aj0 = j0 - HA;
    const bool inner = ai0 >= R && ai0 + AW <= N - R && aj0 >= R && aj0 + AH <= N - R;
    double acc = 0.0;
    bool ran;
    if (inner) {
        FusedOpV<P, NT, AW, AH, true> op(tp, nullptr, bd, kband, mband, N);
        ran = ps_body<TI, J, NT, double, TG>(op, N, x, ec, b, out, dotv, omega, 1, ctl.done, bx, by, U, E, Bt, V, acc);
    } else {
        __shared__ VarTab lt;
        FusedOpV<P, NT, AW, AH, false> op(tp, &lt, bd, kband, mband, N);
        ran = ps_body<TI, J, NT, double, TG>(op, N, x, ec, b, out, dotv, omega, 1, ctl.done, bx, by, U, E, Bt, V, acc);
    }
    if (!ran) return;   // converged: every thread of the block left before its first barrier
    if (dotv) {
        __syncthreads();
        const double s = block_sum<double, NT>(acc);
        if (threadIdx.x == 0) ctl.partial[blk] = s;
    }
}

}  // namespace sbp
