// sbp_fused.hip -- launchers of the temporally blocked V-cycle passes
// (sbp_fused.h), one translation unit for every order, precision and width.
#include <hip/hip_runtime.h>

#include "sbp_fused.h"
#include "sbp_launch.h"

namespace sbp {

namespace {
constexpr int kFusedNT = 256;   // threads per block
// Launches of at most kFusedWideBlocks blocks (the levels of side <= 513 of
// one problem) run the 64-wide shapes with twice the threads: there every
// block is one of a few per CU and the kernel's time is its blocks' chain of
// phases, which more lanes per phase shorten (pre-smoothing + restriction,
// 129^2: 7.8 against 9.7 us per launch; 513^2 on 64 x 8: 10.2 against 11.5);
// with several blocks per CU the extra lanes gain nothing (1025^2 on 64 x 8,
// 2 193 blocks: 16.7 against 16.2).
constexpr int kFusedWideNT = 512;
constexpr int kFusedWideBlocks = 1024;

// LDS of the two kernels' tiles (sbp_fused.h: smooth2_restrict holds b and
// x1 on region A and x2 on region B; prolong_smooth2 the operand on A, the
// coarse patch, and b and the first sweep on B), plus the staged table.
template <int P, int TI, int J, typename T> constexpr int sr_lds_bytes() {
    constexpr int HA = FusedHalo<P>::HA, HB = FusedHalo<P>::HB;
    return (int)sizeof(T) * (2 * (TI + 2 * HA) * (J + 1 + 2 * HA) + (TI + 2 * HB) * (J + 1 + 2 * HB)) +
           (int)sizeof(OpTab);
}
template <int P, int TI, int J, typename T> constexpr int ps_lds_bytes() {
    constexpr int HA = FusedHalo<P>::HA, HB = FusedHalo<P>::HB;
    constexpr int AW = TI + 2 * HA, AH = J + 2 * HA;
    return (int)sizeof(T) * (AW * AH + (AW / 2 + 2) * (AH / 2 + 2) + 2 * (TI + 2 * HB) * (J + 2 * HB)) +
           (int)sizeof(OpTab);
}
// A shape is compiled for order P when both kernels fit the 160 KiB of LDS a
// gfx950 workgroup can hold, in both precisions (the V-cycle of a handle runs
// either).  Orders 2 and 4 fit every shape; p = 6 (halos of 5 and 10 points)
// fits the 64- and 128-wide blocks, not the 256-wide FP64 ones (169 KB).
constexpr int kLdsBytes = 160 * 1024;
template <int P, int TI, int J> constexpr bool fused_compiled() {
    return P <= 6 && sr_lds_bytes<P, TI, J, double>() <= kLdsBytes && ps_lds_bytes<P, TI, J, double>() <= kLdsBytes;
}

template <int P, int TI, int J, typename T>
void sr_launch(const OpTab& tf, const T* b, T* x2, T* rc, T omega, int use_diag, T scale, FusedCtl c,
               hipStream_t st) {
    const int nb = c.nbx * c.nby * c.bt.nprob;
    if constexpr (TI == 64) {
        if (nb <= kFusedWideBlocks) {
            hipLaunchKernelGGL((smooth2_restrict<P, TI, J, kFusedWideNT, T>), dim3(nb), dim3(kFusedWideNT), 0, st, b,
                               x2, rc, omega, use_diag, scale, tf, c);
            return;
        }
    }
    hipLaunchKernelGGL((smooth2_restrict<P, TI, J, kFusedNT, T>), dim3(nb), dim3(kFusedNT), 0, st, b, x2, rc, omega,
                       use_diag, scale, tf, c);
}

template <int P, int TI, int J, typename T>
void ps_launch(const OpTab& tf, const T* x, const T* ec, const T* b, T* out, const T* dotv, T omega, int use_diag,
               FusedCtl c, hipStream_t st) {
    const int nb = c.nbx * c.nby * c.bt.nprob;
    if constexpr (TI == 64) {
        if (nb <= kFusedWideBlocks) {
            hipLaunchKernelGGL((prolong_smooth2<P, TI, J, kFusedWideNT, T>), dim3(nb), dim3(kFusedWideNT), 0, st, x,
                               ec, b, out, dotv, omega, use_diag, tf, c);
            return;
        }
    }
    hipLaunchKernelGGL((prolong_smooth2<P, TI, J, kFusedNT, T>), dim3(nb), dim3(kFusedNT), 0, st, x, ec, b, out,
                       dotv, omega, use_diag, tf, c);
}
}  // namespace

// Compiled block shapes (output columns x fine rows).
#define SBP_FUSED_SHAPES(X) X(64, 4) X(64, 8) X(128, 8) X(256, 8)

int fused_post_blocks(int N, int ti, int tj) { return ((N + ti - 1) / ti) * ((N + tj - 1) / tj); }

bool fused_shape_ok(int ti, int tj, int p) {
#define SBP_FUSED_OK(TI_, J_) \
    if (ti == TI_ && tj == J_) return p <= 4 || (p == 6 && fused_compiled<6, TI_, J_>());
    SBP_FUSED_SHAPES(SBP_FUSED_OK)
#undef SBP_FUSED_OK
    return false;
}

template <typename T>
int launch_smooth2_restrict(int p, int ti, int tj, const OpTab& tf, const T* b, T* x2, T* rc, T omega,
                            int use_diag, T scale, const int* done, int ds, const Batch& bt, hipStream_t st) {
    const int Nc = (tf.N + 1) / 2;
    FusedCtl c{};
    c.done = done;
    c.ds = ds;
    c.nbx = (Nc + ti / 2 - 2) / (ti / 2 - 1);
    c.nby = (Nc + tj / 2 - 1) / (tj / 2);
    c.partial = nullptr;
    c.bt = bt;
#define SBP_SR_CASE(TI_, J_)                                                                  \
    if (ti == TI_ && tj == J_) {                                                              \
        if (p == 2) sr_launch<2, TI_, J_, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);  \
        else if (p == 4) sr_launch<4, TI_, J_, T>(tf, b, x2, rc, omega, use_diag, scale, c, st); \
        else if (p == 6) {                                                                    \
            if constexpr (fused_compiled<6, TI_, J_>())                                       \
                sr_launch<6, TI_, J_, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);       \
            else return 3;                                                                    \
        } else return 3;                                                                      \
        return hipGetLastError() == hipSuccess ? 0 : 4;                                       \
    }
    SBP_FUSED_SHAPES(SBP_SR_CASE)
#undef SBP_SR_CASE
    return 3;
}

template <typename T>
int launch_prolong_smooth2(int p, int ti, int tj, const OpTab& tf, const T* x, const T* ec, const T* b, T* out,
                           const T* dotv, T omega, int use_diag, const int* done, int ds, double* partial,
                           const Batch& bt, hipStream_t st) {
    FusedCtl c{};
    c.done = done;
    c.ds = ds;
    c.nbx = (tf.N + ti - 1) / ti;
    c.nby = (tf.N + tj - 1) / tj;
    c.partial = partial;
    c.bt = bt;
#define SBP_PS_CASE(TI_, J_)                                                                   \
    if (ti == TI_ && tj == J_) {                                                               \
        if (p == 2) ps_launch<2, TI_, J_, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st); \
        else if (p == 4) ps_launch<4, TI_, J_, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st); \
        else if (p == 6) {                                                                     \
            if constexpr (fused_compiled<6, TI_, J_>())                                        \
                ps_launch<6, TI_, J_, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st);     \
            else return 3;                                                                     \
        } else return 3;                                                                       \
        return hipGetLastError() == hipSuccess ? 0 : 4;                                        \
    }
    SBP_FUSED_SHAPES(SBP_PS_CASE)
#undef SBP_PS_CASE
    return 3;
}

template int launch_smooth2_restrict<double>(int, int, int, const OpTab&, const double*, double*, double*, double,
                                             int, double, const int*, int, const Batch&, hipStream_t);
template int launch_smooth2_restrict<float>(int, int, int, const OpTab&, const float*, float*, float*, float, int,
                                            float, const int*, int, const Batch&, hipStream_t);
template int launch_prolong_smooth2<double>(int, int, int, const OpTab&, const double*, const double*,
                                            const double*, double*, const double*, double, int, const int*, int,
                                            double*, const Batch&, hipStream_t);
template int launch_prolong_smooth2<float>(int, int, int, const OpTab&, const float*, const float*, const float*,
                                           float*, const float*, float, int, const int*, int, double*, const Batch&,
                                           hipStream_t);

}  // namespace sbp
