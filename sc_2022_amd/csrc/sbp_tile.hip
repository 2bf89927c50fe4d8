// sbp_tile.hip -- instantiates the stencil pass for one tile shape
// (compile with -DSBP_TI=<threads along i> -DSBP_TJ=<rows per thread>
// -DSBP_TA=<1: line-aligned stores>; with -DSBP_HI=1 also for the orders
// p = 6, 8 and 10).  The residual + restriction kernel is instantiated for
// the unaligned shapes only.
#include <algorithm>

#include "sbp_launch.h"

#ifndef SBP_TI
#error "define SBP_TI, SBP_TJ and SBP_TA"
#endif
#ifndef SBP_TA
#define SBP_TA 0
#endif

namespace sbp {

template <int TI, int J, int TA, typename T, class IN, class OUT>
int launch_tile(int p, const OpTab& tb, const IN& in, const OUT& out, PassCtl c, hipStream_t st) {
    constexpr int ALN = TA ? 128 / (int)sizeof(T) : 0;
    c.nbx = Cols<TI, ALN>::nbx(tb.N);
    c.nby = (tb.N + J - 1) / J;
    {   // interior tiles: every point they load (halo included) lies off the closure rows
        constexpr int TO = Cols<TI, ALN>::TO;
        const int N = tb.N, R = tb.R, H = tb.H;
        const int ylast = N - R - H - J, xlast = N - R - H - TI + ALN;
        c.by0 = (R + H + J - 1) / J;
        c.by1 = ylast >= 0 ? std::min(c.nby, ylast / J + 1) : 0;
        c.bx0 = (R + H + ALN + TO - 1) / TO;
        c.bx1 = xlast >= 0 ? std::min(c.nbx, xlast / TO + 1) : 0;
        if (c.by1 <= c.by0 || c.bx1 <= c.bx0) c.bx0 = c.bx1 = c.by0 = c.by1 = 0;
    }
    const dim3 g(c.nbx * c.nby * c.bt.nprob);
    switch (p) {
    case 2: hipLaunchKernelGGL((stencil_pass<2, TI, J, ALN, T, IN, OUT>), g, dim3(TI), 0, st, in, out, tb, c); break;
    case 4: hipLaunchKernelGGL((stencil_pass<4, TI, J, ALN, T, IN, OUT>), g, dim3(TI), 0, st, in, out, tb, c); break;
#if SBP_HI
    case 6: hipLaunchKernelGGL((stencil_pass<6, TI, J, ALN, T, IN, OUT>), g, dim3(TI), 0, st, in, out, tb, c); break;
    case 8: hipLaunchKernelGGL((stencil_pass<8, TI, J, ALN, T, IN, OUT>), g, dim3(TI), 0, st, in, out, tb, c); break;
    case 10: hipLaunchKernelGGL((stencil_pass<10, TI, J, ALN, T, IN, OUT>), g, dim3(TI), 0, st, in, out, tb, c); break;
#endif
    default: return 3;   // order not compiled for this tile shape
    }
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

template <int TI, int J, typename T>
int launch_resid_restrict(int p, const OpTab& tf, const T* u, const T* b, T* rc, int Nc, T scale, const int* done,
                          int ds, const Batch& bt, hipStream_t st) {
    RestrictCtl c;
    c.done = done;
    c.ds = ds;
    c.Nc = Nc;
    c.nbx = (Nc + TI / 2 - 2) / (TI / 2 - 1);
    c.nby = (Nc + J / 2 - 1) / (J / 2);
    c.bt = bt;
    {   // interior blocks: the fine points they load (halo included) lie off the closure rows
        constexpr int CI = TI / 2 - 1;
        const int N = tf.N, R = tf.R, H = tf.H;
        const int ylast = N - R - H - J, xlast = N - R - H - TI + 1;
        c.by0 = (R + H + 1 + J - 1) / J;
        c.by1 = ylast >= 0 ? std::min(c.nby, ylast / J + 1) : 0;
        c.bx0 = (R + H + 1 + 2 * CI - 1) / (2 * CI);
        c.bx1 = xlast >= 0 ? std::min(c.nbx, xlast / (2 * CI) + 1) : 0;
        if (c.by1 <= c.by0 || c.bx1 <= c.bx0) c.bx0 = c.bx1 = c.by0 = c.by1 = 0;
    }
    const dim3 g(c.nbx * c.nby * bt.nprob);
    switch (p) {
    case 2: hipLaunchKernelGGL((resid_restrict<2, TI, J, T>), g, dim3(TI), 0, st, u, b, rc, scale, tf, c); break;
    case 4: hipLaunchKernelGGL((resid_restrict<4, TI, J, T>), g, dim3(TI), 0, st, u, b, rc, scale, tf, c); break;
#if SBP_HI
    case 6: hipLaunchKernelGGL((resid_restrict<6, TI, J, T>), g, dim3(TI), 0, st, u, b, rc, scale, tf, c); break;
    case 8: hipLaunchKernelGGL((resid_restrict<8, TI, J, T>), g, dim3(TI), 0, st, u, b, rc, scale, tf, c); break;
    case 10: hipLaunchKernelGGL((resid_restrict<10, TI, J, T>), g, dim3(TI), 0, st, u, b, rc, scale, tf, c); break;
#endif
    default: return 3;
    }
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

#if !SBP_TA
template int launch_resid_restrict<SBP_TI, SBP_TJ, double>(int, const OpTab&, const double*, const double*, double*,
                                                           int, double, const int*, int, const Batch&, hipStream_t);
template int launch_resid_restrict<SBP_TI, SBP_TJ, float>(int, const OpTab&, const float*, const float*, float*, int,
                                                          float, const int*, int, const Batch&, hipStream_t);
#endif

#define SBP_INST(T, IN, OUT) \
    template int launch_tile<SBP_TI, SBP_TJ, SBP_TA, T, IN, OUT>(int, const OpTab&, const IN&, const OUT&, PassCtl, \
                                                                 hipStream_t);
#if SBP_TA
SBP_ALIGNED_COMBOS(SBP_INST)
#else
SBP_STENCIL_COMBOS(SBP_INST)
#endif

}  // namespace sbp
