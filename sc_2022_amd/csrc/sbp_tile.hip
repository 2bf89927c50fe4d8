// sbp_tile.hip -- instantiates the stencil pass for one tile shape
// (compile with -DSBP_TI=<threads along i> -DSBP_TJ=<rows per thread>).
#include "sbp_launch.h"

#ifndef SBP_TI
#error "define SBP_TI and SBP_TJ"
#endif

namespace sbp {

template <int TI, int J, typename T, class IN, class OUT>
int launch_tile(int p, const OpTab& tb, const IN& in, const OUT& out, const int* done, double* partial,
                hipStream_t st) {
    PassCtl c;
    c.done = done;
    c.partial = partial;
    c.nbx = (tb.N + TI - 1) / TI;
    c.nby = (tb.N + J - 1) / J;
    if (p == 2)
        hipLaunchKernelGGL((stencil_pass<2, TI, J, T, IN, OUT>), dim3(c.nbx * c.nby), dim3(TI), 0, st, in, out, tb, c);
    else
        hipLaunchKernelGGL((stencil_pass<4, TI, J, T, IN, OUT>), dim3(c.nbx * c.nby), dim3(TI), 0, st, in, out, tb, c);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

#define SBP_INST(T, IN, OUT)                                                                                \
    template int launch_tile<SBP_TI, SBP_TJ, T, IN, OUT>(int, const OpTab&, const IN&, const OUT&, const int*, \
                                                         double*, hipStream_t);
SBP_STENCIL_COMBOS(SBP_INST)

}  // namespace sbp
