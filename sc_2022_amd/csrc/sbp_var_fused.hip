// sbp_var_fused.hip -- launchers of the variable-coefficient two-sweep
// V-cycle kernels (sbp_var_fused.h): orders 2, 4, 6, FP64 and FP32 grid
// functions, the 64-wide block shapes.
#include <hip/hip_runtime.h>

#include "sbp_launch.h"
#include "sbp_var_fused.h"

namespace sbp {

namespace {
constexpr int kVarFusedNT = 256;   // threads per block
// Launches of at most kVarFusedWideBlocks blocks run with twice the threads
// (as the constant-coefficient two-sweep kernels do): few blocks per CU, a
// time set by each block's chain of phases.
constexpr int kVarFusedWideNT = 512;
constexpr int kVarFusedWideBlocks = 1024;

// Compiled block shapes (output columns x fine rows).
#define SBP_VAR_FUSED_SHAPES(X) X(64, 4) X(64, 8)

template <int P, int TI, int J, typename TG>
void vsr_launch(const VarFusedArgs& a, const TG* b, TG* x2, TG* rc, FusedCtl c, hipStream_t st) {
    const int nb = c.nbx * c.nby;
    if (nb <= kVarFusedWideBlocks)
        hipLaunchKernelGGL((var_smooth2_restrict<P, TI, J, kVarFusedWideNT, TG>), dim3(nb), dim3(kVarFusedWideNT), 0,
                           st, b, x2, rc, a.omega, a.scale, a.tab, a.kband, a.mband, a.N, c);
    else
        hipLaunchKernelGGL((var_smooth2_restrict<P, TI, J, kVarFusedNT, TG>), dim3(nb), dim3(kVarFusedNT), 0, st, b,
                           x2, rc, a.omega, a.scale, a.tab, a.kband, a.mband, a.N, c);
}

template <int P, int TI, int J, typename TG>
void vps_launch(const VarFusedArgs& a, const TG* x, const TG* ec, const TG* b, TG* out, const TG* dotv, FusedCtl c,
                hipStream_t st) {
    const int nb = c.nbx * c.nby;
    if (nb <= kVarFusedWideBlocks)
        hipLaunchKernelGGL((var_prolong_smooth2<P, TI, J, kVarFusedWideNT, TG>), dim3(nb), dim3(kVarFusedWideNT), 0,
                           st, x, ec, b, out, dotv, a.omega, a.tab, a.kband, a.mband, a.N, c);
    else
        hipLaunchKernelGGL((var_prolong_smooth2<P, TI, J, kVarFusedNT, TG>), dim3(nb), dim3(kVarFusedNT), 0, st, x,
                           ec, b, out, dotv, a.omega, a.tab, a.kband, a.mband, a.N, c);
}
}  // namespace

bool var_fused_shape_ok(int ti, int tj) {
#define SBP_VF_OK(TI_, J_) \
    if (ti == TI_ && tj == J_) return true;
    SBP_VAR_FUSED_SHAPES(SBP_VF_OK)
#undef SBP_VF_OK
    return false;
}

template <typename TG>
int launch_var_smooth2_restrict(int p, int ti, int tj, const VarFusedArgs& a, const TG* b, TG* x2, TG* rc,
                                hipStream_t st) {
    const int Nc = (a.N + 1) / 2;
    FusedCtl c{};
    c.done = a.done;
    c.ds = 0;
    c.nbx = (Nc + ti / 2 - 2) / (ti / 2 - 1);
    c.nby = (Nc + tj / 2 - 1) / (tj / 2);
    c.partial = nullptr;
    c.bt = Batch{1, 0, 0, 0};
#define SBP_VSR_CASE(TI_, J_)                                              \
    if (ti == TI_ && tj == J_) {                                           \
        if (p == 2) vsr_launch<2, TI_, J_, TG>(a, b, x2, rc, c, st);       \
        else if (p == 4) vsr_launch<4, TI_, J_, TG>(a, b, x2, rc, c, st);  \
        else if (p == 6) vsr_launch<6, TI_, J_, TG>(a, b, x2, rc, c, st);  \
        else return 3;                                                     \
        return hipGetLastError() == hipSuccess ? 0 : 4;                    \
    }
    SBP_VAR_FUSED_SHAPES(SBP_VSR_CASE)
#undef SBP_VSR_CASE
    return 3;
}

template <typename TG>
int launch_var_prolong_smooth2(int p, int ti, int tj, const VarFusedArgs& a, const TG* x, const TG* ec, const TG* b,
                               TG* out, const TG* dotv, double* partial, hipStream_t st) {
    FusedCtl c{};
    c.done = a.done;
    c.ds = 0;
    c.nbx = (a.N + ti - 1) / ti;
    c.nby = (a.N + tj - 1) / tj;
    c.partial = partial;
    c.bt = Batch{1, 0, 0, 0};
#define SBP_VPS_CASE(TI_, J_)                                                        \
    if (ti == TI_ && tj == J_) {                                                     \
        if (p == 2) vps_launch<2, TI_, J_, TG>(a, x, ec, b, out, dotv, c, st);       \
        else if (p == 4) vps_launch<4, TI_, J_, TG>(a, x, ec, b, out, dotv, c, st);  \
        else if (p == 6) vps_launch<6, TI_, J_, TG>(a, x, ec, b, out, dotv, c, st);  \
        else return 3;                                                               \
        return hipGetLastError() == hipSuccess ? 0 : 4;                              \
    }
    SBP_VAR_FUSED_SHAPES(SBP_VPS_CASE)
#undef SBP_VPS_CASE
    return 3;
}

template int launch_var_smooth2_restrict<double>(int, int, int, const VarFusedArgs&, const double*, double*, double*,
                                                 hipStream_t);
template int launch_var_smooth2_restrict<float>(int, int, int, const VarFusedArgs&, const float*, float*, float*,
                                                hipStream_t);
template int launch_var_prolong_smooth2<double>(int, int, int, const VarFusedArgs&, const double*, const double*,
                                                const double*, double*, const double*, double*, hipStream_t);
template int launch_var_prolong_smooth2<float>(int, int, int, const VarFusedArgs&, const float*, const float*,
                                               const float*, float*, const float*, double*, hipStream_t);

}  // namespace sbp
