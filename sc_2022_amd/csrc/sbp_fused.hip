// sbp_fused.hip -- launchers of the temporally blocked V-cycle passes
// (sbp_fused.h), one translation unit for every order, precision and width.
#include <hip/hip_runtime.h>

#include "sbp_fused.h"
#include "sbp_launch.h"

namespace sbp {

namespace {
constexpr int kFusedJ = 8;      // fine rows of a block's output region
constexpr int kFusedNT = 256;   // threads per block

template <int P, int TI, typename T>
void sr_launch(const OpTab& tf, const T* b, T* x2, T* rc, T omega, int use_diag, T scale, FusedCtl c,
               hipStream_t st) {
    hipLaunchKernelGGL((smooth2_restrict<P, TI, kFusedJ, kFusedNT, T>), dim3(c.nbx * c.nby * c.bt.nprob),
                       dim3(kFusedNT), 0, st, b, x2, rc, omega, use_diag, scale, tf, c);
}

template <int P, int TI, typename T>
void ps_launch(const OpTab& tf, const T* x, const T* ec, const T* b, T* out, const T* dotv, T omega, int use_diag,
               FusedCtl c, hipStream_t st) {
    hipLaunchKernelGGL((prolong_smooth2<P, TI, kFusedJ, kFusedNT, T>), dim3(c.nbx * c.nby * c.bt.nprob),
                       dim3(kFusedNT), 0, st, x, ec, b, out, dotv, omega, use_diag, tf, c);
}
}  // namespace

int fused_post_blocks(int N, int ti) { return ((N + ti - 1) / ti) * ((N + kFusedJ - 1) / kFusedJ); }

bool fused_width_ok(int ti) { return ti == 128 || ti == 256; }

template <typename T>
int launch_smooth2_restrict(int p, int ti, const OpTab& tf, const T* b, T* x2, T* rc, T omega, int use_diag,
                            T scale, const int* done, int ds, const Batch& bt, hipStream_t st) {
    const int Nc = (tf.N + 1) / 2;
    FusedCtl c{};
    c.done = done;
    c.ds = ds;
    c.nbx = (Nc + ti / 2 - 2) / (ti / 2 - 1);
    c.nby = (Nc + kFusedJ / 2 - 1) / (kFusedJ / 2);
    c.partial = nullptr;
    c.bt = bt;
    if (p == 2 && ti == 128) sr_launch<2, 128, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);
    else if (p == 2 && ti == 256) sr_launch<2, 256, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);
    else if (p == 4 && ti == 128) sr_launch<4, 128, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);
    else if (p == 4 && ti == 256) sr_launch<4, 256, T>(tf, b, x2, rc, omega, use_diag, scale, c, st);
    else return 3;
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

template <typename T>
int launch_prolong_smooth2(int p, int ti, const OpTab& tf, const T* x, const T* ec, const T* b, T* out,
                           const T* dotv, T omega, int use_diag, const int* done, int ds, double* partial,
                           const Batch& bt, hipStream_t st) {
    FusedCtl c{};
    c.done = done;
    c.ds = ds;
    c.nbx = (tf.N + ti - 1) / ti;
    c.nby = (tf.N + kFusedJ - 1) / kFusedJ;
    c.partial = partial;
    c.bt = bt;
    if (p == 2 && ti == 128) ps_launch<2, 128, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st);
    else if (p == 2 && ti == 256) ps_launch<2, 256, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st);
    else if (p == 4 && ti == 128) ps_launch<4, 128, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st);
    else if (p == 4 && ti == 256) ps_launch<4, 256, T>(tf, x, ec, b, out, dotv, omega, use_diag, c, st);
    else return 3;
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

template int launch_smooth2_restrict<double>(int, int, const OpTab&, const double*, double*, double*, double, int,
                                             double, const int*, int, const Batch&, hipStream_t);
template int launch_smooth2_restrict<float>(int, int, const OpTab&, const float*, float*, float*, float, int, float,
                                            const int*, int, const Batch&, hipStream_t);
template int launch_prolong_smooth2<double>(int, int, const OpTab&, const double*, const double*, const double*,
                                            double*, const double*, double, int, const int*, int, double*,
                                            const Batch&, hipStream_t);
template int launch_prolong_smooth2<float>(int, int, const OpTab&, const float*, const float*, const float*, float*,
                                           const float*, float, int, const int*, int, double*, const Batch&,
                                           hipStream_t);

}  // namespace sbp
