// sbp_var_tab.h -- the device table of the variable-coefficient operator
// (its closure blocks, H weights and coefficient bands; built in sbp_var.h).
#pragma once
#include "sbp_stencil.h"

namespace sbp {

// Closure rows / columns of the variable-coefficient factors:
// p = 2: 3 / 4, p = 4: 6 / 8, p = 6: 9 / 12.
constexpr int kVarMaxR = 9;
constexpr int kVarMaxC = 12;
inline int var_R(int p) { return p == 2 ? 3 : (p == 4 ? 6 : 9); }
inline int var_C(int p) { return p == 2 ? 4 : (p == 4 ? 8 : 12); }
template <int P> struct VarShape;   // the same, at compile time
template <> struct VarShape<2> { static constexpr int R = 3, C = 4; };
template <> struct VarShape<4> { static constexpr int R = 6, C = 8; };
template <> struct VarShape<6> { static constexpr int R = 9, C = 12; };

// One axis' factor on the device: closure rows (first / last R) and the band
// of the rows between.
struct VarAxis {
    double lo[kVarMaxR][kVarMaxC];   // rows r < R, columns c < C
    double hi[kVarMaxR][kVarMaxC];   // rows N-1-r, columns N-1-c
    const double* band;              // (2H+1) x N, offset s = -H..H at band[(s+H) N + r]
};

struct VarTab {
    int N, R, C, H, RH;
    double hw[kMaxR];
    VarAxis k, m;
};

}  // namespace sbp
