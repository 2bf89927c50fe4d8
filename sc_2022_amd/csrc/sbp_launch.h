// sbp_launch.h -- launch entry points of the stencil pass.  Each compiled
// tile shape (TI x J) lives in its own translation unit (sbp_tile.hip built
// once per shape) so that the shapes compile in parallel; the API dispatches
// on the shape chosen for a level.
#pragma once
#include "sbp_stencil.h"

namespace sbp {

using ResidDD = OutResid<double, double>;
using ResidDF = OutResid<double, float>;
using ResidFF = OutResid<float, float>;

// Every (element type, input policy, output policy) the solvers use.
#define SBP_STENCIL_COMBOS(X)                     \
    X(double, InPlain<double>, OutStore<double>)    \
    X(double, InPlain<double>, OutStoreDot<double>) \
    X(double, InAxpy<double>, OutWriteDot<double>)  \
    X(double, InPlain<double>, ResidDD)             \
    X(double, InPlain<double>, ResidDF)             \
    X(double, InSmooth0<double>, OutSmooth<double>) \
    X(double, InPlain<double>, OutSmooth<double>)   \
    X(double, InPlain<double>, OutSmoothDot<double>) \
    X(double, InProlong<double>, OutSmooth<double>) \
    X(double, InProlong<double>, OutSmoothDot<double>) \
    X(double, InAxpy<double>, OutPUpd<double>)      \
    X(double, InSmooth0<double>, OutCopy<double>)   \
    X(float, InPlain<float>, OutStore<float>)       \
    X(float, InPlain<float>, OutStoreDot<float>)    \
    X(float, InAxpy<float>, OutWriteDot<float>)     \
    X(float, InPlain<float>, ResidFF)               \
    X(float, InSmooth0<float>, OutSmooth<float>)    \
    X(float, InPlain<float>, OutSmooth<float>)      \
    X(float, InPlain<float>, OutSmoothDot<float>)   \
    X(float, InProlong<float>, OutSmooth<float>)    \
    X(float, InProlong<float>, OutSmoothDot<float>) \
    X(float, InAxpy<float>, OutPUpd<float>)         \
    X(float, InSmooth0<float>, OutCopy<float>)

// The combinations whose output policy takes line-aligned stores (kAlign):
// the ones the aligned tile shapes are compiled for.
#define SBP_ALIGNED_COMBOS(X)                     \
    X(double, InPlain<double>, OutStore<double>)    \
    X(double, InPlain<double>, OutStoreDot<double>) \
    X(float, InPlain<float>, OutStore<float>)       \
    X(float, InPlain<float>, OutStoreDot<float>)    \
    X(float, InAxpy<float>, OutPUpd<float>)       \
    X(double, InAxpy<double>, OutPUpd<double>)

// Compiled tile shapes (TI threads along i, J rows per thread, A = 1:
// line-aligned stores, the Cols mapping with ALN = 128 bytes / element size).
#define SBP_TILE_SHAPES(X) X(64, 4, 0) X(64, 8, 0) X(128, 8, 0) X(256, 4, 0) X(256, 8, 0) X(256, 16, 0) X(256, 8, 1) X(256, 16, 1) X(512, 8, 1)
// The subset also compiled for p = 6, 8 and 10 (every level's default shape).
#define SBP_HI_TILE_SHAPES(X) X(64, 4, 0) X(128, 8, 0) X(256, 8, 0)

// Elements per 128-byte line for an aligned shape of element size es (0: unaligned).
inline int aln_of(int ta, int es) { return ta ? 128 / es : 0; }

// Tile columns of a pass over an N-wide grid (Cols::nbx).
inline int pass_cols(int N, int ti, int aln) { return aln ? (N + aln - 1 + (ti - aln) - 1) / (ti - aln) : (N + ti - 1) / ti; }

// Blocks (and block partials) of a pass over one N x N problem.
inline int nparts_of(int N, int ti, int tj, int aln = 0) { return pass_cols(N, ti, aln) * ((N + tj - 1) / tj); }

inline int rr_blocks(int Nc, int ti, int tj) { return ((Nc + ti / 2 - 2) / (ti / 2 - 1)) * ((Nc + tj / 2 - 1) / (tj / 2)); }

// rc = scale * FW(b - A u): residual and restriction in one pass (tf = fine
// level table), for bt.nprob stacked problems (done: per-problem flags, stride ds ints).
template <int TI, int J, typename T>
int launch_resid_restrict(int p, const OpTab& tf, const T* u, const T* b, T* rc, int Nc, T scale, const int* done,
                          int ds, const Batch& bt, hipStream_t st);

// Launches stencil_pass<p, TI, J, ALN, T, IN, OUT> (ALN = aln_of(TA, sizeof(T)))
// over the whole N x N grid of tb for every stacked problem (c.done/ds/partial/bt
// given; the tile grid is set here).
template <int TI, int J, int TA, typename T, class IN, class OUT>
int launch_tile(int p, const OpTab& tb, const IN& in, const OUT& out, PassCtl c, hipStream_t st);

// Temporally blocked V-cycle passes (two sweeps per kernel); block shape ti
// output columns x tj fine rows, one of the compiled shapes (fused_shape_ok).
int fused_post_blocks(int N, int ti, int tj);
bool fused_shape_ok(int ti, int tj, int p);
// x2 = S(S(0)) on the fine level and rc = scale * FW(b - A x2).
template <typename T>
int launch_smooth2_restrict(int p, int ti, int tj, const OpTab& tf, const T* b, T* x2, T* rc, T omega,
                            int use_diag, T scale, const int* done, int ds, const Batch& bt, hipStream_t st);
// out = S(S(x + P ec)); with dotv, per-block partials of dotv' out (fused_post_blocks per problem).
template <typename T>
int launch_prolong_smooth2(int p, int ti, int tj, const OpTab& tf, const T* x, const T* ec, const T* b, T* out,
                           const T* dotv, T omega, int use_diag, const int* done, int ds, double* partial,
                           const Batch& bt, hipStream_t st);

// The same two kernels for the variable-coefficient operator (one problem;
// FP64 arithmetic, grid functions of element type TG; block shapes 64 x 4 and
// 64 x 8, var_fused_shape_ok).  tab: the level's device table; kband /
// mband: its (2H+1) x N coefficient bands; done: nullable converged flag.
struct VarTab;
struct VarFusedArgs {
    int N;
    const VarTab* tab;
    const double* kband;
    const double* mband;
    double omega, scale;
    const int* done;
};
bool var_fused_shape_ok(int ti, int tj);
template <typename TG>
int launch_var_smooth2_restrict(int p, int ti, int tj, const VarFusedArgs& a, const TG* b, TG* x2, TG* rc,
                                hipStream_t st);
template <typename TG>
int launch_var_prolong_smooth2(int p, int ti, int tj, const VarFusedArgs& a, const TG* x, const TG* ec, const TG* b,
                               TG* out, const TG* dotv, double* partial, hipStream_t st);

}  // namespace sbp
