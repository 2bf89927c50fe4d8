// sbp_fused.h -- temporally blocked V-cycle passes.
//
// A V-cycle level with two pre- and two post-smoothing sweeps does, on the
// fine grid, (pre) x2 = S(S(0)) then the residual b - A x2 and its
// restriction, and (post) out = S(S(x2 + P e)).  Run as separate stencil
// passes this streams x2, the intermediate sweep and b through HBM several
// times, and on the small levels every pass is a kernel boundary (~5 us
// whatever the size).  Here each block loads its tile of the inputs once,
// with a halo two stencil radii wide, runs both sweeps (and the residual)
// in LDS, recomputing the halo points the neighbouring blocks also compute,
// and writes only the results:
//
//   smooth2_restrict  reads b; writes x2 and rc = scale * FW(b - A x2)
//   prolong_smooth2   reads x2, e_c, b (and dotv); writes out (and dotv'out)
//
// Geometry.  A block's output region C is TI columns x JC rows; the first
// sweep is evaluated on region B = C widened by HB and its operand on region
// A = B widened by HA - HB (FusedHalo: R-1 each, R the closure rows, which
// covers both the stencil radius and the closure rows that reach C columns
// from a face).  Blocks are placed on a regular grid except that the last
// block in each direction is shifted back so its region C ends at the grid's
// far face; each block writes only the points (and coarse points) it owns
// under the unshifted tiling, so shifted blocks overlap their neighbours'
// work but never their writes.
#pragma once
#include "sbp_stencil.h"

namespace sbp {

// Operand read from an LDS 2-D array holding columns [i0, i0 + W) and rows
// [j0, ...) of the grid (row stride W).
template <typename T, int W> struct InGrid {
    const T* a;
    int i0, j0;
    using Raw = T;
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab&, int i, int j) const { return a[(j - j0) * W + (i - i0)]; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab&, int, int, const Raw& r) const { return r; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab&, int i, int j) const { return a[(j - j0) * W + (i - i0)]; }
};

__device__ __forceinline__ bool is_interior(const OpTab& t, int i, int j) {
    return i >= t.R && i < t.N - t.R && j >= t.R && j < t.N - t.R;
}

// (A u) at an interior point of an LDS 2-D array of row stride W, k the
// point's index: the (2p+1)-point cross, same order of operations as the
// tiled passes' interior path (j-direction part first, h-weights 1).
template <int P, typename T, int W>
__device__ __forceinline__ T au_interior(const T* a, int k) {
    constexpr int H = Stencil<P>::H;
    const T uc = a[k];
    T kp = Stencil<P>::template c<T>(1) * ((a[k - W] - uc) + (a[k + W] - uc));
#pragma unroll
    for (int s = 2; s <= H; ++s) kp += Stencil<P>::template c<T>(s) * ((a[k - s * W] - uc) + (a[k + s * W] - uc));
    T mp = Stencil<P>::template c<T>(1) * ((a[k - 1] - uc) + (a[k + 1] - uc));
#pragma unroll
    for (int s = 2; s <= H; ++s) mp += Stencil<P>::template c<T>(s) * ((a[k - s] - uc) + (a[k + s] - uc));
    return kp + mp;
}

// (A u)(i, j) for the operand in an LDS tile covering columns [i0, i0 + W),
// rows [j0, ...): the interior cross where it applies, else the
// closure-aware general form (INNER: every point is interior, the general
// form is not compiled).  The caller's regions contain every point either
// form reads.
template <int P, typename T, bool LT, bool INNER, int W>
__device__ __forceinline__ T au_grid(const OpTab& t, const T* a, int i0, int j0, int i, int j) {
    const int k = (j - j0) * W + (i - i0);
    if (INNER || is_interior(t, i, j)) return au_interior<P, T, W>(a, k);
    const InGrid<T, W> in{a, i0, j0};
    return au_point<P, T, LT>(t, in, i, j);
}

// Halo widths of the fused blocks.  A closure row reads C columns from the
// face, so a region whose points include closure rows must reach the face:
// with R closure rows the first sweep's region B needs R-1 points of halo
// around the output region C and the operand's region A another R-1 (both
// >= the stencil radius H, which interior points need).
template <int R_> struct FusedHaloR {
    static constexpr int R = R_;
    static constexpr int HB = R - 1;
    static constexpr int HA = 2 * (R - 1);
};
template <int P> struct FusedHalo : FusedHaloR<Stencil<P>::R> {};

// Calls f(std::true_type, i, j, k) for the points of the region
// [x0, x0 + W) x [y0, y0 + H) (k = (j - y0) W + (i - x0)) that lie off the
// closure rows of both axes, f(std::false_type, i, j, k) for its other grid
// points and z(k) for its points outside the grid, each set spread over the
// NT threads in turn.  A wave then runs either the plain cross or the
// closure form: the closure points of a face block (a few columns or rows of
// its region) fill a wave or two, instead of one lane of every wave making
// all its lanes wait for the closure form.
template <int NT, int W, int H, class F, class Z>
__device__ __forceinline__ void region_points(int N, int R, int x0, int y0, F&& f, Z&& z) {
    const int tid = threadIdx.x;
    const int gx0 = x0 > 0 ? x0 : 0, gx1 = x0 + W < N ? x0 + W : N;
    const int gy0 = y0 > 0 ? y0 : 0, gy1 = y0 + H < N ? y0 + H : N;
    const int gw = gx1 > gx0 ? gx1 - gx0 : 0, gh = gy1 > gy0 ? gy1 - gy0 : 0;
    // the interior part [cx0, cx1) x [cy0, cy1), clamped into the grid part
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    const int cx0 = clampi(R, gx0, gx0 + gw), cx1 = clampi(N - R, cx0, gx0 + gw);
    const int cy0 = clampi(R, gy0, gy0 + gh), cy1 = clampi(N - R, cy0, gy0 + gh);
    const int wi = cx1 - cx0, hi = cy1 - cy0;
    if (wi > 0 && hi > 0) {
        const float rw = 1.0f / (float)wi;
        for (int m = tid; m < wi * hi; m += NT) {
            const int r = div_small(m, wi, rw);
            const int i = cx0 + (m - r * wi), j = cy0 + r;
            f(std::true_type{}, i, j, (j - y0) * W + (i - x0));
        }
    }
    // the frame: rows [gy0, cy0) and [cy1, gy1) over the grid part's columns,
    // then columns [gx0, cx0) and [cx1, gx1) of the rows [cy0, cy1)
    const int left = cx0 - gx0, right = gx0 + gw - cx1;
    const int n1 = (cy0 - gy0) * gw, n2 = n1 + (gy0 + gh - cy1) * gw;
    const int n3 = n2 + hi * left, n4 = n3 + hi * right;
    if (n4 > 0) {
        const float rgw = gw > 0 ? 1.0f / (float)gw : 0.0f;
        const float rl = left > 0 ? 1.0f / (float)left : 0.0f, rr = right > 0 ? 1.0f / (float)right : 0.0f;
        for (int m = tid; m < n4; m += NT) {
            int i, j;
            if (m < n2) {
                const int mm = m < n1 ? m : m - n1;
                const int r = div_small(mm, gw, rgw);
                i = gx0 + (mm - r * gw);
                j = (m < n1 ? gy0 : cy1) + r;
            } else if (m < n3) {
                const int mm = m - n2;
                const int r = div_small(mm, left, rl);
                i = gx0 + (mm - r * left);
                j = cy0 + r;
            } else {
                const int mm = m - n3;
                const int r = div_small(mm, right, rr);
                i = cx1 + (mm - r * right);
                j = cy0 + r;
            }
            f(std::false_type{}, i, j, (j - y0) * W + (i - x0));
        }
    }
    // outside the grid: whole local rows [0, top) and [H - bot, H), then the
    // local columns [0, lw) and [W - rw, W) of the rows between
    const int top = clampi(-y0, 0, H), bot = clampi(y0 + H - N, 0, H - top);
    const int lw = clampi(-x0, 0, W), rw = clampi(x0 + W - N, 0, W - lw);
    const int hm = H - top - bot;
    const int z1 = top * W, z2 = z1 + bot * W, z3 = z2 + hm * lw, z4 = z3 + hm * rw;
    if (z4 > 0) {
        const float rlw = lw > 0 ? 1.0f / (float)lw : 0.0f, rrw = rw > 0 ? 1.0f / (float)rw : 0.0f;
        for (int m = tid; m < z4; m += NT) {
            int r, c;
            if (m < z2) {
                const int mm = m < z1 ? m : m - z1;
                r = mm / W;
                c = mm - r * W;
                if (m >= z1) r += H - bot;
            } else if (m < z3) {
                const int mm = m - z2, q = div_small(mm, lw, rlw);
                r = top + q;
                c = mm - q * lw;
            } else {
                const int mm = m - z3, q = div_small(mm, rw, rrw);
                r = top + q;
                c = W - rw + (mm - q * rw);
            }
            z(r * W + c);
        }
    }
}

// Scale of the smoothing correction w (b - Au): w / diag (damped Jacobi; the
// inverse diagonal is the constant idint away from the faces) or w (Richardson).
template <typename T, bool LT, bool INNER>
__device__ __forceinline__ T smooth_corr(const OpTab& t, int i, int j, T r, T omega, int use_diag) {
    if (INNER || is_interior(t, i, j)) return r * smooth_scale_int(t, omega, use_diag);
    const T c = omega * r;
    return use_diag ? c / diag_of<T, LT>(t, i, j) : c;
}

// The operator as the fused bodies see it (constant coefficients): its closure
// rows R (the halos), the table staging of face blocks (fetch issues the
// loads, store writes LDS; kStage: the body then waits before the first
// lookup), the smoothing correction and A u on an LDS array.
template <int P, typename T, int NT, bool LT, bool INNER>
struct FusedOpC {
    static constexpr int R = Stencil<P>::R;
    static constexpr bool kStage = LT;
    static constexpr bool kInner = INNER;   // every point the block evaluates is off the closures
    const OpTab& ta;   // kernel-argument table
    OpTab* lt;         // its LDS copy (LT)
    TabStage<P, NT> st;
    __device__ FusedOpC(const OpTab& a, OpTab* l) : ta(a), lt(l) {}
    __device__ __forceinline__ const OpTab& t() const {
        if constexpr (LT) return *lt;
        else return ta;
    }
    __device__ __forceinline__ void fetch(int, int) {
        if constexpr (LT) st.fetch(ta.dev);
    }
    __device__ __forceinline__ void store() {
        if constexpr (LT) st.store(lt);
    }
    // INT: the caller knows (i, j) is off the closure rows of both axes.
    template <bool INT = false>
    __device__ __forceinline__ T corr(int i, int j, T r, T omega, int use_diag) const {
        return smooth_corr<T, LT, INNER || INT>(t(), i, j, r, omega, use_diag);
    }
    template <int W, bool INT = false>
    __device__ __forceinline__ T au(const T* a, int i0, int j0, int i, int j) const {
        return au_grid<P, T, LT, INNER || INT, W>(t(), a, i0, j0, i, j);
    }
};

struct FusedCtl {
    const int* done;    // nullable: per-problem flags, stride ds ints
    int ds;
    int nbx, nby;       // block grid of one problem
    double* partial;    // prolong_smooth2 with dotv: per-block partials of dotv'out
    Batch bt;           // sf: fine elements, sc: coarse elements per problem
};

// ---------------------------------------------------------------------------
// x2 = S(S(0)) and rc = scale * FW(b - A x2)
// ---------------------------------------------------------------------------
// The done flag, the table words (face blocks) and region A's loads are all
// issued before the first wait; a converged problem's block returns false
// after that wait, before any store.  op: the operator (FusedOpC, or the
// variable-coefficient one); T the arithmetic and LDS type, TG the element
// type of the grid functions in global memory.
template <int TI, int J, int NT, typename T, typename TG, class OP>
__device__ __forceinline__ bool sr_body(OP& op, int N, const TG* __restrict__ b, TG* __restrict__ x2g,
                                        TG* __restrict__ rc, T omega, int use_diag, T scale, const int* done, int bx,
                                        int by, int Nc, T* Bt, T* U, T* V) {
    constexpr int HA = FusedHaloR<OP::R>::HA, HB = FusedHaloR<OP::R>::HB;
    constexpr int JR = J + 1;                     // fine residual rows (region C)
    constexpr int CI = TI / 2 - 1, CJ = J / 2;    // coarse points owned per block
    constexpr int AW = TI + 2 * HA, AH = JR + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = JR + 2 * HB;
    const int tid = threadIdx.x;
    int Ic0 = bx * CI < Nc - CI ? bx * CI : Nc - CI;
    int Jc0 = by * CJ < Nc - CJ ? by * CJ : Nc - CJ;
    if (Ic0 < 0) Ic0 = 0;
    if (Jc0 < 0) Jc0 = 0;
    const int ci0 = 2 * Ic0 - 1, cj0 = 2 * Jc0 - 1;   // region C origin
    const int ai0 = ci0 - HA, aj0 = cj0 - HA;   // region A origin
    const int bi0 = ci0 - HB, bj0 = cj0 - HB;   // region B origin
    // A: b and x1 = w D^-1 b (all loads issued before the first use)
    {
        const int dflag = done ? *done : 0;
        op.fetch(ai0, aj0);
        constexpr int NA = (AW * AH + NT - 1) / NT;
        T v[NA];
#pragma unroll
        for (int m = 0; m < NA; ++m) {
            const int k = tid + m * NT;
            const int r = k / AW, c = k - r * AW;
            const int i = ai0 + c, j = aj0 + r;
            v[m] = (k < AW * AH && i >= 0 && i < N && j >= 0 && j < N) ? (T)b[gidx(i, j, N)] : T(0);
        }
        if (dflag) return false;
        op.store();
        if constexpr (OP::kInner) {
            if constexpr (OP::kStage) __syncthreads();
#pragma unroll
            for (int m = 0; m < NA; ++m) {
                const int k = tid + m * NT;
                if (k < AW * AH) {
                    const int r = k / AW, c = k - r * AW;
                    Bt[k] = v[m];
                    U[k] = op.corr(ai0 + c, aj0 + r, v[m], omega, use_diag);
                }
            }
        } else {
            // face block: b to LDS first, then x1 with the closure points
            // (a division by their diagonal each) gathered into a few waves
#pragma unroll
            for (int m = 0; m < NA; ++m) {
                const int k = tid + m * NT;
                if (k < AW * AH) Bt[k] = v[m];
            }
            __syncthreads();
            region_points<NT, AW, AH>(N, OP::R, ai0, aj0, [&](auto in_, int i, int j, int k) {
                constexpr bool INT = decltype(in_)::value;
                U[k] = op.template corr<INT>(i, j, Bt[k], omega, use_diag);
            }, [](int) {});   // (x1 is read at grid points only)
        }
    }
    __syncthreads();
    // B: x2 = x1 + w D^-1 (b - A x1)
    if constexpr (OP::kInner) {
        for (int k = tid; k < BW * BH; k += NT) {
            const int r = k / BW, c = k - r * BW;
            const int i = bi0 + c, j = bj0 + r;
            const int ka = (r + HA - HB) * AW + (c + HA - HB);
            const T Au = op.template au<AW>(U, ai0, aj0, i, j);
            V[k] = U[ka] + op.corr(i, j, Bt[ka] - Au, omega, use_diag);
        }
    } else {
        region_points<NT, BW, BH>(N, OP::R, bi0, bj0, [&](auto in_, int i, int j, int k) {
            constexpr bool INT = decltype(in_)::value;
            const int ka = (j - aj0) * AW + (i - ai0);
            const T Au = op.template au<AW, INT>(U, ai0, aj0, i, j);
            V[k] = U[ka] + op.template corr<INT>(i, j, Bt[ka] - Au, omega, use_diag);
        }, [](int) {});   // (V is read at grid points only: no fill outside the grid)
    }
    __syncthreads();
    // x2 at the fine points this block owns
    {
        const int oi0 = 2 * bx * CI, oj0 = 2 * by * CJ;
        const int oi1 = oi0 + 2 * CI < N ? oi0 + 2 * CI : N;
        const int oj1 = oj0 + 2 * CJ < N ? oj0 + 2 * CJ : N;
        const int ow = oi1 - oi0;
        for (int k = tid; k < ow * (oj1 - oj0); k += NT) {
            const int r = k / ow, c = k - r * ow;
            const int i = oi0 + c, j = oj0 + r;
            x2g[gidx(i, j, N)] = (TG)V[(j - bj0) * BW + (i - bi0)];
        }
    }
    // C: residual r = b - A x2 into U's storage (x1 is no longer read)
    T* R = U;
    if constexpr (OP::kInner) {
        for (int k = tid; k < TI * JR; k += NT) {
            const int r = k / TI, c = k - r * TI;
            const int i = ci0 + c, j = cj0 + r;
            R[k] = Bt[(r + HA) * AW + (c + HA)] - op.template au<BW>(V, bi0, bj0, i, j);
        }
    } else {
        region_points<NT, TI, JR>(N, OP::R, ci0, cj0, [&](auto in_, int i, int j, int k) {
            constexpr bool INT = decltype(in_)::value;
            R[k] = Bt[(j - aj0) * AW + (i - ai0)] - op.template au<BW, INT>(V, bi0, bj0, i, j);
        }, [&](int k) { R[k] = T(0); });
    }
    __syncthreads();
    // Full weighting [1 2 1; 2 4 2; 1 2 1] / 16 (times scale) of the owned coarse points.
    {
        const int I0 = bx * CI, J0 = by * CJ;
        const int I1 = I0 + CI < Nc ? I0 + CI : Nc, J1 = J0 + CJ < Nc ? J0 + CJ : Nc;
        const int cw = I1 - I0;
        for (int o = tid; o < cw * (J1 - J0); o += NT) {
            const int rr = o / cw, cc = o - rr * cw;
            const int I = I0 + cc, Jc = J0 + rr;
            const int lc = 2 * I - ci0, lr = 2 * Jc - cj0;   // fine (2I, 2Jc) in region C
            const T* r0 = R + lr * TI + lc;
            const T s0 = r0[0] + T(0.5) * (r0[-1] + r0[1]);
            const T sm = r0[-TI] + T(0.5) * (r0[-TI - 1] + r0[-TI + 1]);
            const T sp = r0[TI] + T(0.5) * (r0[TI - 1] + r0[TI + 1]);
            rc[gidx(I, Jc, Nc)] = (TG)(scale * T(0.25) * (s0 + T(0.5) * (sm + sp)));
        }
    }
    return true;
}

template <int P, int TI, int J, int NT, typename T>
__global__ __launch_bounds__(NT) void smooth2_restrict(const T* __restrict__ b, T* __restrict__ x2g,
                                                      T* __restrict__ rc, T omega, int use_diag, T scale, OpTab t,
                                                      FusedCtl ctl) {
    constexpr int HA = FusedHalo<P>::HA, HB = FusedHalo<P>::HB;
    constexpr int JR = J + 1;
    constexpr int AW = TI + 2 * HA, AH = JR + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = JR + 2 * HB;
    const int nb = ctl.nbx * ctl.nby;
    const int ball = xcd_remap((int)blockIdx.x, nb * ctl.bt.nprob);
    const int q = ball / nb, blk = ball - q * nb;
    const int* done = ctl.done ? ctl.done + (size_t)q * ctl.ds : nullptr;
    b += q * ctl.bt.sf;
    x2g += q * ctl.bt.sf;
    rc += q * ctl.bt.sc;
    const int bx = blk % ctl.nbx, by = blk / ctl.nbx;
    const int Nc = (t.N + 1) / 2;
    __shared__ T Bt[AW * AH];
    __shared__ T U[AW * AH];
    __shared__ T V[BW * BH];
    constexpr int CI = TI / 2 - 1, CJ = J / 2;
    // Interior block: regions A lie inside the closure-free zone.
    int Ic0 = bx * CI < Nc - CI ? bx * CI : Nc - CI;
    int Jc0 = by * CJ < Nc - CJ ? by * CJ : Nc - CJ;
    if (Ic0 < 0) Ic0 = 0;
    if (Jc0 < 0) Jc0 = 0;
    const int ai0 = 2 * Ic0 - 1 - HA, aj0 = 2 * Jc0 - 1 - HA;
    const bool inner = ai0 >= t.R && ai0 + AW <= t.N - t.R && aj0 >= t.R && aj0 + AH <= t.N - t.R;
    if (inner) {
        FusedOpC<P, T, NT, false, true> op(t, nullptr);
        sr_body<TI, J, NT, T, T>(op, t.N, b, x2g, rc, omega, use_diag, scale, done, bx, by, Nc, Bt, U, V);
    } else {
        __shared__ OpTab ltab;
        FusedOpC<P, T, NT, true, false> op(t, &ltab);
        sr_body<TI, J, NT, T, T>(op, t.N, b, x2g, rc, omega, use_diag, scale, done, bx, by, Nc, Bt, U, V);
    }
}

// ---------------------------------------------------------------------------
// out = S(S(x + P e)), optionally with block partials of dotv' out
// ---------------------------------------------------------------------------
// As sr_body: flag, table and data loads in one round trip; false (nothing
// stored) for a converged problem.
template <int TI, int J, int NT, typename T, typename TG, class OP>
__device__ __forceinline__ bool ps_body(OP& op, int N, const TG* __restrict__ x, const TG* __restrict__ ec,
                                        const TG* __restrict__ b, TG* __restrict__ out, const TG* __restrict__ dotv,
                                        T omega, int use_diag, const int* done, int bx, int by, T* U, T* E, T* Bt,
                                        T* V, T& acc) {
    constexpr int HA = FusedHaloR<OP::R>::HA, HB = FusedHaloR<OP::R>::HB;
    constexpr int AW = TI + 2 * HA, AH = J + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = J + 2 * HB;
    constexpr int EW = AW / 2 + 2, EH = AH / 2 + 2;
    const int Nc = (N + 1) / 2;
    const int tid = threadIdx.x;
    int i0 = bx * TI < N - TI ? bx * TI : N - TI;
    int j0 = by * J < N - J ? by * J : N - J;
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    const int ai0 = i0 - HA, aj0 = j0 - HA;
    const int bi0 = i0 - HB, bj0 = j0 - HB;
    const int ei0 = ai0 >> 1, ej0 = aj0 >> 1;   // coarse patch origin (floor)
    // Coarse patch, x on A and b on B: every load issued before the first use.
    {
        constexpr int NE = (EW * EH + NT - 1) / NT;
        constexpr int NA = (AW * AH + NT - 1) / NT;
        constexpr int NB = (BW * BH + NT - 1) / NT;
        const int dflag = done ? *done : 0;
        op.fetch(ai0, aj0);
        T ve[NE], va[NA], vb[NB];
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            const int k = tid + m * NT;
            const int r = k / EW, c = k - r * EW;
            const int I = ei0 + c, Jc = ej0 + r;
            ve[m] = (k < EW * EH && I >= 0 && I < Nc && Jc >= 0 && Jc < Nc) ? (T)ec[gidx(I, Jc, Nc)] : T(0);
        }
#pragma unroll
        for (int m = 0; m < NA; ++m) {
            const int k = tid + m * NT;
            const int r = k / AW, c = k - r * AW;
            const int i = ai0 + c, j = aj0 + r;
            va[m] = (k < AW * AH && i >= 0 && i < N && j >= 0 && j < N) ? (T)x[gidx(i, j, N)] : T(0);
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int k = tid + m * NT;
            const int r = k / BW, c = k - r * BW;
            const int i = bi0 + c, j = bj0 + r;
            vb[m] = (k < BW * BH && i >= 0 && i < N && j >= 0 && j < N) ? (T)b[gidx(i, j, N)] : T(0);
        }
        if (dflag) return false;
        op.store();   // read after the barriers below
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            const int k = tid + m * NT;
            if (k < EW * EH) E[k] = ve[m];
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int k = tid + m * NT;
            if (k < BW * BH) Bt[k] = vb[m];
        }
        __syncthreads();
        // U = x + P e on A (bilinear prolongation from the coarse patch)
#pragma unroll
        for (int m = 0; m < NA; ++m) {
            const int k = tid + m * NT;
            if (k < AW * AH) {
                const int r = k / AW, c = k - r * AW;
                const int i = ai0 + c, j = aj0 + r;
                T u = T(0);
                if (i >= 0 && i < N && j >= 0 && j < N) {
                    const int ic = i >> 1, jc = j >> 1;
                    const int ic1 = ic + 1 < Nc ? ic + 1 : ic, jc1 = jc + 1 < Nc ? jc + 1 : jc;
                    const T e00 = E[(jc - ej0) * EW + (ic - ei0)], e10 = E[(jc - ej0) * EW + (ic1 - ei0)];
                    const T e01 = E[(jc1 - ej0) * EW + (ic - ei0)], e11 = E[(jc1 - ej0) * EW + (ic1 - ei0)];
                    T pv;
                    if (!(i & 1)) pv = (!(j & 1)) ? e00 : (e00 + e01) * T(0.5);
                    else pv = (!(j & 1)) ? (e00 + e10) * T(0.5) : (e00 + e10 + e01 + e11) * T(0.25);
                    u = va[m] + pv;
                }
                U[k] = u;
            }
        }
    }
    __syncthreads();
    // B: v1 = u + w D^-1 (b - A u)
    if constexpr (OP::kInner) {
        for (int k = tid; k < BW * BH; k += NT) {
            const int r = k / BW, c = k - r * BW;
            const int i = bi0 + c, j = bj0 + r;
            const T Au = op.template au<AW>(U, ai0, aj0, i, j);
            V[k] = U[(r + HA - HB) * AW + (c + HA - HB)] + op.corr(i, j, Bt[k] - Au, omega, use_diag);
        }
    } else {
        region_points<NT, BW, BH>(N, OP::R, bi0, bj0, [&](auto in_, int i, int j, int k) {
            constexpr bool INT = decltype(in_)::value;
            const T Au = op.template au<AW, INT>(U, ai0, aj0, i, j);
            V[k] = U[(j - aj0) * AW + (i - ai0)] + op.template corr<INT>(i, j, Bt[k] - Au, omega, use_diag);
        }, [](int) {});   // (V is read at grid points only: no fill outside the grid)
    }
    __syncthreads();
    // C: out = v1 + w D^-1 (b - A v1) on the owned points
    acc = T(0);
    {
        const int oi0 = bx * TI, oj0 = by * J;
        const int oi1 = oi0 + TI < N ? oi0 + TI : N, oj1 = oj0 + J < N ? oj0 + J : N;
        const int ow = oi1 - oi0;
        auto point = [&](auto in_, int i, int j) {
            constexpr bool INT = decltype(in_)::value;
            const int kb = (j - bj0) * BW + (i - bi0);
            const T Av = op.template au<BW, INT>(V, bi0, bj0, i, j);
            const T o = V[kb] + op.template corr<INT>(i, j, Bt[kb] - Av, omega, use_diag);
            const size_t g = gidx(i, j, N);
            out[g] = (TG)o;
            if (dotv) acc += (T)dotv[g] * o;
        };
        if constexpr (OP::kInner) {
            for (int k = tid; k < ow * (oj1 - oj0); k += NT) {
                const int r = k / ow, c = k - r * ow;
                point(std::false_type{}, oi0 + c, oj0 + r);
            }
        } else {
            // the owned points (the block's whole TI x J tile, cut at the grid's
            // far faces: the points past them go to z, which ignores them)
            region_points<NT, TI, J>(N, OP::R, oi0, oj0,
                                     [&](auto in_, int i, int j, int) { point(in_, i, j); }, [](int) {});
        }
    }
    return true;
}

template <int P, int TI, int J, int NT, typename T>
__global__ __launch_bounds__(NT) void prolong_smooth2(const T* __restrict__ x, const T* __restrict__ ec,
                                                     const T* __restrict__ b, T* __restrict__ out,
                                                     const T* __restrict__ dotv, T omega, int use_diag, OpTab t,
                                                     FusedCtl ctl) {
    constexpr int HA = FusedHalo<P>::HA, HB = FusedHalo<P>::HB;
    constexpr int AW = TI + 2 * HA, AH = J + 2 * HA;
    constexpr int BW = TI + 2 * HB, BH = J + 2 * HB;
    constexpr int EW = AW / 2 + 2, EH = AH / 2 + 2;
    const int nb = ctl.nbx * ctl.nby;
    const int ball = xcd_remap((int)blockIdx.x, nb * ctl.bt.nprob);
    const int q = ball / nb, blk = ball - q * nb;
    const int* done = ctl.done ? ctl.done + (size_t)q * ctl.ds : nullptr;
    x += q * ctl.bt.sf;
    b += q * ctl.bt.sf;
    out += q * ctl.bt.sf;
    ec += q * ctl.bt.sc;
    if (dotv) dotv += q * ctl.bt.sf;
    const int bx = blk % ctl.nbx, by = blk / ctl.nbx;
    const int N = t.N;
    __shared__ T U[AW * AH];
    __shared__ T E[EW * EH];
    __shared__ T Bt[BW * BH];
    __shared__ T V[BW * BH];
    int i0 = bx * TI < N - TI ? bx * TI : N - TI;
    int j0 = by * J < N - J ? by * J : N - J;
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    const int ai0 = i0 - HA, aj0 = j0 - HA;
    const bool inner = ai0 >= t.R && ai0 + AW <= N - t.R && aj0 >= t.R && aj0 + AH <= N - t.R;
    T acc = T(0);
    bool ran;
    if (inner) {
        FusedOpC<P, T, NT, false, true> op(t, nullptr);
        ran = ps_body<TI, J, NT, T, T>(op, N, x, ec, b, out, dotv, omega, use_diag, done, bx, by, U, E, Bt, V, acc);
    } else {
        __shared__ OpTab ltab;
        FusedOpC<P, T, NT, true, false> op(t, &ltab);
        ran = ps_body<TI, J, NT, T, T>(op, N, x, ec, b, out, dotv, omega, use_diag, done, bx, by, U, E, Bt, V, acc);
    }
    if (!ran) return;   // converged: every thread of the block left before its first barrier
    if (dotv) {
        __syncthreads();
        const T s = block_sum<T, NT>(acc);
        if (threadIdx.x == 0) ctl.partial[(size_t)q * nb + blk] = (double)s;
    }
}

}  // namespace sbp
