// sbp_stencil.h -- gfx950 stencil pass for the matrix-free SBP-SAT Poisson path.
//
// The operator (SPD A = -H~(D2 + SAT), preconditioned_CG/level_2_multigrid_new.jl:141-185)
// is applied through its tensor form
//     (A u)(i,j) = hw[i] * sum_j' k[j][j'] u(i,j') + hw[j] * sum_i' m[i][i'] u(i',j)
// with dimensionless 1-D factors built on the host from the reference's SBP
// tables (see sbp_tables.h).  Away from the faces both factors are the
// interior stencil -d (p=2: [-1 2 -1], p=4: [1/12 -4/3 5/2 -4/3 1/12]) and
// hw = 1, so the interior is a (2p+1)-point cross; near the faces the first
// R rows of k and m are a dense R x C closure block (mirrored at the far face).
//
// Kernel shape (the hot path): a workgroup owns a tile of TI consecutive i
// (fast, coalesced) by J consecutive j.  Each thread loads its own column
// u(i, j0-H .. j0+J-1+H) into registers up front (all loads independent, so
// a wave keeps J+2H row loads in flight), stores the J centre values in an
// LDS tile (plus H halo columns each side), waits once, and then forms the
// j-direction stencil from registers and the i-direction stencil from LDS.
// HBM sees each element once; the H-row halo re-reads of j-adjacent tiles are
// served by the XCD's L2 because the block index is remapped so that each XCD
// walks a contiguous band of tiles (dispatch is round-robin over the 8 XCDs).
//
// What the pass does with (u, Au) is a policy: IN says how u(i,j) is formed
// (plain load, Jacobi-from-zero, x + P e_c, r + beta p), OUT what is written
// (sign * Au, residual, Jacobi update, fused inner products).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

namespace sbp {

constexpr int kMaxR = 11;  // closure rows
constexpr int kMaxC = 16;  // closure columns

struct OpTab {
    int N;      // points per side
    int R;      // closure rows of k and m (p = 2, 4, 6, 8, 10: 3, 4, 6, 8, 11)
    int C;      // closure columns (4, 6, 9, 12, 16)
    int RH;     // number of non-unit H weights (1, 4, 6, 8, 11)
    int H;      // half-width of the interior stencil (p / 2)
    double cint[6];   // interior stencil of the 1-D factors: cint[s] at offsets +-s
    double hw[kMaxR];
    double mb[kMaxR][kMaxC];
    double kb[kMaxR][kMaxC];
    double md[kMaxR];   // mb[r][r]
    double kd[kMaxR];   // kb[r][r]
    double mrs[kMaxR];  // row sums of mb (exactly 0: m 1 = 0)
    double krs[kMaxR];  // row sums of kb (the Dirichlet SAT "mass")
    double mid;         // interior diagonal of the 1-D factor (2 or 5/2)
    double idint;       // 1 / (2 mid): inverse diagonal of A away from the faces
    // Device-memory copy of this table.  Kernels take the table by value in
    // the kernel-argument segment for its scalar fields; lookups with a
    // per-thread row index go through this pointer (an indexed access into
    // the argument segment would make the compiler copy the table to scratch).
    const OpTab* dev;
};

// Every row of the 1-D factors is evaluated in difference form,
//     sum_c a[r][c] u_c = rowsum_r u_r + sum_{c != r} a[r][c] (u_c - u_r),
// which is exact algebra but keeps the zero row sums of the interior and of
// the Neumann-side closure exact in any precision: with rounded FP32
// coefficients the plain sum leaves a spurious O(1e-8) diagonal term, as
// large as the smallest eigenvalue of A at N ~ 4k, while differences of
// neighbouring values of a smooth field are formed almost exactly.

// Problems stacked in one launch (a batched handle): problem q's grid
// functions start q * sf elements (this level) or q * sc elements (the next
// coarser level) after problem 0's, and its device scalar block q * ss doubles
// after problem 0's.  A single problem is nprob = 1.
struct Batch {
    int nprob;
    long long sf;
    long long sc;
    int ss;
};

// Interior 1-D stencil s[-H..H] of -d (diagonal_sbp.jl:193, 201, 213, 227, 394),
// the closure block width C of the order (columns of the first R rows) and R.
template <int P> struct Stencil;
template <> struct Stencil<2> {
    static constexpr int H = 1, C = 4, R = 3;
    template <typename T> __device__ static T c(int t) { return t == 0 ? T(2) : T(-1); }
};
template <> struct Stencil<4> {
    static constexpr int H = 2, C = 6, R = 4;
    template <typename T> __device__ static T c(int t) {
        return t == 0 ? T(2.5) : ((t == 1 || t == -1) ? T(-4.0 / 3.0) : T(1.0 / 12.0));
    }
};
template <> struct Stencil<6> {
    static constexpr int H = 3, C = 9, R = 6;
    template <typename T> __device__ static T c(int t) {
        const int a = t < 0 ? -t : t;
        return a == 0 ? T(49.0 / 18.0) : a == 1 ? T(-3.0 / 2.0) : a == 2 ? T(3.0 / 20.0) : T(-1.0 / 90.0);
    }
};
template <> struct Stencil<8> {
    static constexpr int H = 4, C = 12, R = 8;
    template <typename T> __device__ static T c(int t) {
        const int a = t < 0 ? -t : t;
        return a == 0 ? T(205.0 / 72.0)
             : a == 1 ? T(-8.0 / 5.0)
             : a == 2 ? T(1.0 / 5.0)
             : a == 3 ? T(-8.0 / 315.0)
                      : T(1.0 / 560.0);
    }
};
template <> struct Stencil<10> {
    static constexpr int H = 5, C = 16, R = 11;
    template <typename T> __device__ static T c(int t) {
        const int a = t < 0 ? -t : t;
        return a == 0 ? T(5269.0 / 1800.0)
             : a == 1 ? T(-5.0 / 3.0)
             : a == 2 ? T(5.0 / 21.0)
             : a == 3 ? T(-5.0 / 126.0)
             : a == 4 ? T(5.0 / 1008.0)
                      : T(-1.0 / 3150.0);
    }
};

__device__ __forceinline__ size_t gidx(int i, int j, int N) { return (size_t)j * (size_t)N + (size_t)i; }

// m / d for 0 <= m < 2^20, 1 <= d <= 2^10, given rd = 1.0f / d (a float
// estimate, corrected by one): the index arithmetic of point loops whose
// widths are known only at run time.
__device__ __forceinline__ int div_small(int m, int d, float rd) {
    int r = (int)((float)m * rd);
    if (r * d > m) --r;
    else if ((r + 1) * d <= m) ++r;
    return r;
}

// Storage location of grid point (i, j): the offset of its row and its column
// (0 <= i < N where it is accessed).  The policies read and write grid
// functions only through ld/st, so the addressing form is chosen here.
// (Measured: buffer-resource loads with the row in the scalar base and a
// shared 32-bit column offset cut vector registers but doubled the scalar
// instructions per wave and ran the smoothing passes 2-17 % slower.)
struct Loc {
    size_t row;
    int i;
};
__device__ __forceinline__ Loc loc(int i, int j, int N) { return Loc{(size_t)j * (size_t)N, i}; }
template <typename T>
__device__ __forceinline__ T* elem(T* a, Loc q) { return a + q.row + (size_t)q.i; }
template <typename T>
__device__ __forceinline__ T ld(const T* a, Loc q) { return *elem(a, q); }
template <typename T>
__device__ __forceinline__ void st(T* a, Loc q, T v) { *elem(a, q) = v; }

// Where per-thread table lookups go: LT = false, the table's device-memory
// copy (t.dev); LT = true, the table object itself, which the caller has
// placed in LDS (its fields are then plain LDS reads).
template <bool LT>
__device__ __forceinline__ const OpTab& lut(const OpTab& t) {
    if constexpr (LT) return t;
    else return *t.dev;
}

template <typename T, bool LT = false>
__device__ __forceinline__ T hw_of(const OpTab& t, int r) {
    const int rr = r < t.N - 1 - r ? r : t.N - 1 - r;   // distance to the nearer face
    return rr < t.RH ? (T)lut<LT>(t).hw[rr] : T(1);
}

template <typename T, bool LT = false>
__device__ __forceinline__ T diag_of(const OpTab& t, int i, int j) {
    const int jj = j < t.N - 1 - j ? j : t.N - 1 - j;
    const int ii = i < t.N - 1 - i ? i : t.N - 1 - i;
    const T kj = jj < t.R ? (T)lut<LT>(t).kd[jj] : (T)t.mid;
    const T mi = ii < t.R ? (T)lut<LT>(t).md[ii] : (T)t.mid;
    return hw_of<T, LT>(t, i) * kj + hw_of<T, LT>(t, j) * mi;
}

// Rows 0..N-1 are remapped so that the 8 XCDs (round-robin dispatch) each get
// a contiguous run of logical blocks.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int full = nb & ~7;
    if (b >= full) return b;
    const int per = full >> 3;
    return (b & 7) * per + (b >> 3);
}

// Scale of a smoothing correction away from the faces, where the inverse
// diagonal is the constant idint: w idint (damped Jacobi) or w (Richardson).
// Uniform over a launch, so the compiler hoists it and each point costs one
// multiply.
template <typename T>
__device__ __forceinline__ T smooth_scale_int(const OpTab& t, T omega, int use_diag) {
    return use_diag ? omega * (T)t.idint : omega;
}

// The correction w D^-1 r (or w r) at (i, j); INT: (i, j) is off the closure rows.
template <typename T, bool INT, bool LT>
__device__ __forceinline__ T smooth_corr_at(const OpTab& t, int i, int j, T r, T omega, int use_diag) {
    if constexpr (INT) return r * smooth_scale_int(t, omega, use_diag);
    const T c = omega * r;
    return use_diag ? c / diag_of<T, LT>(t, i, j) : c;
}

// ---------------------------------------------------------------------------
// Input policies: value of the operand at (i, j)
// ---------------------------------------------------------------------------
// Every input policy splits a value into a raw fetch (plain loads only) and a
// transform, so that the guarded face-tile paths issue all their loads before
// the first use (a transform inside a guarded load would serialise the
// round trips).  load() = make(fetch()).
template <typename T> struct InPlain {
    const T* u;
    __device__ void at(int q, const Batch& bt) { u += q * bt.sf; }
    static constexpr bool kCol = false;
    static constexpr bool kRawRhs = false;
    using Raw = T;
    __device__ void init() {}
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab& t, int i, int j) const { return ld(u, loc(i, j, t.N)); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab&, int, int, const Raw& r) const { return r; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab& t, int i, int j) const { return ld(u, loc(i, j, t.N)); }
};

// First smoothing sweep from x = 0: x1 = w D^-1 b (damped Jacobi) or
// x1 = w b (Richardson, use_diag = false).
template <typename T> struct InSmooth0 {
    const T* b; T omega; int use_diag;
    __device__ void at(int q, const Batch& bt) { b += q * bt.sf; }
    static constexpr bool kCol = false;
    static constexpr bool kRawRhs = true;   // the raw fetch at a point is b there
    using Raw = T;
    __device__ void init() {}
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab& t, int i, int j) const { return ld(b, loc(i, j, t.N)); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab& t, int i, int j, const Raw& r) const {
        if constexpr (INT) return r * smooth_scale_int(t, omega, use_diag);
        const T v = omega * r;
        return use_diag ? v / diag_of<T, LT>(t, i, j) : v;
    }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab& t, int i, int j) const {
        return make<INT, LT>(t, i, j, fetch<INT, LT>(t, i, j));
    }
};

// x + P e_c with P the bilinear prolongation from the coarse side Nc = (N+1)/2.
template <typename T> struct InProlong {
    const T* x; const T* ec; int Nc;
    __device__ void at(int q, const Batch& bt) { x += q * bt.sf; ec += q * bt.sc; }
    static constexpr bool kCol = true;
    static constexpr bool kRawRhs = false;
    __device__ void init() {}
    // Column window w[k] = x(i, jbeg+k) + (P ec)(i, jbeg+k), k < W, with jbeg of
    // parity PAR: each coarse row is loaded once (W/2+2 rows) instead of four
    // coarse loads per fine point.  No bounds tests: rows jbeg .. jbeg+W-1 and
    // the column i lie inside the grid (an odd i < N has i/2 + 1 < Nc).
    template <int W, int PAR>
    __device__ __forceinline__ void load_col(const OpTab& t, int i, int jbeg, T* w) const {
        constexpr int NC = ((PAR + W - 1) >> 1) + 2;
        const int ic = i >> 1, jc0 = jbeg >> 1;
        const bool odd = i & 1;
        T c[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            const T e0 = ld(ec, loc(ic, jc0 + m, Nc));
            c[m] = odd ? (e0 + ld(ec, loc(ic + 1, jc0 + m, Nc))) * T(0.5) : e0;
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int s = PAR + k;
            const T v = (s & 1) ? (c[s >> 1] + c[(s >> 1) + 1]) * T(0.5) : c[s >> 1];
            w[k] = ld(x, loc(i, jbeg + k, t.N)) + v;
        }
    }
    struct Raw { T x, e00, e10, e01, e11; };
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab& t, int i, int j) const {
        const int ic = i >> 1, jc = j >> 1;
        const int ic1 = ic + 1 < Nc ? ic + 1 : ic, jc1 = jc + 1 < Nc ? jc + 1 : jc;
        Raw r;
        r.x = ld(x, loc(i, j, t.N));
        r.e00 = ld(ec, loc(ic, jc, Nc));
        r.e10 = ld(ec, loc(ic1, jc, Nc));
        r.e01 = ld(ec, loc(ic, jc1, Nc));
        r.e11 = ld(ec, loc(ic1, jc1, Nc));
        return r;
    }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab&, int i, int j, const Raw& r) const {
        T v;
        if (!(i & 1)) v = (!(j & 1)) ? r.e00 : (r.e00 + r.e01) * T(0.5);
        else v = (!(j & 1)) ? (r.e00 + r.e10) * T(0.5) : (r.e00 + r.e10 + r.e01 + r.e11) * T(0.25);
        return r.x + v;
    }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab& t, int i, int j) const {
        return make<INT, LT>(t, i, j, fetch<INT, LT>(t, i, j));
    }
};

// r + beta * p with beta read from the device scalar block.
template <typename T> struct InAxpy {
    const T* r; const T* p; const double* beta; T bt;
    __device__ void at(int q, const Batch& bt) { r += q * bt.sf; p += q * bt.sf; beta += q * bt.ss; }
    static constexpr bool kCol = false;
    static constexpr bool kRawRhs = false;
    __device__ void init() { bt = (T)(*beta); }
    struct Raw { T r, p; };
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ Raw fetch(const OpTab& t, int i, int j) const {
        const Loc q = loc(i, j, t.N);
        Raw v;
        v.r = ld(r, q);
        v.p = ld(p, q);
        return v;
    }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T make(const OpTab&, int, int, const Raw& v) const { return v.r + bt * v.p; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ T load(const OpTab& t, int i, int j) const {
        const Loc q = loc(i, j, t.N);
        return ld(r, q) + bt * ld(p, q);
    }
};

// ---------------------------------------------------------------------------
// Output policies
// ---------------------------------------------------------------------------
// A policy stores kNS grid functions per point: vals() gives their values at
// (i, j) from the operand u, (A u)(i, j) and what pre() loaded there, dst<K>()
// the K-th destination array (its element type may differ from T: the value
// is converted on store), and with kDot, dterm() the point's term of the fused
// inner product (accumulated in T).  Values are pure functions of those
// inputs, so a pass may form them at the computing thread and store them from
// another lane (the line-aligned stores below).  kAlign: the policy runs on
// the line-aligned tiles when its level has them (measured per pass at
// 4097^2: the apply and the Krylov pass gain; the smoothing passes, whose
// held values raise the register count, lose).
template <typename T> struct OutStore {            // o = s * Au
    T* o; T s;
    __device__ void at(int q, const Batch& bt) { o += q * bt.sf; }
    static constexpr bool kDot = false;
    static constexpr bool kAlign = true;
    static constexpr int kNS = 1;
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc) const { return T(0); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T, T Au, const Aux&, T* v) const { v[0] = s * Au; }
    template <int K> __device__ __forceinline__ T* dst() const { return o; }
};

template <typename T> struct OutStoreDot {         // o = Au ; acc += u Au
    T* o;
    __device__ void at(int q, const Batch& bt) { o += q * bt.sf; }
    static constexpr bool kDot = true;
    static constexpr bool kAlign = true;
    static constexpr int kNS = 1;
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc) const { return T(0); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T, T Au, const Aux&, T* v) const { v[0] = Au; }
    __device__ __forceinline__ T dterm(T u, T Au, const Aux&) const { return u * Au; }
    template <int K> __device__ __forceinline__ T* dst() const { return o; }
};

template <typename T> struct OutWriteDot {         // w = u (materialise the formed operand); o = Au ; acc += u Au
    T* o; T* w;
    __device__ void at(int q, const Batch& bt) { o += q * bt.sf; w += q * bt.sf; }
    static constexpr bool kDot = true;
    static constexpr bool kAlign = false;
    static constexpr int kNS = 2;
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc) const { return T(0); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T u, T Au, const Aux&, T* v) const {
        v[0] = u; v[1] = Au;
    }
    __device__ __forceinline__ T dterm(T u, T Au, const Aux&) const { return u * Au; }
    template <int K> __device__ __forceinline__ T* dst() const {
        if constexpr (K == 0) return w;
        else return o;
    }
};

template <typename T, typename TO> struct OutResid {  // r = b - Au (optionally cast), acc += r^2
    const T* b; TO* r;
    __device__ void at(int q, const Batch& bt) { b += q * bt.sf; r += q * bt.sf; }
    static constexpr bool kDot = true;
    static constexpr bool kAlign = false;
    static constexpr int kNS = 1;
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc q) const { return ld(b, q); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T, T Au, const Aux& bq, T* v) const { v[0] = bq - Au; }
    __device__ __forceinline__ T dterm(T, T Au, const Aux& bq) const { const T rv = bq - Au; return rv * rv; }
    template <int K> __device__ __forceinline__ TO* dst() const { return r; }
};

// Smoothing sweep x_new = u + w (b - Au) / diag (damped Jacobi,
// jacobi_brittany!) or u + w (b - Au) (Richardson, use_diag = 0).
template <typename T> struct OutSmooth {
    const T* b; T* xo; T omega; int use_diag;
    __device__ void at(int q, const Batch& bt) { b += q * bt.sf; xo += q * bt.sf; }
    static constexpr bool kDot = false;
    static constexpr bool kAlign = false;
    static constexpr int kNS = 1;
    static constexpr bool kPreRhs = true;   // pre() is b at the point: an input whose raw
                                            // fetch is b (kRawRhs) supplies it without a load
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc q) const { return ld(b, q); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab& t, int i, int j, T u, T Au, const Aux& bq, T* v) const {
        v[0] = u + smooth_corr_at<T, INT, LT>(t, i, j, bq - Au, omega, use_diag);
    }
    template <int K> __device__ __forceinline__ T* dst() const { return xo; }
};

// The same sweep, also accumulating dotv' x_new (r'z of the last pass of a V-cycle).
template <typename T> struct OutSmoothDot {
    const T* b; T* xo; T omega; const T* dotv; int use_diag;
    __device__ void at(int q, const Batch& bt) { b += q * bt.sf; xo += q * bt.sf; dotv += q * bt.sf; }
    static constexpr bool kDot = true;
    static constexpr bool kAlign = false;
    static constexpr int kNS = 1;
    struct Aux { T b, d; };
    __device__ Aux pre(const OpTab&, Loc q) const {
        Aux a;
        a.b = ld(b, q);
        a.d = dotv == b ? a.b : ld(dotv, q);   // r'z on the level-0 right-hand side r itself
        return a;
    }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab& t, int i, int j, T u, T Au, const Aux& a, T* v) const {
        v[0] = u + smooth_corr_at<T, INT, LT>(t, i, j, a.b - Au, omega, use_diag);
    }
    // kDotOfValue: the pass hands dterm the stored value x_new in place of Au
    static constexpr bool kDotOfValue = true;
    __device__ __forceinline__ T dterm(T, T xn, const Aux& a) const { return a.d * xn; }
    template <int K> __device__ __forceinline__ T* dst() const { return xo; }
};

// Fused Krylov step on the search direction: the operand is p_new = z + beta p_old
// (InAxpy); writes p_new and A p_new, applies the pending x += alpha p_old of
// the previous iteration, and accumulates p_new' A p_new.
template <typename T> struct OutPUpd {
    T* pnew; T* Ap; T* x; const T* pold; const double* alpha; T al;
    __device__ void at(int q, const Batch& bt) { pnew += q * bt.sf; Ap += q * bt.sf; x += q * bt.sf; pold += q * bt.sf; alpha += q * bt.ss; }
    static constexpr bool kDot = true;
    static constexpr bool kAlign = true;
    static constexpr int kNS = 3;
    struct Aux { T x, p; };
    __device__ void init() { al = (T)(*alpha); }
    __device__ Aux pre(const OpTab&, Loc q) const { Aux a; a.x = ld(x, q); a.p = ld(pold, q); return a; }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T u, T Au, const Aux& a, T* v) const {
        v[0] = u; v[1] = Au; v[2] = a.x + al * a.p;
    }
    __device__ __forceinline__ T dterm(T u, T Au, const Aux&) const { return u * Au; }
    template <int K> __device__ __forceinline__ T* dst() const {
        if constexpr (K == 0) return pnew;
        else if constexpr (K == 1) return Ap;
        else return x;
    }
};

// o = u (materialises a formed operand).
template <typename T> struct OutCopy {
    T* o;
    __device__ void at(int q, const Batch& bt) { o += q * bt.sf; }
    static constexpr bool kDot = false;
    static constexpr bool kAlign = false;
    static constexpr int kNS = 1;
    using Aux = T;
    __device__ Aux pre(const OpTab&, Loc) const { return T(0); }
    template <bool INT = false, bool LT = false>
    __device__ __forceinline__ void vals(const OpTab&, int, int, T u, T, const Aux&, T* v) const { v[0] = u; }
    template <int K> __device__ __forceinline__ T* dst() const { return o; }
};

template <class X, class = void> struct dot_of_value { static constexpr bool value = false; };
template <class X> struct dot_of_value<X, decltype(void(X::kDotOfValue))> { static constexpr bool value = X::kDotOfValue; };

template <int K, class OUT, typename T>
__device__ __forceinline__ void store_vals(const OUT& out, Loc q, const T* v) {
    auto* d = out.template dst<K>();
    using E = std::remove_pointer_t<decltype(d)>;
    st(d, q, (E)v[K]);
    if constexpr (K + 1 < OUT::kNS) store_vals<K + 1>(out, q, v);
}

// Stores every output at its own point and accumulates the inner product.
template <bool INT, bool LT, class OUT, typename T>
__device__ __forceinline__ void emit(const OUT& out, const OpTab& t, int i, int j, Loc q, T u, T Au,
                                     const typename OUT::Aux& a, T& acc) {
    T v[OUT::kNS];
    out.template vals<INT, LT>(t, i, j, u, Au, a, v);
    store_vals<0>(out, q, v);
    if constexpr (OUT::kDot) acc += out.dterm(u, dot_of_value<OUT>::value ? v[0] : Au, a);
}

// ---------------------------------------------------------------------------
// Block reduction helpers (wave64 shuffles, then LDS across waves)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v) {
    constexpr int NW = (NT + 63) / 64;
    __shared__ T red[NW];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NW; ++k) s += red[k];
    }
    return s;  // valid on thread 0
}

// ---------------------------------------------------------------------------
// The stencil pass
// ---------------------------------------------------------------------------
// Closure rows via direct (cache-served) loads, used only for face rows and
// columns; uc is the operand at (i, j).
// (The closure width is the order's compile-time C, so the unrolled loops
// carry no dead iterations for the narrower orders.)
// Inputs whose raw fetch is more than one value (x + P e: five, r + beta p:
// two) form each operand of the closures and of the face-row tile as it is
// loaded instead of holding every raw fetch at once: few blocks take those
// paths, but the kernel's register allocation is the largest over all its
// paths, so a face path holding C x 5 values would cap the occupancy of
// every tile.
template <class IN, typename T>
constexpr bool multi_raw() { return sizeof(typename IN::Raw) > sizeof(T); }

template <int P, typename T, bool LT = false, class IN>
__device__ __forceinline__ T k_closure(const OpTab& t, const IN& in, int i, int j, T uc) {
    constexpr int CW = Stencil<P>::C;
    const bool lo = j < t.R;
    const int r = lo ? j : t.N - 1 - j;
    T s = (T)lut<LT>(t).krs[r] * uc;
    if constexpr (multi_raw<IN, T>()) {
#pragma unroll
        for (int c = 0; c < CW; ++c)
            if (c != r) s += (T)lut<LT>(t).kb[r][c] * (in.template load<false, LT>(t, i, lo ? c : t.N - 1 - c) - uc);
        return s;
    }
    typename IN::Raw raw[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) raw[c] = in.template fetch<false, LT>(t, i, lo ? c : t.N - 1 - c);
#pragma unroll
    for (int c = 0; c < CW; ++c)
        if (c != r)
            s += (T)lut<LT>(t).kb[r][c] * (in.template make<false, LT>(t, i, lo ? c : t.N - 1 - c, raw[c]) - uc);
    return s;
}

template <int P, typename T, bool LT = false, class IN>
__device__ __forceinline__ T m_closure(const OpTab& t, const IN& in, int i, int j, T uc) {
    constexpr int CW = Stencil<P>::C;
    const bool lo = i < t.R;
    const int r = lo ? i : t.N - 1 - i;
    T s = (T)lut<LT>(t).mrs[r] * uc;
    if constexpr (multi_raw<IN, T>()) {
#pragma unroll
        for (int c = 0; c < CW; ++c)
            if (c != r) s += (T)lut<LT>(t).mb[r][c] * (in.template load<false, LT>(t, lo ? c : t.N - 1 - c, j) - uc);
        return s;
    }
    typename IN::Raw raw[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) raw[c] = in.template fetch<false, LT>(t, lo ? c : t.N - 1 - c, j);
#pragma unroll
    for (int c = 0; c < CW; ++c)
        if (c != r)
            s += (T)lut<LT>(t).mb[r][c] * (in.template make<false, LT>(t, lo ? c : t.N - 1 - c, j, raw[c]) - uc);
    return s;
}

struct PassCtl {
    const int* done;     // nullable: problem q skips the pass when done[q * ds] != 0
    int ds;              // ints between problems' done flags
    double* partial;     // nullable: problem q's block partials at partial[q * nbx * nby + tile]
    int nbx, nby;        // tile grid of one problem
    int bx0, bx1, by0, by1;   // tiles [bx0, bx1) x [by0, by1) are interior (tile_order)
    Batch bt;
};

// Order of a problem's tiles: the face tiles first, then the interior ones.
// A face tile (closure rows or columns, bounds tests, staged tables) lives
// longer than an interior one; dispatched last, the face rows at the end of
// the grid set the kernel's tail.  Dispatched first, their latency hides
// behind the interior tiles.  b: the block's index within its problem on an
// nbx x nby tile grid whose tiles [bx0, bx1) x [by0, by1) are interior; the
// interior tiles follow in row-major order, remapped (remap) so that each
// XCD walks its own band of them.
__device__ __forceinline__ void tile_order(int nbx, int nby, int bx0, int bx1, int by0, int by1, int b, bool remap,
                                           int& bx, int& by) {
    const int nix = bx1 - bx0, niy = by1 - by0;
    const int nface = nbx * nby - nix * niy;
    if (b >= nface) {
        int k = b - nface;
        if (remap) k = xcd_remap(k, nix * niy);
        by = by0 + k / nix;
        bx = bx0 + (k - (by - by0) * nix);
        return;
    }
    const int fr = (nby - niy) * nbx;   // tiles of the face rows
    if (b < fr) {
        const int rb = b / nbx;
        bx = b - rb * nbx;
        by = rb < by0 ? rb : by1 + (rb - by0);
    } else {
        const int f = b - fr, ncol = nbx - nix;
        const int r = f / ncol, col = f - r * ncol;
        by = by0 + r;
        bx = col < bx0 ? col : bx1 + (col - bx0);
    }
}

// Copies the device table into LDS (blockDim.x threads, one word each).
// The parts of the table an order-P kernel reads (the scalars, hw, the first
// R rows of mb and kb, and the per-row arrays), copied into LDS in 8-byte
// words by NT threads, so the copy does not grow with the largest order's
// closure.  fetch() only issues the loads (into registers) and store() writes
// LDS, so a kernel can overlap the table's round trip with its data loads.
template <int P, int NT>
struct TabStage {
    static_assert(sizeof(OpTab) % 8 == 0 && offsetof(OpTab, mb) % 8 == 0 && offsetof(OpTab, kb) % 8 == 0 &&
                      offsetof(OpTab, md) % 8 == 0,
                  "table is copied in 8-byte words");
    static constexpr int rows = Stencil<P>::R * kMaxC;
    static constexpr int a1 = (int)(offsetof(OpTab, mb) / 8) + rows;   // scalars, hw, mb rows < R
    static constexpr int b0 = (int)(offsetof(OpTab, kb) / 8);          // kb rows < R
    static constexpr int c0 = (int)(offsetof(OpTab, md) / 8);          // md .. dev
    static constexpr int n = a1 + rows + ((int)(sizeof(OpTab) / 8) - c0);
    static constexpr int K = (n + NT - 1) / NT;
    unsigned long long v[K];
    __device__ static int word(int k) { return k < a1 ? k : (k < a1 + rows ? b0 + (k - a1) : c0 + (k - a1 - rows)); }
    __device__ __forceinline__ void fetch(const OpTab* src) {
        const unsigned long long* s = reinterpret_cast<const unsigned long long*>(src);
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int k = (int)threadIdx.x + m * NT;
            if (k < n) v[m] = s[word(k)];
        }
    }
    __device__ __forceinline__ void store(OpTab* dst) const {
        unsigned long long* d = reinterpret_cast<unsigned long long*>(dst);
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const int k = (int)threadIdx.x + m * NT;
            if (k < n) d[word(k)] = v[m];
        }
    }
};

// Fetch, store and barrier in one (callers with nothing to overlap).
template <int P, int NT>
__device__ __forceinline__ void stage_table(const OpTab& t, OpTab* l) {
    TabStage<P, NT> st;
    st.fetch(t.dev);
    st.store(l);
    __syncthreads();
}

template <class X, class = void> struct pre_rhs { static constexpr bool value = false; };
template <class X> struct pre_rhs<X, decltype(void(X::kPreRhs))> { static constexpr bool value = X::kPreRhs; };

template <class X, class = void> struct has_init { static constexpr bool value = false; };
template <class X> struct has_init<X, decltype(void(&X::init))> { static constexpr bool value = true; };

// The pass's LDS tile: J rows of the operand on the block's TI columns, a
// right halo of H columns and a left halo of HL = C columns (the closure
// width), so that the closure rows of m on a face-column block read their C
// operand values from the tile: at the left face they are the first block's
// own columns, at the right face the last C columns, which a last block
// narrower than C reaches through that halo.  Interior blocks load H of the
// left halo.
template <int P, int TI> constexpr int tile_w() { return TI + Stencil<P>::C + Stencil<P>::H; }

// m_closure with the operand read from the tile row (columns [i0 - HL,
// i0 + TI + H)); a column outside it (the block before a last block wider
// than H, only for N not of the form 2^k + 1) from memory.  Same order of
// operations as m_closure.
template <int P, int TI, typename T, bool LT, class IN>
__device__ __forceinline__ T m_closure_tile(const OpTab& t, const IN& in, int i, int j, int i0, const T* trow, T uc) {
    constexpr int CW = Stencil<P>::C;
    constexpr int H = Stencil<P>::H;
    constexpr int HL = Stencil<P>::C;
    const bool lo = i < t.R;
    const int r = lo ? i : t.N - 1 - i;
    T s = (T)lut<LT>(t).mrs[r] * uc;
#pragma unroll
    for (int c = 0; c < CW; ++c)
        if (c != r) {
            const int col = lo ? c : t.N - 1 - c;
            const T v = col < i0 + TI + H ? trow[col - i0 + HL] : in.template load<false, LT>(t, col, j);
            s += (T)lut<LT>(t).mb[r][c] * (v - uc);
        }
    return s;
}

// ---------------------------------------------------------------------------
// Line-aligned stores (ALN > 0)
// ---------------------------------------------------------------------------
// ALN is the number of elements in a 128-byte line.  A block of TI threads
// computes the TI columns cbase .. cbase + TI - 1, cbase = bx TO - ALN with
// TO = TI - ALN, but on row j stores only the TO columns that start at
// bx TO - s_j, s_j being the offset of the destination's element (0, j)
// within its line.  TO is a multiple of ALN, so every block's stored range
// starts on a line boundary, the ranges of one row's blocks tile it exactly,
// and each line is written whole by one wave.  (With N = 2^k + 1 every row
// starts mid-line; a line written in two parts from blocks on different XCDs
// is merged by a read-modify-write in memory, which costs an unaligned pass
// about a tenth of its bandwidth.)  A value travels from the thread that
// computed it to the storing lane through the LDS tile; the stores are
// non-temporal.  ALN = 0: each thread stores its own column (cbase = bx TI).
template <int TI, int ALN> struct Cols {
    static constexpr int TO = ALN ? TI - ALN : TI;   // columns stored per row
    __host__ __device__ static int base(int bx) { return ALN ? bx * TO - ALN : bx * TI; }
    __host__ __device__ static int nbx(int N) { return ALN ? (N + ALN - 1 + TO - 1) / TO : (N + TI - 1) / TI; }
};

template <int ALN, typename E>
__device__ __forceinline__ int line_shift(const E* row) {
    return (int)(((uintptr_t)row / sizeof(E)) & (uintptr_t)(ALN - 1));
}

// Whether the calling thread's column belongs to its block's stored range on
// row j (by the first destination's lines): the fused inner products count
// each point once, at the block that stores it.
template <int TI, int ALN, class OUT>
__device__ __forceinline__ bool owns_col(const OUT& out, int j, int N) {
    if constexpr (ALN == 0) {
        return true;
    } else {
        const int s = line_shift<ALN>(out.template dst<0>() + (size_t)j * N);
        return (unsigned)((int)threadIdx.x - (ALN - s)) < (unsigned)(TI - ALN);
    }
}

// Stores destination K (and, recursively, the ones after it) of rows j0 ..
// j0 + J - 1: every thread puts its own column's value vv[r][K] in the tile,
// then the block stores each row's TO columns from cbase + ALN - s_j as
// 16-byte vectors (non-temporal; single elements at the grid's edges).
template <typename E> struct Vec16 {
    static constexpr int V = 16 / (int)sizeof(E);
    typedef E type __attribute__((ext_vector_type(16 / sizeof(E))));
};

template <int K, int TI, int J, int ALN, int NS, int TW, typename T, class OUT>
__device__ __forceinline__ void aligned_store(const OUT& out, int N, T (*tile)[TW], int cbase, int j0,
                                              const T (*vv)[NS]) {
    constexpr int TO = TI - ALN;
    const int tid = threadIdx.x;
    __syncthreads();   // the tile's previous contents have been read
#pragma unroll
    for (int r = 0; r < J; ++r) tile[r][tid] = vv[r][K];
    __syncthreads();
    auto* d = out.template dst<K>();
    using E = std::remove_pointer_t<decltype(d)>;
    constexpr int V = Vec16<E>::V;
    static_assert(TO % V == 0, "stored range is whole vectors");
    constexpr int NV = TO / V;        // vectors per row
#pragma unroll
    for (int k0 = 0; k0 < J * NV; k0 += TI) {
        const int k = k0 + tid;
        if (k < J * NV) {
            const int r = k / NV, e = k - r * NV;
            const int j = j0 + r;
            if (j < N) {
                E* row = d + (size_t)j * N;
                const int s = line_shift<ALN>(row);
                const int col = cbase + ALN - s + e * V;
                const T* src = &tile[r][ALN - s + e * V];
                if (col >= 0 && col + V <= N) {
                    typename Vec16<E>::type v;
#pragma unroll
                    for (int m = 0; m < V; ++m) v[m] = (E)src[m];
                    __builtin_nontemporal_store(v, reinterpret_cast<typename Vec16<E>::type*>(row + col));
                } else {
#pragma unroll
                    for (int m = 0; m < V; ++m)
                        if (col + m >= 0 && col + m < N) row[col + m] = (E)src[m];
                }
            }
        }
    }
    if constexpr (K + 1 < NS) aligned_store<K + 1, TI, J, ALN>(out, N, tile, cbase, j0, vv);
}

// The values of one point for the aligned path: vals() into vv, and with
// kDot the point's inner-product term when the block stores the point.
template <int TI, int ALN, bool INT, bool LT, class OUT, typename T>
__device__ __forceinline__ void point_vals(const OUT& out, const OpTab& t, int i, int j, T u, T Au,
                                           const typename OUT::Aux& a, T* vv, T& acc) {
    out.template vals<INT, LT>(t, i, j, u, Au, a, vv);
    if constexpr (OUT::kDot)
        if (owns_col<TI, ALN>(out, j, t.N)) acc += out.dterm(u, dot_of_value<OUT>::value ? vv[0] : Au, a);
}

// Tile whose rows (and their halo) lie inside the grid, away from the closure
// rows in j.  With ICHK = false the columns are interior too: straight-line
// code with no bounds tests and a constant inverse diagonal (the common
// case).  With ICHK = true (the first/last column of tiles) each thread tests
// its own column: out-of-grid threads idle, threads on the i-closure rows take
// the closure sum and the H weight; everything else stays on the fast path.
template <int P, int TI, int J, int ALN, bool ICHK, typename T, bool LT = false, class IN, class OUT>
__device__ __forceinline__ T pass_interior(const IN& in, const OUT& out, const OpTab& t, int i0, int j0,
                                           T (*tile)[tile_w<P, TI>()]) {
    constexpr int H = Stencil<P>::H;
    constexpr int W = J + 2 * H;
    constexpr int HL = Stencil<P>::C;          // centre column of thread tid: tid + HL
    constexpr int NL = ICHK ? HL : H;           // left halo columns loaded
    constexpr bool INT = !ICHK;
    const int tid = threadIdx.x;
    const int i = i0 + tid;
    const int N = t.N;
    const bool act = !ICHK || (i >= 0 && i < N);
    T w[W];
    // the smoothing right-hand side at the tile's points is already in the
    // input's raw window: no second load of it
    constexpr bool kRhsRaw = IN::kRawRhs && pre_rhs<OUT>::value;
    typename IN::Raw rhs[kRhsRaw ? J : 1];
    if (act) {
        if constexpr (IN::kCol) {
            // j0 is a multiple of the even J; the window's rows lie inside the
            // grid (the tile is off the j closure rows), and so do the coarse
            // rows and columns an active column reads
            in.template load_col<W, (H & 1)>(t, i, j0 - H, w);
        } else {
            typename IN::Raw raw[W];
#pragma unroll
            for (int k = 0; k < W; ++k) raw[k] = in.template fetch<INT, LT>(t, i, j0 - H + k);
#pragma unroll
            for (int k = 0; k < W; ++k) w[k] = in.template make<INT, LT>(t, i, j0 - H + k, raw[k]);
            if constexpr (kRhsRaw) {
#pragma unroll
                for (int r = 0; r < J; ++r) rhs[r] = raw[r + H];
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = T(0);
    }
    // NL + H halo columns x J rows, spread over the first threads.
#pragma unroll
    for (int e0 = 0; e0 < (NL + H) * J; e0 += TI) {
        const int e = e0 + tid;
        if (e < (NL + H) * J) {
            const int c = e / J, r = e % J;
            const int ii = c < NL ? i0 - NL + c : i0 + TI + (c - NL);
            T v = T(0);
            if (!ICHK || (ii >= 0 && ii < N)) v = in.template load<INT, LT>(t, ii, j0 + r);
            tile[r][c < NL ? HL - NL + c : HL + TI + (c - NL)] = v;
        }
    }
    typename OUT::Aux aux[J];
#pragma unroll
    for (int r = 0; r < J; ++r) {
        tile[r][tid + HL] = w[r + H];
        if constexpr (kRhsRaw) {
            aux[r] = rhs[r];
        } else {
            if (act) aux[r] = out.pre(t, loc(i, j0 + r, N));
        }
    }
    __syncthreads();
    T acc = T(0);
    T vv[ALN ? J : 1][OUT::kNS];
    if (act) {
        const bool iint = !ICHK || ((i >= t.R) && (i < N - t.R));
        const T hwi = ICHK ? hw_of<T, LT>(t, i) : T(1);
#pragma unroll
        for (int r = 0; r < J; ++r) {
            const T uc = w[r + H];
            T kp = Stencil<P>::template c<T>(1) * ((w[r + H - 1] - uc) + (w[r + H + 1] - uc));
#pragma unroll
            for (int s = 2; s <= H; ++s) kp += Stencil<P>::template c<T>(s) * ((w[r + H - s] - uc) + (w[r + H + s] - uc));
            T mp;
            if (iint) {
                const T* row = &tile[r][tid + HL];
                mp = Stencil<P>::template c<T>(1) * ((row[-1] - uc) + (row[1] - uc));
#pragma unroll
                for (int s = 2; s <= H; ++s) mp += Stencil<P>::template c<T>(s) * ((row[-s] - uc) + (row[s] - uc));
            } else {
                mp = m_closure_tile<P, TI, T, LT>(t, in, i, j0 + r, i0, tile[r], uc);
            }
            const T Au = ICHK ? hwi * kp + mp : kp + mp;   // hw_j = 1 on interior rows
            if constexpr (ALN == 0) emit<INT, LT>(out, t, i, j0 + r, loc(i, j0 + r, N), uc, Au, aux[r], acc);
            else point_vals<TI, ALN, INT, LT>(out, t, i, j0 + r, uc, Au, aux[r], vv[r], acc);
        }
    }
    if constexpr (ALN > 0) aligned_store<0, TI, J, ALN>(out, N, tile, i0, j0, vv);
    return acc;
}

// General tile: bounds tests, face closures and H weights.
template <int P, int TI, int J, int ALN, typename T, bool LT = false, class IN, class OUT>
__device__ __forceinline__ T pass_general(const IN& in, const OUT& out, const OpTab& t, int i0, int j0,
                                          T (*tile)[tile_w<P, TI>()]) {
    constexpr int H = Stencil<P>::H;
    constexpr int W = J + 2 * H;
    const int tid = threadIdx.x;
    const int N = t.N;
    const int i = i0 + tid;
    const bool act = i >= 0 && i < N;
    T w[W];
    if constexpr (multi_raw<IN, T>()) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int jj = j0 - H + k;
            w[k] = (act && jj >= 0 && jj < N) ? in.template load<false, LT>(t, i, jj) : T(0);
        }
    } else {
        typename IN::Raw raw[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int jj = j0 - H + k;
            if (act && jj >= 0 && jj < N) raw[k] = in.template fetch<false, LT>(t, i, jj);
        }
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int jj = j0 - H + k;
            w[k] = (act && jj >= 0 && jj < N) ? in.template make<false, LT>(t, i, jj, raw[k]) : T(0);
        }
    }
    if (tid < 2 * H) {
        const int side = tid / H, off = tid % H;
        const int ii = side == 0 ? i0 - H + off : i0 + TI + off;
        const int col = side == 0 ? off : TI + H + off;
        if constexpr (multi_raw<IN, T>()) {
#pragma unroll
            for (int r = 0; r < J; ++r) {
                const int jj = j0 + r;
                tile[r][col] = (ii >= 0 && ii < N && jj < N) ? in.template load<false, LT>(t, ii, jj) : T(0);
            }
        } else {
            typename IN::Raw raw[J];
#pragma unroll
            for (int r = 0; r < J; ++r)
                if (ii >= 0 && ii < N && j0 + r < N) raw[r] = in.template fetch<false, LT>(t, ii, j0 + r);
#pragma unroll
            for (int r = 0; r < J; ++r) {
                const int jj = j0 + r;
                tile[r][col] = (ii >= 0 && ii < N && jj < N) ? in.template make<false, LT>(t, ii, jj, raw[r]) : T(0);
            }
        }
    }
    typename OUT::Aux aux[J];
#pragma unroll
    for (int r = 0; r < J; ++r) {
        tile[r][tid + H] = w[r + H];
        const int jj = j0 + r;
        if (act && jj < N) aux[r] = out.pre(t, loc(i, jj, N));
    }
    __syncthreads();
    T acc = T(0);
    T vv[ALN ? J : 1][OUT::kNS];
    if (act) {
        const bool iint = (i >= t.R) && (i < N - t.R);
        const T hwi = hw_of<T, LT>(t, i);
#pragma unroll
        for (int r = 0; r < J; ++r) {
            const int j = j0 + r;
            if (j >= N) break;
            const bool jint = (j >= t.R) && (j < N - t.R);
            const T uc = w[r + H];
            T kp, mp;
            if (jint) {
                kp = Stencil<P>::template c<T>(1) * ((w[r + H - 1] - uc) + (w[r + H + 1] - uc));
#pragma unroll
                for (int s = 2; s <= H; ++s) kp += Stencil<P>::template c<T>(s) * ((w[r + H - s] - uc) + (w[r + H + s] - uc));
            } else {
                kp = k_closure<P, T, LT>(t, in, i, j, uc);
            }
            if (iint) {
                const T* row = &tile[r][tid + H];
                mp = Stencil<P>::template c<T>(1) * ((row[-1] - uc) + (row[1] - uc));
#pragma unroll
                for (int s = 2; s <= H; ++s) mp += Stencil<P>::template c<T>(s) * ((row[-s] - uc) + (row[s] - uc));
            } else {
                mp = m_closure<P, T, LT>(t, in, i, j, uc);
            }
            const T Au = (iint ? kp : hwi * kp) + (jint ? mp : hw_of<T, LT>(t, j) * mp);
            if constexpr (ALN == 0) emit<false, LT>(out, t, i, j, loc(i, j, N), uc, Au, aux[r], acc);
            else point_vals<TI, ALN, false, LT>(out, t, i, j, uc, Au, aux[r], vv[r], acc);
        }
    }
    if constexpr (ALN > 0) aligned_store<0, TI, J, ALN>(out, N, tile, i0, j0, vv);
    return acc;
}

// Face tiles go first (face_first); the interior tiles are remapped so that
// each XCD walks its own band of them (workgroups go to the 8 XCDs
// round-robin), where the tiles above and below a tile find the halo rows
// they share in the XCD's L2 -- except for FP32 line-aligned tiles, which
// measured faster taken in launch order (the whole chip sweeping one band of
// rows at a time).
template <int P, int TI, int J, int ALN, typename T, class IN, class OUT>
__global__ __launch_bounds__(TI) void stencil_pass(IN in, OUT out, OpTab t, PassCtl ctl) {
    constexpr int H = Stencil<P>::H;
    const int nbt = ctl.nbx * ctl.nby;
    const int q = (int)blockIdx.x / nbt;
    if (ctl.done && ctl.done[(size_t)q * ctl.ds]) return;
    int bx, by;
    tile_order(ctl.nbx, ctl.nby, ctl.bx0, ctl.bx1, ctl.by0, ctl.by1, (int)blockIdx.x - q * nbt,
               !(ALN && sizeof(T) < 8), bx, by);
    const int b = by * ctl.nbx + bx;
    const int N = t.N;
    const int i0 = Cols<TI, ALN>::base(bx), j0 = by * J;
    in.at(q, ctl.bt);
    out.at(q, ctl.bt);
    in.init();
    if constexpr (has_init<OUT>::value) out.init();
    __shared__ T tile[J][tile_w<P, TI>()];
    // every point the tile loads (halo included) lies outside the closure rows
    const bool jin = j0 - H >= t.R && j0 + J + H <= N - t.R;
    const bool iin = i0 - H >= t.R && i0 + TI + H <= N - t.R;
    T acc;
    if (jin && iin) {
        acc = pass_interior<P, TI, J, ALN, false, T>(in, out, t, i0, j0, tile);
    } else {
        // Face tiles look the closure tables up per thread: stage them in LDS
        // first (one load per thread) so those lookups cost LDS, not global,
        // latency -- on the small multigrid levels every tile is a face tile.
        __shared__ OpTab ltab;
        stage_table<P, TI>(t, &ltab);
        if (jin) acc = pass_interior<P, TI, J, ALN, true, T, true>(in, out, ltab, i0, j0, tile);
        else acc = pass_general<P, TI, J, ALN, T, true>(in, out, ltab, i0, j0, tile);
    }
    if constexpr (OUT::kDot) {
        if (ctl.partial) {
            __syncthreads();
            T s = block_sum<T, TI>(acc);
            if (threadIdx.x == 0) ctl.partial[(size_t)q * nbt + b] = (double)s;
        }
    }
}

// ---------------------------------------------------------------------------
// Residual + full-weighting restriction in one pass
// ---------------------------------------------------------------------------
// (A u)(i, j) at a single point, closure-aware, all loads through IN.
template <int P, typename T, bool LT = false, class IN>
__device__ __forceinline__ T au_point(const OpTab& t, const IN& in, int i, int j) {
    constexpr int H = Stencil<P>::H;
    const int N = t.N;
    const T uc = in.template load<false, LT>(t, i, j);
    const bool jint = (j >= t.R) && (j < N - t.R);
    const bool iint = (i >= t.R) && (i < N - t.R);
    T kp = T(0), mp = T(0);
    if (jint) {
#pragma unroll
        for (int s = 1; s <= H; ++s)
            kp += Stencil<P>::template c<T>(s) * ((in.template load<false, LT>(t, i, j - s) - uc) + (in.template load<false, LT>(t, i, j + s) - uc));
    } else {
        kp = k_closure<P, T, LT>(t, in, i, j, uc);
    }
    if (iint) {
#pragma unroll
        for (int s = 1; s <= H; ++s)
            mp += Stencil<P>::template c<T>(s) * ((in.template load<false, LT>(t, i - s, j) - uc) + (in.template load<false, LT>(t, i + s, j) - uc));
    } else {
        mp = m_closure<P, T, LT>(t, in, i, j, uc);
    }
    return (iint ? kp : hw_of<T, LT>(t, i) * kp) + (jint ? mp : hw_of<T, LT>(t, j) * mp);
}

struct RestrictCtl {
    const int* done;   // nullable, per problem (stride ds ints)
    int ds;
    int Nc;
    int nbx, nby;      // block grid over coarse points: TI/2-1 columns x J/2 rows per block
    int bx0, bx1, by0, by1;   // blocks [bx0, bx1) x [by0, by1) are interior (tile_order)
    Batch bt;          // sf: fine elements, sc: coarse elements per problem
};

// rc = scale * FW(b - A u) on the coarse grid of side Nc = (N+1)/2 (b null:
// rc = scale * FW(-A u)).  A block
// owns TI/2-1 coarse columns and J/2 coarse rows; its TI threads form the fine
// residual on the TI x (J+1) fine points those coarse points weigh (one
// column/row overlap with the neighbouring blocks), keep it in LDS and
// restrict it there, so the fine residual never goes to HBM.
template <int P, int TI, int J, typename T>
__global__ __launch_bounds__(TI) void resid_restrict(const T* __restrict__ u, const T* __restrict__ b,
                                                    T* __restrict__ rc, T scale, OpTab t, RestrictCtl ctl) {
    constexpr int H = Stencil<P>::H;
    constexpr int JR = J + 1;            // fine residual rows per block
    constexpr int W = JR + 2 * H;        // column window
    constexpr int CI = TI / 2 - 1;       // coarse columns per block
    constexpr int CJ = J / 2;            // coarse rows per block
    const int nb = ctl.nbx * ctl.nby;
    const int q = (int)blockIdx.x / nb;
    if (ctl.done && ctl.done[(size_t)q * ctl.ds]) return;
    u += q * ctl.bt.sf;
    if (b) b += q * ctl.bt.sf;
    rc += q * ctl.bt.sc;
    int bx, by;
    tile_order(ctl.nbx, ctl.nby, ctl.bx0, ctl.bx1, ctl.by0, ctl.by1, (int)blockIdx.x - q * nb, true, bx, by);
    const int N = t.N, Nc = ctl.Nc;
    const int Ic0 = bx * CI, Jc0 = by * CJ;
    const int tid = threadIdx.x;
    const int fi0 = 2 * Ic0 - 1, fj0 = 2 * Jc0 - 1;   // fine origin (may be -1)
    const int fi = fi0 + tid;
    const InPlain<T> in{u};
    __shared__ T ut[JR][TI + 2 * H];
    T (*rt)[TI + 2 * H] = ut;           // the residual reuses the operand tile after a barrier
    T res[JR];
    const bool jin = fj0 - H >= t.R && fj0 + JR + H <= N - t.R;
    const bool iin = fi0 - H >= t.R && fi0 + TI + H <= N - t.R;
    __shared__ OpTab ltab;
    if (!(jin && iin)) stage_table<P, TI>(t, &ltab);   // per-thread closure lookups from LDS
    const OpTab& tl = ltab;
    if (jin) {
        // Rows interior: register window along j, LDS tile along i; with
        // columns at the faces (iin false) each thread tests its own column.
        const bool act = iin || (fi >= 0 && fi < N);
        T w[W];
#pragma unroll
        for (int k = 0; k < W; ++k) w[k] = act ? ld(u, loc(fi, fj0 - H + k, N)) : T(0);
#pragma unroll
        for (int e0 = 0; e0 < 2 * H * JR; e0 += TI) {
            const int e = e0 + tid;
            if (e < 2 * H * JR) {
                const int side = e / (H * JR), rem = e % (H * JR);
                const int off = rem / JR, r = rem % JR;
                const int ii = side == 0 ? fi0 - H + off : fi0 + TI + off;
                ut[r][side == 0 ? off : TI + H + off] = (iin || (ii >= 0 && ii < N)) ? *elem(u, loc(ii, fj0 + r, N)) : T(0);
            }
        }
        T bq[JR];
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            ut[r][tid + H] = w[r + H];
            bq[r] = (act && b) ? ld(b, loc(fi, fj0 + r, N)) : T(0);
        }
        __syncthreads();
        const bool iint = iin || ((fi >= t.R) && (fi < N - t.R));
        const T hwi = iin ? T(1) : (act ? hw_of<T, true>(tl, fi) : T(0));
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            const T uc = w[r + H];
            T kp = Stencil<P>::template c<T>(1) * ((w[r + H - 1] - uc) + (w[r + H + 1] - uc));
#pragma unroll
            for (int s = 2; s <= H; ++s) kp += Stencil<P>::template c<T>(s) * ((w[r + H - s] - uc) + (w[r + H + s] - uc));
            T mp;
            if (iint) {
                const T* row = &ut[r][tid + H];
                mp = Stencil<P>::template c<T>(1) * ((row[-1] - uc) + (row[1] - uc));
#pragma unroll
                for (int s = 2; s <= H; ++s) mp += Stencil<P>::template c<T>(s) * ((row[-s] - uc) + (row[s] - uc));
            } else {
                mp = act ? m_closure<P, T, true>(tl, in, fi, fj0 + r, uc) : T(0);
            }
            res[r] = act ? bq[r] - (hwi * kp + mp) : T(0);
        }
    } else {
#pragma unroll
        for (int r = 0; r < JR; ++r) {
            const int fj = fj0 + r;
            T v = T(0);
            if (fi >= 0 && fi < N && fj >= 0 && fj < N)
                v = (b ? ld(b, loc(fi, fj, N)) : T(0)) - au_point<P, T, true>(tl, in, fi, fj);
            res[r] = v;
        }
    }
    __syncthreads();                     // every thread is done reading the operand tile
#pragma unroll
    for (int r = 0; r < JR; ++r) rt[r][tid] = res[r];
    __syncthreads();
    // Full weighting [1 2 1; 2 4 2; 1 2 1] / 16 (times scale) around fine (2I, 2Jc).
    for (int o = tid; o < CI * CJ; o += TI) {
        const int cI = o % CI, cJ = o / CI;
        const int I = Ic0 + cI, Jc = Jc0 + cJ;
        if (I >= Nc || Jc >= Nc) continue;
        const int lc = 2 * cI + 1, lr = 2 * cJ + 1;
        const T s0 = rt[lr][lc] + T(0.5) * (rt[lr][lc - 1] + rt[lr][lc + 1]);
        const T sm = rt[lr - 1][lc] + T(0.5) * (rt[lr - 1][lc - 1] + rt[lr - 1][lc + 1]);
        const T sp = rt[lr + 1][lc] + T(0.5) * (rt[lr + 1][lc - 1] + rt[lr + 1][lc + 1]);
        *elem(rc, loc(I, Jc, Nc)) = scale * T(0.25) * (s0 + T(0.5) * (sm + sp));
    }
}

}  // namespace sbp
