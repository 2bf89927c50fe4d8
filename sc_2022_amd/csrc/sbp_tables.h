// sbp_tables.h -- host-side construction of the dimensionless 1-D factors.
//
// From the reference's diagonal-norm SBP second-derivative tables
// (diagonal_sbp.jl:191-208: bhinv, interior d, closure bd, boundary row BS)
// this builds, for a side of N points,
//     hw = H / h                        (H = diag(h / bhinv) near the faces, h inside)
//     m  = h (BS - H D2)                (the 1-D "stiffness", symmetric PSD, m 1 = 0)
//     k  = m - (E bs + bs^T E) + 13 E   (m plus the Dirichlet SAT with tau = -13/h, beta = 1)
// so that the reference's A = -H~(D2 + SAT) (level_2_multigrid_new.jl:141-185)
// equals  k (x) hw  +  hw (x) m  (first factor on the slow index j).
// The Neumann SAT with tau_N = -1 cancels the BS term of the fast index exactly.
// Only the first R rows / C columns differ from the interior stencil; they
// are extracted from a small dense build and checked for shape.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "sbp_kernels.h"

namespace sbp {

struct Sbp1D {
    int p = 0, bm = 0, bn = 0, nbs = 0;
    double bhinv[4] = {0, 0, 0, 0};
    double d[5] = {0, 0, 0, 0, 0};       // interior D2 stencil (times h^2)
    double bd[4][6] = {};                 // closure of D2 (times h^2)
    double bs[4] = {0, 0, 0, 0};          // boundary derivative row (times h)
};

inline bool sbp_1d(int p, Sbp1D* s) {
    std::memset(s, 0, sizeof(*s));
    s->p = p;
    if (p == 2) {
        s->bm = 1; s->bn = 3; s->nbs = 3;
        s->bhinv[0] = 2.0;
        const double d[3] = {1.0, -2.0, 1.0};
        for (int k = 0; k < 3; ++k) { s->d[k] = d[k]; s->bd[0][k] = d[k]; }
        s->bs[0] = 3.0 / 2.0; s->bs[1] = -2.0; s->bs[2] = 1.0 / 2.0;
        return true;
    }
    if (p == 4) {
        s->bm = 4; s->bn = 6; s->nbs = 4;
        const double bh[4] = {48.0 / 17.0, 48.0 / 59.0, 48.0 / 43.0, 48.0 / 49.0};
        const double d[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
        const double bd[4][6] = {
            {2.0, -5.0, 4.0, -1.0, 0.0, 0.0},
            {1.0, -2.0, 1.0, 0.0, 0.0, 0.0},
            {-4.0 / 43.0, 59.0 / 43.0, -110.0 / 43.0, 59.0 / 43.0, -4.0 / 43.0, 0.0},
            {-1.0 / 49.0, 0.0, 59.0 / 49.0, -118.0 / 49.0, 64.0 / 49.0, -4.0 / 49.0}};
        for (int k = 0; k < 4; ++k) s->bhinv[k] = bh[k];
        for (int k = 0; k < 5; ++k) s->d[k] = d[k];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 6; ++c) s->bd[r][c] = bd[r][c];
        s->bs[0] = 11.0 / 6.0; s->bs[1] = -3.0; s->bs[2] = 3.0 / 2.0; s->bs[3] = -1.0 / 3.0;
        return true;
    }
    return false;
}

// Dense dimensionless factors (hw, m, k) for side n (n >= 2*bm, n >= bn).
inline void dense_factors(const Sbp1D& s, int n, std::vector<double>& hw, std::vector<double>& m,
                          std::vector<double>& k) {
    std::vector<double> D2(n * n, 0.0), BS(n * n, 0.0);
    const int half = s.p / 2;
    for (int r = s.bm; r < n - s.bm; ++r)
        for (int t = 0; t <= s.p; ++t) D2[r * n + (r - half + t)] = s.d[t];
    for (int r = 0; r < s.bm; ++r)
        for (int c = 0; c < s.bn; ++c) {
            D2[r * n + c] = s.bd[r][c];
            D2[(n - 1 - r) * n + (n - 1 - c)] = s.bd[r][c];
        }
    for (int c = 0; c < s.nbs; ++c) {
        BS[0 * n + c] = s.bs[c];                  // -S0 row 1
        BS[(n - 1) * n + (n - 1 - c)] = s.bs[c];  //  SN row N
    }
    hw.assign(n, 1.0);
    for (int r = 0; r < s.bm; ++r) {
        hw[r] = 1.0 / s.bhinv[r];
        hw[n - 1 - r] = 1.0 / s.bhinv[r];
    }
    m.assign(n * n, 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) m[r * n + c] = BS[r * n + c] - hw[r] * D2[r * n + c];
    k = m;
    for (int c = 0; c < n; ++c) {
        k[0 * n + c] -= BS[0 * n + c];
        k[(n - 1) * n + c] -= BS[(n - 1) * n + c];
        k[c * n + 0] -= BS[0 * n + c];
        k[c * n + (n - 1)] -= BS[(n - 1) * n + c];
    }
    k[0] += 13.0;
    k[(n - 1) * n + (n - 1)] += 13.0;
}

// Fills the kernel table for side N. Returns false when N is too small.
inline bool build_optab(int p, int N, OpTab* t, RhsTab* rt) {
    Sbp1D s;
    if (!sbp_1d(p, &s)) return false;
    const int R = p == 2 ? 3 : 4;
    const int C = p == 2 ? 4 : 6;
    if (N < 2 * C || N < 2 * s.bm || N < s.bn) return false;
    const int n0 = 2 * C + 4;
    std::vector<double> hw, m, k;
    dense_factors(s, n0, hw, m, k);
    std::memset(t, 0, sizeof(*t));
    t->N = N; t->R = R; t->C = C; t->RH = s.bm;
    for (int r = 0; r < s.bm; ++r) t->hw[r] = hw[r];
    for (int r = 0; r < R; ++r) {
        for (int c = 0; c < C; ++c) {
            t->mb[r][c] = m[r * n0 + c];
            t->kb[r][c] = k[r * n0 + c];
        }
        for (int c = C; c < n0; ++c)
            if (m[r * n0 + c] != 0.0 || k[r * n0 + c] != 0.0) return false;
        t->md[r] = m[r * n0 + r];
        t->kd[r] = k[r * n0 + r];
        double ms = 0.0, ks = 0.0, mmax = 0.0;
        for (int c = 0; c < C; ++c) {
            ms += t->mb[r][c];
            ks += t->kb[r][c];
            mmax = std::fmax(mmax, std::fabs(t->mb[r][c]));
        }
        if (std::fabs(ms) > 1e-12 * mmax) return false;   // m 1 = 0 must hold
        t->mrs[r] = 0.0;
        t->krs[r] = ks;
    }
    t->mid = -s.d[p / 2];
    t->idint = 1.0 / (2.0 * t->mid);
    if (rt) {
        std::memset(rt, 0, sizeof(*rt));
        rt->nbs = s.nbs;
        for (int c = 0; c < s.nbs; ++c) rt->bs[c] = s.bs[c];
        rt->h = 1.0 / (N - 1);
    }
    return true;
}

// Dense assembled A (row-major, n = N*N) for a small side; used for the
// coarsest-level factorisation.
inline void dense_operator(const OpTab& t, std::vector<double>& A) {
    const int N = t.N, n = N * N;
    A.assign((size_t)n * n, 0.0);
    auto hwf = [&](int r) {
        if (r < t.RH) return t.hw[r];
        if (r >= N - t.RH) return t.hw[N - 1 - r];
        return 1.0;
    };
    const int H = t.C == 6 ? 2 : 1;
    auto coef = [&](bool is_k, int r, int c) -> double {
        if (r < t.R) return (c < t.C) ? (is_k ? t.kb[r][c] : t.mb[r][c]) : 0.0;
        if (r >= N - t.R) {
            const int rr = N - 1 - r, cc = N - 1 - c;
            return (cc >= 0 && cc < t.C) ? (is_k ? t.kb[rr][cc] : t.mb[rr][cc]) : 0.0;
        }
        const int s = c - r;
        if (s < -H || s > H) return 0.0;
        if (s == 0) return t.mid;
        if (H == 1) return -1.0;
        return (s == 1 || s == -1) ? -4.0 / 3.0 : 1.0 / 12.0;
    };
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const int row = i + j * N;
            for (int c = 0; c < N; ++c) {
                const double kv = coef(true, j, c);
                if (kv != 0.0) A[(size_t)row * n + (i + c * N)] += hwf(i) * kv;
                const double mv = coef(false, i, c);
                if (mv != 0.0) A[(size_t)row * n + (c + j * N)] += hwf(j) * mv;
            }
        }
}

// In-place inverse of an SPD matrix via Cholesky (row-major n x n).
inline bool spd_inverse(std::vector<double>& A, int n) {
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / ljj;
        }
    }
    // Solve L L^T X = I column by column.
    std::vector<double> X((size_t)n * n, 0.0), y(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double v = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) v -= L[(size_t)i * n + k] * y[k];
            y[i] = v / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double v = y[i];
            for (int k = i + 1; k < n; ++k) v -= L[(size_t)k * n + i] * X[(size_t)k * n + c];
            X[(size_t)i * n + c] = v / L[(size_t)i * n + i];
        }
    }
    // Symmetrise (the inverse of a symmetric matrix).
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (X[(size_t)i * n + j] + X[(size_t)j * n + i]);
            X[(size_t)i * n + j] = X[(size_t)j * n + i] = v;
        }
    A.swap(X);
    return true;
}

}  // namespace sbp
