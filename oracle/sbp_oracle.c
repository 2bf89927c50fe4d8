/*
 * sbp_oracle.c -- CPU oracle in C for the SBP-SAT 2-D Poisson path
 * (test infrastructure and CPU baseline only; never linked into the product).
 *
 * Restates, with plain loops (OpenMP across grid lines):
 *   - the 1-D SBP second-derivative operator and norm of diagonal_sbp.jl:189-454
 *     (p = 2 :191-197, p = 4 :198-208; H :421-428; S0/SN :446-448), built row
 *     by row for any side N;
 *   - the SPD operator A = -H~ (D2 + SAT_W + SAT_E + SAT_S + SAT_N) of
 *     preconditioned_CG/level_2_multigrid_new.jl:141-185, both matrix-free and as
 *     CSR (Operators_2d / Assembling_matrix :48-188);
 *   - the right-hand side b of :164-167,181,185 for given data (f, g);
 *   - CG_CPU (level_2_multigrid_new.jl:258-291) on the CSR matrix;
 *   - MG-preconditioned CG (mg_preconditioned_CG :480-524) with the V-cycle
 *     of damped-Jacobi sweeps (jacobi_brittany! :335-344), full weighting
 *     times a scale (restriction_matrix_normal :220-233), bilinear
 *     prolongation (prolongation_matrix :190-203) and an exact coarse solve
 *     (dense Cholesky), as the GPU library builds it;
 *   - the transpose of the right-hand-side map (the adjoint of :164-167,181,185
 *     with respect to f and the four boundary data vectors) and, from it, the
 *     FP64 gradient of L = 1/2 (u - u_obs)' H~ (u - u_obs) with u = A^-1 b(f, g)
 *     (adjoint MG-PCG with the same V-cycle; A is symmetric);
 *   - the seeded Fourier-mode data fields (sum of cosine modes; the mode list
 *     is the caller's) by direct evaluation, and the H~ norm.
 *
 * Pinned in tests/test_c_oracle.py against the numpy/scipy oracle
 * (oracle/sbp_oracle.py), itself pinned in tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- 1-D operators --------------------------------------------------- */
typedef struct {
    int N, p;
    double *hw;      /* H / h                         (N)      */
    int *mc0, *kc0;  /* first column of row r's band  (N)      */
    int *mw, *kw;    /* band width of row r           (N)      */
    double *mv, *kv; /* band values, 8 slots per row  (8 N)    */
    double *mrs, *krs; /* exact row sums of m and k (N): 0 except the Dirichlet SAT rows of k */
} op1d;

static int tables(int p, int *bm, int *bn, int *nbs, double *bhinv, double *d, double bd[4][6], double *bs) {
    memset(bd, 0, sizeof(double) * 24);
    if (p == 2) {
        *bm = 1; *bn = 3; *nbs = 3;
        bhinv[0] = 2.0;
        d[0] = 1.0; d[1] = -2.0; d[2] = 1.0;
        bd[0][0] = 1.0; bd[0][1] = -2.0; bd[0][2] = 1.0;
        bs[0] = 1.5; bs[1] = -2.0; bs[2] = 0.5;
        return 0;
    }
    if (p == 4) {
        static const double BD[4][6] = {
            {2.0, -5.0, 4.0, -1.0, 0.0, 0.0},
            {1.0, -2.0, 1.0, 0.0, 0.0, 0.0},
            {-4.0 / 43.0, 59.0 / 43.0, -110.0 / 43.0, 59.0 / 43.0, -4.0 / 43.0, 0.0},
            {-1.0 / 49.0, 0.0, 59.0 / 49.0, -118.0 / 49.0, 64.0 / 49.0, -4.0 / 49.0}};
        *bm = 4; *bn = 6; *nbs = 4;
        bhinv[0] = 48.0 / 17.0; bhinv[1] = 48.0 / 59.0; bhinv[2] = 48.0 / 43.0; bhinv[3] = 48.0 / 49.0;
        d[0] = -1.0 / 12.0; d[1] = 4.0 / 3.0; d[2] = -5.0 / 2.0; d[3] = 4.0 / 3.0; d[4] = -1.0 / 12.0;
        memcpy(bd, BD, sizeof(BD));
        bs[0] = 11.0 / 6.0; bs[1] = -3.0; bs[2] = 1.5; bs[3] = -1.0 / 3.0;
        return 0;
    }
    return -1;
}

static void op1d_free(op1d *o) {
    free(o->hw); free(o->mc0); free(o->kc0); free(o->mw); free(o->kw); free(o->mv); free(o->kv);
    free(o->mrs); free(o->krs);
}

/* Dense row r of the dimensionless D2 (times h^2) and BS (times h), placed in
 * a window starting at column c0 of width 8. */
static int op1d_build(int p, int N, op1d *o) {
    int bm, bn, nbs;
    double bhinv[4], d[5], bd[4][6], bs[4];
    if (tables(p, &bm, &bn, &nbs, bhinv, d, bd, bs)) return -1;
    if (N < 2 * bm || N < bn || N < 12) return -2;
    o->N = N; o->p = p;
    o->hw = malloc(sizeof(double) * N);
    o->mc0 = malloc(sizeof(int) * N); o->kc0 = malloc(sizeof(int) * N);
    o->mw = malloc(sizeof(int) * N); o->kw = malloc(sizeof(int) * N);
    o->mv = calloc(8 * (size_t)N, sizeof(double)); o->kv = calloc(8 * (size_t)N, sizeof(double));
    o->mrs = calloc(N, sizeof(double)); o->krs = calloc(N, sizeof(double));
    for (int r = 0; r < N; ++r) o->hw[r] = 1.0;
    for (int r = 0; r < bm; ++r) { o->hw[r] = 1.0 / bhinv[r]; o->hw[N - 1 - r] = 1.0 / bhinv[r]; }
    const int half = p / 2;
    for (int r = 0; r < N; ++r) {
        double win[16]; memset(win, 0, sizeof(win));
        int c0;
        /* window start: cover closure columns or the centred stencil */
        if (r < 6) c0 = 0;
        else if (r >= N - 6) c0 = N - 8;
        else c0 = r - half;
        /* -hw[r] * D2 row */
        if (r < bm) {
            for (int c = 0; c < bn; ++c) win[c - c0] -= o->hw[r] * bd[r][c];
        } else if (r >= N - bm) {
            const int rr = N - 1 - r;
            for (int c = 0; c < bn; ++c) win[(N - 1 - c) - c0] -= o->hw[r] * bd[rr][c];
        } else {
            for (int t = 0; t <= p; ++t) win[(r - half + t) - c0] -= d[t];
        }
        /* + BS row (rows 0 and N-1 only): BS = SN - S0 */
        if (r == 0) for (int c = 0; c < nbs; ++c) win[c - c0] += bs[c];
        if (r == N - 1) for (int c = 0; c < nbs; ++c) win[(N - 1 - c) - c0] += bs[c];
        double kwin[16]; memcpy(kwin, win, sizeof(win));
        /* k = m - (E BS + BS^T E) + 13 E  (Dirichlet SAT, tau = -13/h, beta = 1) */
        if (r == 0) { for (int c = 0; c < nbs; ++c) kwin[c - c0] -= bs[c]; kwin[0 - c0] += 13.0; }
        if (r == N - 1) { for (int c = 0; c < nbs; ++c) kwin[(N - 1 - c) - c0] -= bs[c]; kwin[(N - 1) - c0] += 13.0; }
        if (r < nbs) kwin[0 - c0] -= bs[r];
        if (N - 1 - r < nbs) kwin[(N - 1) - c0] -= bs[N - 1 - r];
        /* Exact row sums (from the rational tables, not from the rounded
         * entries): D2 1 = 0 and BS 1 = 0 give m 1 = 0; k adds the Dirichlet
         * SAT, whose rows sum to 13 delta_r0 - bs[r] (r < nbs, mirrored). */
        o->mrs[r] = 0.0;
        o->krs[r] = 0.0;
        if (r == 0 || r == N - 1) o->krs[r] += 13.0;
        if (r < nbs) o->krs[r] -= bs[r];
        if (N - 1 - r < nbs) o->krs[r] -= bs[N - 1 - r];
        o->mc0[r] = o->kc0[r] = c0;
        o->mw[r] = o->kw[r] = 8;
        for (int c = 0; c < 8; ++c) { o->mv[8 * (size_t)r + c] = win[c]; o->kv[8 * (size_t)r + c] = kwin[c]; }
    }
    return 0;
}

static double dot_(const double *a, const double *b, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t q = 0; q < n; ++q) s += a[q] * b[q];
    return s;
}

/* out = A u (matrix-free, tensor form).  Each 1-D row is evaluated as
 *     rowsum_r u_r + sum_{c != r} a_rc (u_c - u_r)
 * with the exact row sum of the rational operator.  This is the same
 * operator, but it keeps the zero row sums of diagonal_sbp.jl's D2 and BS
 * exactly: the rounded FP64 entries themselves do not sum to zero (p = 4
 * interior: 1/12 - 4/3 + 5/2 - 4/3 + 1/12 = 1.4e-16 in FP64), and a plain
 * sum of them acts as a diagonal shift of that size, which moves the
 * solution of A x = b by about 1.4e-16 / lambda_min(A) -- 2e-10..4e-10
 * relative at 4097^2.  apply_plain_ keeps the plain sum for the test that
 * shows the shift. */
static void apply_(const op1d *o, const double *u, double *out) {
    const int N = o->N;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        const double *kv = o->kv + 8 * (size_t)j;
        const int kc0 = o->kc0[j];
        for (int i = 0; i < N; ++i) {
            const double *mv = o->mv + 8 * (size_t)i;
            const int mc0 = o->mc0[i];
            const double uc = u[i + (size_t)j * N];
            double ks = o->krs[j] * uc, ms = o->mrs[i] * uc;
            for (int c = 0; c < 8; ++c) {
                if (kv[c] != 0.0 && kc0 + c != j) ks += kv[c] * (u[i + (size_t)(kc0 + c) * N] - uc);
                if (mv[c] != 0.0 && mc0 + c != i) ms += mv[c] * (u[(mc0 + c) + (size_t)j * N] - uc);
            }
            out[i + (size_t)j * N] = o->hw[i] * ks + o->hw[j] * ms;
        }
    }
}

static void apply_plain_(const op1d *o, const double *u, double *out) {
    const int N = o->N;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        const double *kv = o->kv + 8 * (size_t)j;
        const int kc0 = o->kc0[j];
        for (int i = 0; i < N; ++i) {
            const double *mv = o->mv + 8 * (size_t)i;
            const int mc0 = o->mc0[i];
            double ks = 0.0, ms = 0.0;
            for (int c = 0; c < 8; ++c) {
                if (kv[c] != 0.0) ks += kv[c] * u[i + (size_t)(kc0 + c) * N];
                if (mv[c] != 0.0) ms += mv[c] * u[(mc0 + c) + (size_t)j * N];
            }
            out[i + (size_t)j * N] = o->hw[i] * ks + o->hw[j] * ms;
        }
    }
}

static double diag_(const op1d *o, int i, int j) {
    return o->hw[i] * o->kv[8 * (size_t)j + (j - o->kc0[j])] + o->hw[j] * o->mv[8 * (size_t)i + (i - o->mc0[i])];
}

/* ---- public: operator ----------------------------------------------------- */
static int g_threads = 0;

void orc_set_threads(int t) {
    g_threads = t;
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#endif
}

int orc_apply(int p, int N, const double *u, double *out) {
    op1d o;
    int rc = op1d_build(p, N, &o);
    if (rc) return rc;
    apply_(&o, u, out);
    op1d_free(&o);
    return 0;
}

/* The plain-sum evaluation of the same entries (see apply_). */
int orc_apply_plain(int p, int N, const double *u, double *out) {
    op1d o;
    int rc = op1d_build(p, N, &o);
    if (rc) return rc;
    apply_plain_(&o, u, out);
    op1d_free(&o);
    return 0;
}

int orc_diag(int p, int N, double *d) {
    op1d o;
    int rc = op1d_build(p, N, &o);
    if (rc) return rc;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) d[i + (size_t)j * N] = diag_(&o, i, j);
    op1d_free(&o);
    return 0;
}

/* CSR of A with sorted columns; returns nnz (call with rp == NULL to count). */
long orc_csr(int p, int N, int32_t *rp, int32_t *col, double *val) {
    op1d o;
    if (op1d_build(p, N, &o)) return -1;
    long nnz = 0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            /* gather (column, value) pairs, columns ascending */
            long cs[24]; double vs[24]; int n = 0;
            const double *kv = o.kv + 8 * (size_t)j, *mv = o.mv + 8 * (size_t)i;
            for (int c = 0; c < 8; ++c) {
                const int jj = o.kc0[j] + c;
                if (kv[c] != 0.0 && jj < j) { cs[n] = i + (long)jj * N; vs[n++] = o.hw[i] * kv[c]; }
            }
            for (int c = 0; c < 8; ++c) {
                const int ii = o.mc0[i] + c;
                double v = (mv[c] != 0.0) ? o.hw[j] * mv[c] : 0.0;
                if (ii == i) v += o.hw[i] * kv[j - o.kc0[j]];
                if (v != 0.0) { cs[n] = ii + (long)j * N; vs[n++] = v; }
            }
            for (int c = 0; c < 8; ++c) {
                const int jj = o.kc0[j] + c;
                if (kv[c] != 0.0 && jj > j) { cs[n] = i + (long)jj * N; vs[n++] = o.hw[i] * kv[c]; }
            }
            if (rp) {
                const long row = i + (long)j * N;
                rp[row] = (int32_t)nnz;
                for (int a = 0; a < n; ++a) { col[nnz + a] = (int32_t)cs[a]; val[nnz + a] = vs[a]; }
            }
            nnz += n;
        }
    if (rp) rp[(size_t)N * N] = (int32_t)nnz;
    op1d_free(&o);
    return nnz;
}

/* b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN) */
int orc_rhs(int p, int N, const double *f, const double *gW, const double *gE, const double *gS, const double *gN,
            double *b) {
    int bm, bn, nbs;
    double bhinv[4], d[5], bd[4][6], bs[4];
    if (tables(p, &bm, &bn, &nbs, bhinv, d, bd, bs)) return -1;
    op1d o;
    if (op1d_build(p, N, &o)) return -2;
    const double h = 1.0 / (N - 1), tauD = -13.0 / h, tauN = -1.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            /* t = f + SAT_r g (the reference's un-normed right side) */
            double t = f ? f[i + (size_t)j * N] : 0.0;
            /* Dirichlet W: HI_x (tau E_W + beta BS^T E_W) e_W gW, HI = 1/(h hw) */
            if (gW && j < nbs) t += (1.0 / (h * o.hw[j])) * ((j == 0 ? tauD : 0.0) + bs[j] / h) * gW[i];
            if (gE && N - 1 - j < nbs) {
                const int k = N - 1 - j;
                t += (1.0 / (h * o.hw[j])) * ((k == 0 ? tauD : 0.0) + bs[k] / h) * gE[i];
            }
            /* Neumann S/N: HI_y tau E e g */
            if (gS && i == 0) t += (1.0 / (h * o.hw[0])) * tauN * gS[j];
            if (gN && i == N - 1) t += (1.0 / (h * o.hw[N - 1])) * tauN * gN[j];
            b[i + (size_t)j * N] = -(h * o.hw[i]) * (h * o.hw[j]) * t;
        }
    op1d_free(&o);
    return 0;
}

/* ---- CG_CPU on CSR (level_2_multigrid_new.jl:258-291) ------------------- */
static void spmv_(long n, const int32_t *rp, const int32_t *col, const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (long r = 0; r < n; ++r) {
        double s = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

int orc_cg_csr(long n, const int32_t *rp, const int32_t *col, const double *val, const double *b, double *x,
               double abstol, int maxiter, double *resnorm) {
    double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
    spmv_(n, rp, col, val, x, Ap);
    for (long q = 0; q < n; ++q) { r[q] = b[q] - Ap[q]; p[q] = r[q]; }
    double rsold = dot_(r, r, n);
    int it = 0;
    double rn = sqrt(rsold);
    for (int s = 0; s < maxiter; ++s) {
        ++it;
        spmv_(n, rp, col, val, p, Ap);
        const double alpha = rsold / dot_(p, Ap, n);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) { x[q] += alpha * p[q]; r[q] -= alpha * Ap[q]; }
        const double rsnew = dot_(r, r, n);
        rn = sqrt(rsnew);
        if (rn < abstol) break;
        const double beta = rsnew / rsold;
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) p[q] = r[q] + beta * p[q];
        rsold = rsnew;
    }
    if (resnorm) *resnorm = rn;
    free(r); free(p); free(Ap);
    return it;
}

/* ---- multigrid ------------------------------------------------------------ */
typedef struct {
    int nlev;
    op1d op[16];
    double *x[16], *y[16], *b[16], *r[16];
    double *Cinv;   /* coarsest inverse (dense) */
    int nc;
} mg_t;

static void mg_free(mg_t *g) {
    for (int l = 0; l < g->nlev; ++l) {
        op1d_free(&g->op[l]);
        free(g->x[l]); free(g->y[l]); free(g->b[l]); free(g->r[l]);
    }
    free(g->Cinv);
}

static int chol_inverse(double *A, int n) {
    double *L = calloc((size_t)n * n, sizeof(double));
    for (int j = 0; j < n; ++j) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0)) { free(L); return -1; }
        L[(size_t)j * n + j] = sqrt(s);
        for (int i = j + 1; i < n; ++i) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / L[(size_t)j * n + j];
        }
    }
    double *y = malloc(sizeof(double) * n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double v = (i == c);
            for (int k = 0; k < i; ++k) v -= L[(size_t)i * n + k] * y[k];
            y[i] = v / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double v = y[i];
            for (int k = i + 1; k < n; ++k) v -= L[(size_t)k * n + i] * A[(size_t)k * n + c];
            A[(size_t)i * n + c] = v / L[(size_t)i * n + i];
        }
    }
    free(y); free(L);
    return 0;
}

static int mg_build(int p, int N, int coarse_n, mg_t *g) {
    memset(g, 0, sizeof(*g));
    int n = N, l = 0;
    for (;;) {
        if (op1d_build(p, n, &g->op[l])) return -1;
        size_t sz = (size_t)n * n;
        g->x[l] = calloc(sz, sizeof(double)); g->y[l] = calloc(sz, sizeof(double));
        g->b[l] = calloc(sz, sizeof(double)); g->r[l] = calloc(sz, sizeof(double));
        ++l;
        if (n == coarse_n) break;
        if (n < coarse_n || ((n - 1) & 1)) return -2;
        n = (n + 1) / 2;
    }
    g->nlev = l;
    /* dense coarsest operator from unit vectors, then its inverse */
    const int nc = n * n;
    g->nc = nc;
    g->Cinv = calloc((size_t)nc * nc, sizeof(double));
    double *e = calloc(nc, sizeof(double)), *col = malloc(sizeof(double) * nc);
    for (int c = 0; c < nc; ++c) {
        e[c] = 1.0;
        apply_(&g->op[l - 1], e, col);
        for (int r = 0; r < nc; ++r) g->Cinv[(size_t)r * nc + c] = col[r];
        e[c] = 0.0;
    }
    free(e); free(col);
    return chol_inverse(g->Cinv, nc);
}

/* x <- x + omega D^-1 (b - A x), nu times (ping-pong through y) */
static void jacobi_(const op1d *o, double *x, double *y, const double *b, int nu, double omega) {
    const int N = o->N;
    const size_t n = (size_t)N * N;
    for (int s = 0; s < nu; ++s) {
        apply_(o, x, y);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const size_t q = i + (size_t)j * N;
                y[q] = x[q] + omega * (b[q] - y[q]) / diag_(o, i, j);
            }
        memcpy(x, y, n * sizeof(double));
    }
}

static void restrict_(const double *rf, int Nf, double *rc, int Nc, double scale) {
#pragma omp parallel for schedule(static)
    for (int J = 0; J < Nc; ++J)
        for (int I = 0; I < Nc; ++I) {
            double s = 0.0;
            for (int bj = -1; bj <= 1; ++bj)
                for (int bi = -1; bi <= 1; ++bi) {
                    const int fi = 2 * I + bi, fj = 2 * J + bj;
                    if (fi < 0 || fj < 0 || fi >= Nf || fj >= Nf) continue;
                    s += (bi ? 0.5 : 1.0) * (bj ? 0.5 : 1.0) * rf[fi + (size_t)fj * Nf];
                }
            rc[I + (size_t)J * Nc] = scale * 0.25 * s;
        }
}

static void prolong_add_(const double *ec, int Nc, double *xf, int Nf) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nf; ++j)
        for (int i = 0; i < Nf; ++i) {
            const int ic = i / 2, jc = j / 2;
            const double *e = ec + ic + (size_t)jc * Nc;
            double v;
            if (i % 2 == 0 && j % 2 == 0) v = e[0];
            else if (i % 2 == 1 && j % 2 == 0) v = 0.5 * (e[0] + e[1]);
            else if (i % 2 == 0) v = 0.5 * (e[0] + e[Nc]);
            else v = 0.25 * (e[0] + e[1] + e[Nc] + e[Nc + 1]);
            xf[i + (size_t)j * Nf] += v;
        }
}

static void vcycle_(mg_t *g, int l, const double *b, double *out, int nu1, int nu2, double omega, double rscale) {
    const op1d *o = &g->op[l];
    const int N = o->N;
    const size_t n = (size_t)N * N;
    if (l == g->nlev - 1) {
        const int nc = g->nc;
        for (int r = 0; r < nc; ++r) {
            double s = 0.0;
            for (int k = 0; k < nc; ++k) s += g->Cinv[(size_t)r * nc + k] * b[k];
            out[r] = s;
        }
        return;
    }
    double *x = g->x[l], *y = g->y[l], *r = g->r[l];
    memset(x, 0, n * sizeof(double));
    jacobi_(o, x, y, b, nu1, omega);
    apply_(o, x, r);
    for (size_t q = 0; q < n; ++q) r[q] = b[q] - r[q];
    const int Nc = g->op[l + 1].N;
    restrict_(r, N, g->b[l + 1], Nc, rscale);
    double *ec = g->r[l + 1] ? g->r[l + 1] : NULL;
    vcycle_(g, l + 1, g->b[l + 1], ec, nu1, nu2, omega, rscale);
    prolong_add_(ec, Nc, x, N);
    jacobi_(o, x, y, b, nu2, omega);
    memcpy(out, x, n * sizeof(double));
}

/* MG-PCG; returns iterations (< 0 on error). */
int orc_mgpcg(int p, int N, const double *b, double *x, double abstol, int maxiter, int nu_pre, int nu_post,
              double omega, double rscale, double *resnorm) {
    mg_t g;
    if (mg_build(p, N, 17, &g)) { mg_free(&g); return -1; }
    const size_t n = (size_t)N * N;
    double *r = malloc(n * sizeof(double)), *z = malloc(n * sizeof(double));
    double *pv = malloc(n * sizeof(double)), *Ap = malloc(n * sizeof(double));
    apply_(&g.op[0], x, Ap);
    for (size_t q = 0; q < n; ++q) r[q] = b[q] - Ap[q];
    double rn = sqrt(dot_(r, r, n));
    int it = 0;
    if (!(rn < abstol)) {
        vcycle_(&g, 0, r, z, nu_pre, nu_post, omega, rscale);
        memcpy(pv, z, n * sizeof(double));
        double rz = dot_(r, z, n);
        for (int s = 0; s < maxiter; ++s) {
            ++it;
            apply_(&g.op[0], pv, Ap);
            const double alpha = rz / dot_(pv, Ap, n);
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; ++q) { x[q] += alpha * pv[q]; r[q] -= alpha * Ap[q]; }
            rn = sqrt(dot_(r, r, n));
            if (rn < abstol) break;
            vcycle_(&g, 0, r, z, nu_pre, nu_post, omega, rscale);
            const double rznew = dot_(r, z, n);
            const double beta = rznew / rz;
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; ++q) pv[q] = z[q] + beta * pv[q];
            rz = rznew;
        }
    }
    if (resnorm) *resnorm = rn;
    free(r); free(z); free(pv); free(Ap);
    mg_free(&g);
    return it;
}

/* ---- H~ norm and the transposed right-hand-side map ---------------------- */
/* sqrt(v' H~ v), H~ = h^2 hw_i hw_j (matrix_free_mg_cg.jl:464). */
double orc_hnorm(int p, int N, const double *v) {
    op1d o;
    if (op1d_build(p, N, &o)) return -1.0;
    const double h = 1.0 / (N - 1);
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int j = 0; j < N; ++j) {
        double sj = 0.0;
        for (int i = 0; i < N; ++i) {
            const double x = v[i + (size_t)j * N];
            sj += o.hw[i] * x * x;
        }
        s += o.hw[j] * sj;
    }
    op1d_free(&o);
    return h * sqrt(s);
}

/* Adjoint of orc_rhs: with b = -H~ (f + sum_X SAT_X_r g_X),
 *   df = -H~ lam,   dg_X = -(H~ SAT_X_r)' lam.
 * Row by row from the entries orc_rhs adds (Dirichlet W/E: the columns
 * j < nbs from each face carry HI_x (tau_D E + beta BS^T E) e; Neumann S/N:
 * the rows i = 0 / N-1 carry HI_y tau_N).  Any output may be NULL. */
int orc_rhs_transpose(int p, int N, const double *lam, double *df, double *dgW, double *dgE, double *dgS,
                      double *dgN) {
    int bm, bn, nbs;
    double bhinv[4], d[5], bd[4][6], bs[4];
    if (tables(p, &bm, &bn, &nbs, bhinv, d, bd, bs)) return -1;
    op1d o;
    if (op1d_build(p, N, &o)) return -2;
    const double h = 1.0 / (N - 1), tauD = -13.0 / h, tauN = -1.0;
    if (df) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i)
                df[i + (size_t)j * N] = -(h * o.hw[i]) * (h * o.hw[j]) * lam[i + (size_t)j * N];
    }
    /* b(i,j) gets -(h hw_i)(h hw_j) (1/(h hw_j)) c_j g_W(i) = -(h hw_i) c_j g_W(i) */
    for (int i = 0; i < N; ++i) {
        double sW = 0.0, sE = 0.0;
        for (int k = 0; k < nbs; ++k) {
            const double c = (k == 0 ? tauD : 0.0) + bs[k] / h;
            sW += c * lam[i + (size_t)k * N];
            sE += c * lam[i + (size_t)(N - 1 - k) * N];
        }
        if (dgW) dgW[i] = -(h * o.hw[i]) * sW;
        if (dgE) dgE[i] = -(h * o.hw[i]) * sE;
    }
    /* b(0,j) gets -(h hw_0)(h hw_j) (1/(h hw_0)) tau_N g_S(j) = -(h hw_j) tau_N g_S(j) */
    for (int j = 0; j < N; ++j) {
        if (dgS) dgS[j] = -(h * o.hw[j]) * tauN * lam[0 + (size_t)j * N];
        if (dgN) dgN[j] = -(h * o.hw[j]) * tauN * lam[(N - 1) + (size_t)j * N];
    }
    op1d_free(&o);
    return 0;
}

/* Gradient of L = 1/2 (u - u_obs)' H~ (u - u_obs) for u = A^-1 b(f, g):
 * lam = A^-1 H~ (u - u_obs) by MG-PCG from zero to ||r||_2 < abstol (a
 * negative abstol: -abstol times ||H~ (u - u_obs)||_2) (2 + 2
 * damped Jacobi, full weighting times 4, 17^2 Cholesky), then the transposed
 * right-hand-side map.  loss / lam / gradients nullable.  Returns the MG-PCG
 * iterations (< 0 on error). */
int orc_adjoint_grad(int p, int N, const double *u, const double *u_obs, double abstol, int maxiter, double *loss,
                     double *lam, double *df, double *dgW, double *dgE, double *dgS, double *dgN) {
    op1d o;
    if (op1d_build(p, N, &o)) return -2;
    const size_t n = (size_t)N * N;
    const double h = 1.0 / (N - 1);
    double *w = malloc(n * sizeof(double));
    double *l = lam ? lam : malloc(n * sizeof(double));
    double L = 0.0;
#pragma omp parallel for reduction(+ : L) schedule(static)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t q = i + (size_t)j * N;
            const double e = u[q] - u_obs[q];
            const double hq = (h * o.hw[i]) * (h * o.hw[j]);
            w[q] = hq * e;
            L += 0.5 * hq * e * e;
        }
    op1d_free(&o);
    if (abstol < 0.0) abstol = -abstol * sqrt(dot_(w, w, n));   /* relative to ||H~ (u - u_obs)|| */
    memset(l, 0, n * sizeof(double));
    double rn = 0.0;
    const int it = orc_mgpcg(p, N, w, l, abstol, maxiter, 2, 2, 2.0 / 3.0, 4.0, &rn);
    if (it >= 0) orc_rhs_transpose(p, N, l, df, dgW, dgE, dgS, dgN);
    if (loss) *loss = L;
    free(w);
    if (!lam) free(l);
    return it;
}

/* ---- seeded Fourier-mode fields ------------------------------------------- */
static const double kPi = 3.14159265358979323846;

/* F(i, j) = sum_m a_m cos(pi (kx_m x_j + ky_m y_i) + ph_m) with x_j = j/(N-1)
 * (slow index) and y_i = i/(N-1) (fast index); modes = nm rows (kx, ky, a, ph).
 * g(k) = sum_m a_m cos(pi k_m t_k + ph_m); bmodes = nb rows (k, a, ph). */
void orc_fourier_field(int N, int nm, const double *modes, double *F) {
    const double h = 1.0 / (N - 1);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        const double x = j == N - 1 ? 1.0 : j * h;
        for (int i = 0; i < N; ++i) {
            const double y = i == N - 1 ? 1.0 : i * h;
            double s = 0.0;
            for (int m = 0; m < nm; ++m) {
                const double *md = modes + 4 * (size_t)m;
                s += md[2] * cos(kPi * (md[0] * x + md[1] * y) + md[3]);
            }
            F[i + (size_t)j * N] = s;
        }
    }
}

void orc_fourier_edge(int N, int nb, const double *bmodes, double *g) {
    const double h = 1.0 / (N - 1);
    for (int k = 0; k < N; ++k) {
        const double t = k == N - 1 ? 1.0 : k * h;
        double s = 0.0;
        for (int m = 0; m < nb; ++m) s += bmodes[3 * m + 1] * cos(kPi * bmodes[3 * m] * t + bmodes[3 * m + 2]);
        g[k] = s;
    }
}
