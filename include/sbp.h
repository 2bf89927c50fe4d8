/*
 * sbp.h -- C ABI of the MI355X (gfx950) matrix-free SBP-SAT 2-D Poisson solver.
 *
 * Grid and layout (the reference's): a level-k problem has N = 2^k + 1 points
 * per side on [0,1]^2 (the apply and CG accept any N >= 2*closure; the
 * multigrid entry points need N = 2^k + 1 >= 17, >= 33 for p >= 6).  Grid functions are dense
 * device arrays of N*N elements with element (i, j) at i + j*N, i fast
 * (Julia column-major (Nx, Ny)).  The fast index i carries the Neumann faces
 * (i = 0 and i = N-1), the slow index j the Dirichlet faces (j = 0, j = N-1).
 *
 * The operator is the reference's matrix
 *     A = -H~ (D2 + SAT_W + SAT_E + SAT_S + SAT_N)      (tau_D = -13/h, tau_N = -1)
 * (preconditioned_CG/level_2_multigrid_new.jl:141-185) for SBP order p = 2, 4,
 * 6, 8 or 10 (diagonal_sbp.jl:191-408).  It is symmetric positive definite
 * for p <= 6; for p = 8, 10 the reference's tau_D is too small for their
 * wider boundary closures and A is symmetric indefinite, so the apply is
 * exact but the CG-type solvers may stop with SBP_ERR_BREAKDOWN.  The
 * two-sweep fused multigrid kernels are compiled for p = 2, 4 and (64- and
 * 128-wide blocks) p = 6, the single-workgroup bottom kernels for p = 2, 4; other
 * orders use the separate passes.
 *
 * Conventions: every pointer argument that names a grid function is a device
 * pointer owned by the caller; work is ordered on the handle's stream (the
 * null stream unless sbp_set_stream was called); scalar outputs (iteration
 * counts, norms) are host pointers and the call synchronises the stream before
 * it returns them.  A handle is not thread-safe; use one handle per stream.
 * Every entry point returns an sbp_status.  No entry point falls back to host
 * computation.
 *
 * Batches: a handle made by sbp_create_batch holds nprob independent problems
 * of the same N and p.  Their grid functions are stacked in one device array,
 * problem q at element offset q*N_l*N_l on level l, and the operator,
 * multigrid-component and batched-solver entry points act on all of them in
 * each launch (per-problem scalars are arrays of nprob).  sbp_diag,
 * sbp_assemble_rhs, sbp_rhs_transpose, sbp_hnorm, sbp_dot and the CSR path act
 * on one problem's arrays; the single-problem solvers need nprob = 1.
 */
#ifndef SBP_H
#define SBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    SBP_OK = 0,
    SBP_ERR_INVALID_SIZE = 1,   /* N too small, or not 2^k+1 where multigrid needs it */
    SBP_ERR_INVALID_ORDER = 2,  /* p not in {2, 4, 6, 8, 10} ({2, 4, 6} for sbp_create_var) */
    SBP_ERR_INVALID_ARG = 3,    /* null pointer, bad dtype/sign/level, bad parameter */
    SBP_ERR_HIP = 4,            /* a HIP runtime call failed */
    SBP_ERR_NONFINITE = 5,      /* NaN/Inf met in a residual or inner product */
    SBP_ERR_BREAKDOWN = 6,      /* p'Ap <= 0 (operator or preconditioner not SPD) */
    SBP_ERR_NOT_CONVERGED = 7,  /* maxiter reached; x holds the last iterate */
    SBP_ERR_NO_DEVICE = 8,      /* no HIP device visible */
    SBP_ERR_VENDOR = 9,         /* a rocSPARSE call failed (sbp_csr_vendor_*) */
    SBP_ERR_SCRATCH = 10        /* sbp_apply_dot_noalloc: the handle's block partials exceed the
                                   library's static scratch array */
} sbp_status;

typedef enum { SBP_F64 = 0, SBP_F32 = 1 } sbp_dtype;
/* Or-ed into sbp_set_tiles' dtype: the shape of the fused two-sweep kernels. */
#define SBP_FUSED 2
/* Or-ed into sbp_set_tiles' dtype: the line-aligned variant of a stencil tile
 * shape (256 x 8, 256 x 16 and 512 x 8; p = 2 and 4): each block stores every grid row
 * in whole 128-byte lines with non-temporal stores. */
#define SBP_ALIGNED 4

typedef struct {
    int nu_pre;              /* pre-smoothing sweeps per level (default 2) */
    int nu_post;             /* post-smoothing sweeps per level (default 2) */
    double omega;            /* damped-Jacobi weight (default 2/3) */
    int coarse_n;            /* coarsest grid side, 2^k+1 (default 17; raised to 33 for p >= 6) */
    double restrict_scale;   /* coarse rhs = restrict_scale * full_weighting(r);
                                4 (default) gives R = P^T, 1 the reference's scale */
    int smoother;            /* 0 = damped Jacobi (default), 1 = Richardson */
    double richardson_omega; /* Richardson weight (reference: 0.15) */
    int poll_every;          /* host convergence poll interval in iterations; 0 (default):
                                after 2 iterations, then where the measured rate of
                                ||r|| predicts convergence (at most 16 ahead) */
    double inner_rtol;       /* mixed precision: inner FP32 solve relative tolerance (default 1e-5) */
    int inner_maxiter;       /* mixed precision: inner iteration cap (default 50) */
    int max_outer;           /* mixed precision: refinement step cap (default 10) */
    int tile_i;              /* apply tile width in i (0 = autotuned) */
    int tile_j;              /* apply rows per thread (0 = autotuned) */
    int use_graph;           /* replay Krylov iterations from a captured hipGraph (default 1) */
    double ir_rtol;          /* mixed precision: also stop when the correction-based error
                                estimate ||d_k|| min(1, ||d_k||/||d_k-1||) <= ir_rtol ||x||
                                (0 = residual test only, the default) */
    int bottom_n;            /* multigrid levels of side <= bottom_n (at most 65) above the
                                coarsest run as two single-workgroup kernels per problem
                                (down and up leg, level data in LDS) around the coarse
                                solve, for at most three such levels (default 0 = off) */
    int fuse_n;              /* with nu_pre = 2 / nu_post = 2, levels whose points over all the
                                handle's problems number at most fuse_n^2 run their pre-smoothing
                                + residual + restriction, and their prolongation + post-smoothing,
                                as one temporally blocked kernel each (default 1025, 0 = off) */
    int drop_coarse_rhs;     /* 1: on every level below the finest, the residual restricted to
                                the next coarser level is -A v, without that level's right-hand
                                side -- the reference's three-level cycle
                                (matrix_free_Three_level_multigrid, initial_guess_cg.jl:281-283);
                                such handles run the separate passes (no fused or bottom
                                kernels).  0 (default): the standard residual rhs - A v. */
} sbp_params;

typedef struct sbp_handle_s* sbp_handle;

/* Library / handle management ------------------------------------------- */
const char* sbp_version(void);
const char* sbp_status_string(int status);
void sbp_default_params(sbp_params* prm);
/* Builds the operator tables for side N and order p and allocates the fine
 * work vectors and, when N = 2^k+1 >= 17, the multigrid hierarchy.
 * Replaces the per-call set-up of Assembling_matrix (level_2_multigrid_new.jl:125)
 * and the per-call CuArray allocations in initial_guess_cg.jl:152-161. */
int sbp_create(int N, int p, const sbp_params* prm, sbp_handle* out);
/* A handle for nprob stacked problems of side N and order p (sbp_create is
 * nprob = 1).  No counterpart in the reference, which solves one problem per
 * call of matrix_free_MGCG (initial_guess_cg.jl:303-346). */
int sbp_create_batch(int N, int p, int nprob, const sbp_params* prm, sbp_handle* out);
/* A handle for nprob problems of one order p and UNEVEN sides (the ragged
 * batch): sides[q] = 2^k + 1, sorted largest first, all on the multigrid
 * hierarchy of sides[0] down to the coarse side (<= 33).  Grid functions are
 * ragged arrays: problem q at element offset sum_{q' < q} sides[q']^2.  The
 * multigrid levels are shared across sizes: level l holds every problem whose
 * side is at least the level's (a 513^2 problem's finest level and a 1025^2
 * problem's first coarse level sit in the same stack), so each V-cycle pass
 * is ONE launch over all the problems present on its level, and the Krylov
 * passes are one launch per level over the problems whose finest level it
 * is.  Serves sbp_mgpcg_batch and sbp_mgpcg_mixed_batch (per-problem
 * tolerances and masks, problems in this order), sbp_set_tiles, sbp_get_info,
 * sbp_get_batch, sbp_set_stream and sbp_destroy; every other entry point
 * returns SBP_ERR_INVALID_ARG.  No counterpart in the reference (one problem
 * per matrix_free_MGCG call, initial_guess_cg.jl:303-346). */
int sbp_create_ragged(int p, int nprob, const int* sides, const sbp_params* prm, sbp_handle* out);
/* Variable-coefficient operator d/dx (bx d/dx) + d/dy (by d/dy), p = 2, 4 or 6,
 * from variable_diagonal_sbp_D2 (diagonal_sbp.jl:470-740) on both axes as
 * get_operators uses it (legacy_functions.jl:100-174): bx (host, N values)
 * along the slow, Dirichlet index j, by along the fast, Neumann index i, both
 * > 0.  The Dirichlet penalty is -13 b_face / h (the reference's -13/h at
 * b = 1) and the Neumann data are the fluxes b du/dn.  Such a handle serves
 * sbp_apply (FP64, level 0), sbp_apply_dot, sbp_cg, sbp_assemble_rhs,
 * sbp_rhs_transpose and sbp_hnorm, and for N = 2^k + 1 also sbp_vcycle
 * (FP64) and sbp_mgpcg: FP64 MG-PCG on the hierarchy rediscretised with the
 * profiles injected on every level (b_l[i] = b[2^l i]; damped Jacobi with
 * nu_pre / nu_post / omega, full weighting times restrict_scale, bilinear
 * prolongation, dense coarsest solve at coarse_n <= 33, raised to 33 for
 * p = 6).  This hierarchy has no counterpart in the reference, which builds
 * no multigrid for this operator.  The mixed-precision, batched, smoother,
 * diag and CSR entry points return SBP_ERR_INVALID_ARG /
 * SBP_ERR_INVALID_SIZE.  N >= 8, 16, 24 for p = 2, 4, 6. */
int sbp_create_var(int N, int p, const double* bx, const double* by, const sbp_params* prm,
                   sbp_handle* out);
int sbp_get_batch(sbp_handle h, int* nprob);
int sbp_destroy(sbp_handle h);
int sbp_set_stream(sbp_handle h, void* hip_stream);
int sbp_get_info(sbp_handle h, int* N, int* p, int* nlevels);
/* Stencil tile shape (tile_i threads along i x tile_j rows per thread) used
 * on `level` for element type `dtype`; shapes not compiled in are rejected.
 * With SBP_ALIGNED or-ed into dtype: the shape's line-aligned variant (each
 * block computes tile_i columns and stores tile_i - 128 / element-size of
 * them per row, starting on a 128-byte line).
 * With SBP_FUSED or-ed into dtype: the block shape (tile_i output columns x
 * tile_j fine rows) of the level's fused two-sweep kernels (64 x 4, 64 x 8,
 * 128 x 8, 256 x 8; for p = 6 all but 256 x 8, whose FP64 tiles exceed the
 * 160 KiB of workgroup LDS). */
int sbp_set_tiles(sbp_handle h, int level, int dtype, int tile_i, int tile_j);

/* Operator --------------------------------------------------------------- */
/* Au = sign * A u on multigrid level `level` (0 = finest).  sign = -1 gives
 * the reference's matrix_free_A_full_GPU(idata, odata) (split_matrix_free.jl:444-486),
 * which returns -A u; sign = +1 gives the SPD A u.  dtype selects the element
 * type of both u and Au. */
int sbp_apply(sbp_handle h, int level, const void* u, void* Au, int sign, int dtype);
/* Au = A u and dot[q] = u_q'A u_q in one pass (FP64; dot: nprob host doubles). */
int sbp_apply_dot(sbp_handle h, const double* u, double* Au, double* dot);
/* sbp_apply_dot without device memory of its own: Au = A u and dot[q] =
 * u_q'A u_q (FP64), with the block partials and the per-problem results in
 * the library's file-scope static device array (1 MiB per GPU, held until the
 * process ends) and the results written to `dot`, a DEVICE array of nprob
 * doubles; nothing is allocated and nothing waits for the stream, so the call
 * may be captured into a hipGraph.  The handle (made beforehand) supplies the
 * operator tables and the stream; its device must be current.  Handles whose
 * finest pass has more than 2^17 - nprob block partials over all problems
 * return SBP_ERR_SCRATCH (4097^2 with the default 256 x 8 tiles uses 8721 per
 * problem).  Every call on one GPU shares the array: calls that overlap on one
 * GPU (two streams, two host threads) race.  Variable-coefficient handles:
 * SBP_ERR_INVALID_ARG.  No counterpart in the reference; for callers that
 * cannot allocate (code under stream capture, tight serving loops). */
int sbp_apply_dot_noalloc(sbp_handle h, const double* u, double* Au, double* dot);
/* d = diag(A) on the finest level (FP64). */
int sbp_diag(sbp_handle h, double* d);
/* b = -H~ (f + SAT_W_r gW + SAT_E_r gE + SAT_S_r gS + SAT_N_r gN)
 * (level_2_multigrid_new.jl:164-167,181,185).  f: N*N; gW, gE: N values
 * indexed by i; gS, gN: N values indexed by j.  Any g may be NULL (zero). */
int sbp_assemble_rhs(sbp_handle h, const double* f, const double* gW, const double* gE,
                     const double* gS, const double* gN, double* b);
/* Transpose of sbp_assemble_rhs: given lambda, df = -H~ lambda and
 * dg_X = -(H~ SAT_X_r)^T lambda.  Any output may be NULL. */
int sbp_rhs_transpose(sbp_handle h, const double* lam, double* df, double* dgW,
                      double* dgE, double* dgS, double* dgN);
/* *out = sqrt(v' H~ v) (matrix_free_mg_cg.jl:464). */
int sbp_hnorm(sbp_handle h, const double* v, double* out);
/* The 1-D norm weights of the finest grid, hw[k] = H_kk / h for k < N (host
 * array): the diagonal of H from diagonal_sbp_D1/D2 (diagonal_sbp.jl:421-428),
 * so that the reference's H~ (level_2_multigrid_new.jl:120, returned by
 * Assembling_matrix) is h^2 hw[i] hw[j] at (i, j).  Host-only, no device work. */
int sbp_h_weights(sbp_handle h, double* hw);
/* *out = u'v (FP64). */
int sbp_dot(sbp_handle h, const double* u, const double* v, double* out);

/* Multigrid components (level l has side N_l = (N_{l-1}+1)/2) ------------ */
/* rc = restrict_scale * FW(rf): full weighting [1 2 1;2 4 2;1 2 1]/16 with zero
 * padding (restriction_2D_kernel, initial_guess_cg.jl:98-131), fine level l -> l+1. */
int sbp_restrict(sbp_handle h, int level, const void* rf, void* rc, int dtype);
/* xf += P ec, bilinear (prolongation_2D_kernel, initial_guess_cg.jl:41-77), l+1 -> l. */
int sbp_prolong_add(sbp_handle h, int level, const void* ec, void* xf, int dtype);
/* nsweeps damped-Jacobi sweeps x <- x + omega D^-1 (b - A x) on level l
 * (jacobi_brittany!, level_2_multigrid_new.jl:335-344); smoother = 1 uses
 * Richardson x <- x + w (b - A x) (matrix_free_richardson, initial_guess_cg.jl:134-144). */
int sbp_smooth(sbp_handle h, int level, const void* b, void* x, int nsweeps, int dtype);
/* Exact coarsest-level solve x = A_c^-1 b in FP64 (replaces the UMFPACK
 * A_2h \ f of initial_guess_cg.jl:170). */
int sbp_coarse_solve(sbp_handle h, const double* b, double* x);
/* One V-cycle z ~ A^-1 r from a zero initial guess (matrix_free_Two_level_multigrid,
 * initial_guess_cg.jl:147-180, generalised to all levels). */
int sbp_vcycle(sbp_handle h, const void* r, void* z, int dtype);

/* Solvers (x holds the initial guess on entry and the solution on exit) --- */
/* Matrix-free CG on the SPD A, stop when ||r||_2 < abstol
 * (CG_Matrix_Free_GPU_v2 / CG_full_GPU, split_matrix_free.jl:647-714, with the
 * sign fixed).  hist (nullable, maxiter+1 doubles) receives ||r_k||_2. */
int sbp_cg(sbp_handle h, const double* b, double* x, double abstol, int maxiter,
           int* iters, double* resnorm, double* hist);
/* FP64 matrix-free MG-preconditioned CG (matrix_free_MGCG, initial_guess_cg.jl:303-346). */
int sbp_mgpcg(sbp_handle h, const double* b, double* x, double abstol, int maxiter,
              int* iters, double* resnorm, double* hist);
/* FP32 inner MG-PCG inside FP64 iterative refinement: the outer residual
 * r = b - A x, the update x += d, the norms and the coarse solve are FP64;
 * the inner preconditioned solve of A d = r is FP32 (to inner_rtol ||r||).
 * Stops when the FP64 ||r||_2 < abstol, or (ir_rtol > 0) when the error
 * estimate from successive corrections drops below ir_rtol ||x||.
 * outer_iters receives the number of corrections applied, inner_iters
 * (nullable) the total inner iterations, resnorm the last FP64 residual norm
 * evaluated.  On a variable-coefficient handle it is instead FP64 MG-PCG
 * whose V-cycle runs in FP32 (r cast down, z cast up; the operator products,
 * the Krylov vectors and the inner products stay FP64), for at most
 * inner_maxiter * max(max_outer, 1) iterations; outer_iters and inner_iters
 * both receive the iteration count. */
int sbp_mgpcg_mixed(sbp_handle h, const double* b, double* x, double abstol, int max_outer,
                    int* outer_iters, int* inner_iters, double* resnorm);
/* Batched forms of sbp_mgpcg and sbp_mgpcg_mixed over the handle's nprob
 * problems: one launch sequence drives all of them, each with its own
 * tolerance abstol[q] and convergence mask (a converged problem's work is
 * skipped on the device while the others iterate).  Host outputs are arrays of
 * nprob (each nullable); seconds[q] is the device wall-clock time from the
 * solve's start to the moment problem q stopped iterating (-1 if it never
 * did), status[q] its own sbp_status.  Returns the worst status over the
 * problems.  No counterpart in the
 * reference (one problem per matrix_free_MGCG call, initial_guess_cg.jl:303-346). */
int sbp_mgpcg_batch(sbp_handle h, const double* b, double* x, const double* abstol, int maxiter,
                    int* iters, double* resnorm, double* seconds, int* status);
int sbp_mgpcg_mixed_batch(sbp_handle h, const double* b, double* x, const double* abstol,
                          int max_outer, int* outer_iters, int* inner_iters, double* resnorm,
                          double* seconds, int* status);

/* Gradient of L = 1/2 (u - u_obs)' H~ (u - u_obs) with u = A^-1 b(f, g)
 * (u: the forward solution): the adjoint solve A^T lambda = H~ (u - u_obs)
 * (FP64 MG-PCG; A is symmetric) to ||r|| < abstol, then dL/df, dL/dg_X from
 * the transpose of sbp_assemble_rhs.  lam (nullable) receives lambda; loss
 * (nullable, host) receives L; any gradient output may be NULL.  No
 * counterpart in the reference (its solves are not differentiated). */
int sbp_adjoint_grad(sbp_handle h, const double* u, const double* u_obs, double abstol, int maxiter,
                     double* loss, double* lam, double* df, double* dgW, double* dgE, double* dgS,
                     double* dgN, int* iters);

/* Initial-guess interpolation CG (initial_guess_interpolation_CG /
 * _CG_GPU / _CG_Matrix_Free_GPU, initial_guess_cg.jl:394-430; with nlev = 3,
 * initial_guess_interpolation_three_level_CG_GPU, :447-456).  nlev = 2: solve
 * A_2h x_2h = b2h exactly on the grid of side (N+1)/2 (the reference: sparse
 * LU; here MG-PCG to ||r|| < 1e-13 ||b2h||), set x = P x_2h (bilinear) and run
 * matrix-free CG on A x = b from there to ||r||_2 < abstol.  nlev = 3: the
 * exact solve is on the 4h grid (b4h), CG on the 2h grid from P x_4h to
 * abstol, then CG on this grid from P x_2h.  b2h / b4h are the right-hand
 * sides assembled on the coarser grids (the reference's Assembling_matrix at
 * level - 1 / - 2).  iters and resnorm (nullable, nlev - 1 entries, coarse to
 * fine) receive the CG iterations and final ||r|| per CG grid; hist (nullable,
 * maxiter + 1 doubles) the fine grid's ||r_k||.  x is output only.  Needs
 * N = 2^k+1 with (N+1)/2 (and for nlev = 3 also (N+3)/4) >= coarse_n; the
 * reference's reversal of the interpolated guess (:416) maps between its
 * kernel's array orientation and its assembled ordering, which coincide in
 * this layout. */
int sbp_cg_interp(sbp_handle h, int nlev, const double* b, const double* b2h, const double* b4h,
                  double* x, double abstol, int maxiter, int* iters, double* resnorm, double* hist);

/* CSR comparison path (the assembled operator, as CG_GPU_sparse,
 * split_matrix_free.jl:744-769) ------------------------------------------ */
/* Fills device CSR arrays (int32 indices) of the assembled A; call first with
 * row_ptr = col = val = NULL to get *nnz. */
int sbp_csr_assemble(sbp_handle h, int32_t* row_ptr, int32_t* col, double* val, int64_t* nnz);
/* y = A x with a hand-written CSR SpMV kernel. */
int sbp_csr_spmv(sbp_handle h, const int32_t* row_ptr, const int32_t* col, const double* val,
                 const double* x, double* y);
/* CG with the CSR SpMV in place of the matrix-free apply. */
int sbp_csr_cg(sbp_handle h, const int32_t* row_ptr, const int32_t* col, const double* val,
               const double* b, double* x, double abstol, int maxiter, int* iters,
               double* resnorm);

/* The same CSR matrix through rocSPARSE (the vendor SpMV: the reference's
 * comparison path multiplies with CUSPARSE, Poisson_CG.jl:217-271,
 * kernel_comparisons/SPMV_CSR.jl:197-202, CG_GPU_sparse split_matrix_free.jl:744-769).
 * attach: keeps the caller's CSR arrays by reference (they must stay alive
 * until sbp_destroy or the next attach) and runs rocSPARSE's analysis once
 * (alg 0: csr_adaptive, 1: csr_rowsplit); single-problem handles. */
int sbp_csr_vendor_attach(sbp_handle h, const int32_t* row_ptr, const int32_t* col, const double* val, int64_t nnz,
                          int alg);
/* y = A x (rocsparse_spmv on the handle's stream). */
int sbp_csr_vendor_spmv(sbp_handle h, const double* x, double* y);
/* CG as sbp_csr_cg with the rocSPARSE SpMV. */
int sbp_csr_vendor_cg(sbp_handle h, const double* b, double* x, double abstol, int maxiter, int* iters,
                      double* resnorm);

/* Streaming-copy calibration kernel (copy_naive!, split_matrix_free.jl:51-77). */
int sbp_copy(sbp_handle h, const void* src, void* dst, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* SBP_H */
